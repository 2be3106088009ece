"""Named fixed-source workloads, built with the product's own quadrature and
seeded-media generator (no oracle involved).

* ``slab_five_region`` -- 1D, 8 cm slab, 400 uniform cells, five regions
  (sigma_t 50/5/0/1/1, sigma_s 0/0/0/0.9/0.9, Q 50/0/0/1/0), S16 Gauss-Legendre,
  diamond difference, one group.
* ``random_blocks`` -- 3D cube of 1 cm cells with seeded 8^3-cell material
  blocks: sigma_t ~ U[0.5, 2] per (block, group), c ~ U[0.3, 0.9], 0.7 of the
  scattering in-group and 0.3 to the next group (the last group keeps only its
  in-group share), unit source in group 0 on the central 8^3 cells; GC product
  quadrature np x na per octant; theta-weighted DD.  64^3, 3 groups, 4x4 is the
  standard instance.
* ``multigroup_full`` -- 128^3, 8 groups, 8x4 per octant (256 directions),
  log-uniform sigma_t in [0.2, 5], c ~ U[0.5, 0.95], full scattering matrix with
  down-scatter and up-scatter between the last two groups, Q ~ U[0, 1] in 5% of
  blocks.
* ``high_scatter`` -- the same geometry with c ~ U[0.95, 0.995], sigma_t
  log-uniform in [0.5, 10] and 98 % of the scattering in-group (down-scatter
  only), diamond difference: the learned-correction workload.
* ``batched_uneven`` -- independent seeded problems of uneven size for the
  batched path.
* ``kinetic_blocks`` -- ``fissile_blocks`` (subcritical) with transient data:
  inverse velocities per group, two delayed groups on the fissile materials,
  step sources of every type on the others.
* ``fissile_blocks`` -- ``random_blocks`` with fission data on a seeded share of
  the blocks: chi_p concentrated in the fastest groups, nu*sigma_f a seeded
  multiple of each group's absorption; below critical for fixed-source solves
  (``nu_ratio`` < 1), near critical for power iteration.
"""
from __future__ import annotations

import numpy as np

from . import capi
from .capi import MediaSpec, Problem

SEED_BASE = 20261202


def quadrature_3d(np_: int, na: int):
    """GCAngularQuadrature(2na, 2np, 3) as the reference's 3D solver builds it."""
    return capi.gc_quadrature(2 * na, 2 * np_, 3)


def quadrature_2d(np_: int, na: int):
    return capi.gc_quadrature(2 * na, 2 * np_, 2)


def quadrature_1d(np_: int):
    mu, w = capi.gl_quadrature(2 * np_)
    return mu, np.zeros_like(mu), np.zeros_like(mu), w


def slab_five_region(np_: int = 8, scale: int = 1) -> Problem:
    """400*scale cells of 0.02/scale cm over the regions [0,2),[2,3),[3,5),[5,6),[6,8]."""
    widths = [2.0, 1.0, 2.0, 1.0, 2.0]
    cells = [100 * scale, 50 * scale, 100 * scale, 50 * scale, 100 * scale]
    lx = np.concatenate([np.full(n, w / n) for w, n in zip(widths, cells)])
    mat = np.concatenate([np.full(n, m, dtype=np.int32) for m, n in enumerate(cells)])
    sigt = np.array([[50.0], [5.0], [0.0], [1.0], [1.0]])
    sigs = np.array([0.0, 0.0, 0.0, 0.9, 0.9]).reshape(5, 1, 1)
    src = np.array([[50.0], [0.0], [0.0], [1.0], [0.0]])
    has = np.ones(5, dtype=np.uint8)
    mu, eta, xi, w = quadrature_1d(np_)
    return Problem(1, lx, np.ones(1), np.ones(1), 1, capi.SCHEME_DD, mat, sigt, sigs, src, has,
                   mu, eta, xi, w, name=f"slab_five_region_{len(lx)}_S{2 * np_}")


def kobayashi1(refinement: int = 0, np_: int = 10, na: int = 10,
               scheme: int = capi.SCHEME_DD) -> Problem:
    """The reference's cases/3D_kobayashi_1.xml: a 200 cm cube of interval widths
    50/40/20/40/50 cm with 5/4/2/4/5 cells each, times (refinement + 1)
    (InputParameters.cpp ix * (refinement + 1)); block 0 (the central 20 cm cube)
    is the unit source with sigma_t 0.1, sigma_s 0.05; block 1 (the surrounding
    60 cm shell) is a near-void duct region (1e-4 / 5e-5); block 2 is the shield
    (0.1 / 0.05); vacuum boundaries, one group, GC np x na per octant."""
    widths = [50.0, 40.0, 20.0, 40.0, 50.0]
    ivals = [k * (refinement + 1) for k in (5, 4, 2, 4, 5)]
    l = np.concatenate([np.full(n, w / n) for w, n in zip(widths, ivals)])
    blocks = np.full((5, 5, 5), 2, dtype=np.int32)
    blocks[1:4, 1:4, 1:4] = 1
    blocks[2, 2, 2] = 0
    idx = np.repeat(np.arange(5), ivals)
    mat = blocks[idx][:, idx][:, :, idx].astype(np.int32)  # [k][j][i]
    sigt = np.array([[1e-1], [1e-4], [1e-1]])
    sigs = np.array([5e-2, 5e-5, 5e-2]).reshape(3, 1, 1)
    src = np.array([[1.0], [0.0], [0.0]])
    has = np.ones(3, dtype=np.uint8)
    mu, eta, xi, w = quadrature_3d(np_, na)
    return Problem(3, l, l.copy(), l.copy(), 1, scheme, mat.reshape(-1), sigt, sigs, src, has,
                   mu, eta, xi, w, name=f"kobayashi1_{len(l)}c_{8 * np_ * na}d")


def cell_at_point(widths, x):
    """Index of the first cell whose closed interval [x_c - l/2, x_c + l/2] holds x,
    with the reference's centroids x_c = l_0/2 + sum_(j<i) l_j accumulated in
    order (BrickMesh3D.cpp:83-101, CartesianCell3D.cpp:60-67, fluxAtPoint
    BrickMesh3D.cpp:202-213 returns the first such cell in cell order)."""
    xc = 0.5 * widths[0]
    for i, li in enumerate(widths):
        if xc - 0.5 * li <= x <= xc + 0.5 * li:
            return i
        xc += li
    raise ValueError(f"point {x} lies outside the mesh")


def _media(spec: MediaSpec):
    return capi.media_generate(spec)


def block_media_problem(n: int, groups: int, np_: int, na: int, scheme: int, spec: MediaSpec,
                        name: str) -> Problem:
    mat, sigt, sigs, src, has = _media(spec)
    mu, eta, xi, w = quadrature_3d(np_, na)
    l = np.full(n, spec.cell_width)
    return Problem(3, l, l.copy(), l.copy(), groups, scheme, mat, sigt, sigs, src, has, mu, eta,
                   xi, w, name=name, meta={"seed": int(spec.seed)})


def random_blocks(n: int = 64, groups: int = 3, np_: int = 4, na: int = 4,
                  scheme: int = capi.SCHEME_TW_DD, seed: int = SEED_BASE + 1,
                  block: int = 8) -> Problem:
    half = n // 2
    spec = MediaSpec(n=n, block=block, groups=groups, sigt_log=0, sigt_lo=0.5, sigt_hi=2.0,
                     c_lo=0.3, c_hi=0.9, scatter_rule=capi.SCATTER_DOWN1,
                     source_rule=capi.SOURCE_CENTRAL, src_lo=half - 4, src_hi=half + 4,
                     source_fraction=0.0, cell_width=1.0, seed=seed)
    return block_media_problem(n, groups, np_, na, scheme, spec,
                               f"random_blocks_{n}c_{groups}g_{8 * np_ * na}d")


def multigroup_full(n: int = 128, groups: int = 8, np_: int = 8, na: int = 4,
                    scheme: int = capi.SCHEME_TW_DD, seed: int = SEED_BASE + 2,
                    c_lo: float = 0.5, c_hi: float = 0.95, sigt_lo: float = 0.2,
                    sigt_hi: float = 5.0, block: int = 8) -> Problem:
    spec = MediaSpec(n=n, block=block, groups=groups, sigt_log=1, sigt_lo=sigt_lo,
                     sigt_hi=sigt_hi, c_lo=c_lo, c_hi=c_hi, scatter_rule=capi.SCATTER_FULL,
                     source_rule=capi.SOURCE_BLOCKS, src_lo=0, src_hi=0, source_fraction=0.05,
                     cell_width=1.0, seed=seed)
    return block_media_problem(n, groups, np_, na, scheme, spec,
                               f"multigroup_full_{n}c_{groups}g_{8 * np_ * na}d")


def high_scatter(n: int = 128, groups: int = 8, np_: int = 8, na: int = 4,
                 seed: int = SEED_BASE + 4, scheme: int = capi.SCHEME_DD,
                 scatter_rule: int = capi.SCATTER_INGROUP) -> Problem:
    """128^3 cells of 1 cm, 16^3 blocks of 8^3 cells, sigma_t log-uniform in
    [0.5, 10] and c ~ U[0.95, 0.995] per (block, group), Q ~ U[0, 1] in 5 % of
    blocks.  The default scatter rule keeps 98 % of c in-group (down-scatter
    only), so the in-group source iterations -- spectral radius up to ~0.99 --
    dominate the solve; SCATTER_FULL gives the full-matrix split instead."""
    spec = MediaSpec(n=n, block=8, groups=groups, sigt_log=1, sigt_lo=0.5, sigt_hi=10.0,
                     c_lo=0.95, c_hi=0.995, scatter_rule=scatter_rule,
                     source_rule=capi.SOURCE_BLOCKS, src_lo=0, src_hi=0, source_fraction=0.05,
                     cell_width=1.0, seed=seed)
    tag = "" if scatter_rule == capi.SCATTER_INGROUP else "_full"
    return block_media_problem(n, groups, np_, na, scheme, spec,
                               f"high_scatter{tag}_{n}c_{groups}g_{8 * np_ * na}d")


def diffusive(n: int = 128, np_: int = 8, na: int = 4, seed: int = SEED_BASE + 5,
              c_lo: float = 0.95, c_hi: float = 0.995) -> Problem:
    """One group, in-group scattering ratio c ~ U[c_lo, c_hi] per block, log-uniform
    sigma_t in [0.5, 10], sources in 5% of blocks: diffusive media where plain
    source iteration converges slowly (spectral radius near c)."""
    spec = MediaSpec(n=n, block=8, groups=1, sigt_log=1, sigt_lo=0.5, sigt_hi=10.0, c_lo=c_lo,
                     c_hi=c_hi, scatter_rule=capi.SCATTER_DOWN1, source_rule=capi.SOURCE_BLOCKS,
                     src_lo=0, src_hi=0, source_fraction=0.05, cell_width=1.0, seed=seed)
    return block_media_problem(n, 1, np_, na, capi.SCHEME_TW_DD, spec,
                               f"diffusive_{n}c_1g_{8 * np_ * na}d")


def batched_uneven(count: int = 512, seed: int = SEED_BASE + 3, np_: int = 4, na: int = 4,
                   sizes=(16, 24, 32, 48, 64), groups=(1, 2, 4)):
    """Independent problems: cube edge and group count drawn per problem."""
    rng = np.random.default_rng(seed)
    out = []
    for b in range(count):
        n = int(rng.choice(sizes))
        g = int(rng.choice(groups))
        spec = MediaSpec(n=n, block=8, groups=g, sigt_log=0, sigt_lo=0.5, sigt_hi=2.0, c_lo=0.9,
                         c_hi=0.99, scatter_rule=capi.SCATTER_DOWN1,
                         source_rule=capi.SOURCE_BLOCKS, src_lo=0, src_hi=0,
                         source_fraction=0.05, cell_width=1.0, seed=seed * 1000 + b)
        out.append(block_media_problem(n, g, np_, na, capi.SCHEME_TW_DD, spec,
                                       f"batched_{b}_{n}c_{g}g"))
    return out


def fission_for(p: Problem, seed: int, fissile_fraction: float = 0.5,
                nu_ratio=(0.4, 0.8)) -> capi.Fission:
    """Fission data for the materials of `p`: a seeded `fissile_fraction` of them
    (always including those of the first and the middle cell) get chi_p (0.75 / 0.25 over the two fastest groups; all in group 0 with one
    group) and nu*sigma_f = r * absorption with r ~ U[nu_ratio]; the others have
    neither.  No delayed groups (prompt fraction 1)."""
    rng = np.random.default_rng(seed)
    M, G = p.num_mats, p.groups
    sigt = np.asarray(p.sigt, dtype=np.float64).reshape(M, G)
    sigs = np.asarray(p.sigs, dtype=np.float64).reshape(M, G, G)
    absorb = np.maximum(sigt - sigs.sum(axis=1), 0.05 * sigt)  # out of each source group
    fis = rng.random(M) < fissile_fraction
    mat = np.asarray(p.mat).reshape(-1)
    fis[mat[0]] = fis[mat[mat.size // 2]] = True  # corner and centre cells are fissile
    chi = np.zeros((M, G))
    chi[:, 0] = 1.0 if G == 1 else 0.75
    if G > 1:
        chi[:, 1] = 0.25
    chi[~fis] = 0.0
    prod = rng.uniform(nu_ratio[0], nu_ratio[1], (M, G)) * absorb
    prod[~fis] = 0.0
    return capi.Fission(chi, prod, np.ones((M, G)), fis.astype(np.uint8), fis.astype(np.uint8))


def fissile_blocks(n: int = 32, groups: int = 3, np_: int = 2, na: int = 2,
                   scheme: int = capi.SCHEME_TW_DD, seed: int = SEED_BASE + 6,
                   nu_ratio=(0.4, 0.8), dims: int = 3) -> Problem:
    """dims 3: an n^3 ``random_blocks`` cube; dims 2: its k = 0 plane as an n x n
    square with the 2D product quadrature; dims 1: its first row as an n-cell slab
    with Gauss-Legendre S_(2 np_) (diamond difference)."""
    block = max(b for b in range(1, 9) if n % b == 0)
    p = random_blocks(n, groups, np_, na, scheme, seed, block=block)
    if dims < 3:
        mat = np.asarray(p.mat).reshape(n, n, n)[0]
        if dims == 2:
            mu, eta, xi, w = quadrature_2d(np_, na)
            p = Problem(2, p.lx, p.ly, np.ones(1), groups, scheme, mat.reshape(-1), p.sigt,
                        p.sigs, p.src, p.has_src, mu, eta, xi, w, name=p.name, meta=p.meta)
        else:
            mu, eta, xi, w = quadrature_1d(np_)
            p = Problem(1, p.lx, np.ones(1), np.ones(1), groups, capi.SCHEME_DD,
                        mat[n // 2].copy(), p.sigt, p.sigs, p.src, p.has_src, mu, eta, xi, w,
                        name=p.name, meta=p.meta)
        # the central source lies off the slice: put a unit source in every material
        p.src = np.zeros((p.num_mats, groups))
        p.src[:, 0] = 1.0
        p.has_src = np.ones(p.num_mats, dtype=np.uint8)
    p.fission = fission_for(p, seed, nu_ratio=nu_ratio)
    p.name = f"fissile_blocks_{dims}d_{n}c_{groups}g"
    return p


def kinetic_blocks(n: int = 16, groups: int = 2, np_: int = 2, na: int = 2,
                   scheme: int = capi.SCHEME_TW_DD, seed: int = SEED_BASE + 7, dims: int = 3,
                   dt: float = 0.01) -> Problem:
    """Transient workload: inv_v_g = 10^(-9 + 2g) s/cm (fast to thermal), two
    delayed groups (beta 0.002/0.004 per source group, lambda 0.08/1.2 /s, chi_d
    in the second-fastest group) on fissile materials, prompt fraction 1 - sum
    beta; non-fissile materials carry step sources cycling through both/insert/
    remove with times on the step grid of width dt."""
    p = fissile_blocks(n, groups, np_, na, scheme, seed, nu_ratio=(0.3, 0.6), dims=dims)
    M, G, K = p.num_mats, groups, capi.MAX_DELAYED
    f = p.fission
    inv_v = np.tile(10.0 ** (-9.0 + 2.0 * np.arange(G)), (M, 1))
    nd = np.where(f.has_chi > 0, 2, 0).astype(np.int32)
    chi_d, beta, lam = np.zeros((M, G, K)), np.zeros((M, G, K)), np.zeros((M, K))
    for m in range(M):
        if nd[m]:
            chi_d[m, min(1, G - 1), :2] = 1.0
            beta[m, :, 0], beta[m, :, 1] = 0.002, 0.004
            lam[m, :2] = (0.08, 1.2)
    prompt = np.ones((M, G))
    prompt[nd > 0] = 1.0 - (0.002 + 0.004)
    p.fission = capi.Fission(f.chi, f.prod, prompt, f.has_chi, f.has_prod)
    p.kinetics = capi.Kinetics(inv_v, nd, chi_d, beta, lam)
    ty = np.full(M, capi.STEP_NONE, dtype=np.int32)
    t_in, t_out, s_src = np.zeros(M), np.zeros(M), np.zeros((M, G))
    others = [m for m in range(M) if nd[m] == 0]
    for j, m in enumerate(others):
        ty[m] = (capi.STEP_BOTH, capi.STEP_INSERT, capi.STEP_REMOVE)[j % 3]
        t_in[m], t_out[m] = dt, 2 * dt
        s_src[m, 0] = 10.0
    p.steps = capi.StepSources(ty, t_in, t_out, s_src)
    p.name = f"kinetic_blocks_{dims}d_{n}c_{G}g"
    return p
