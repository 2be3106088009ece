"""Pins the oracle's fission source and power iteration with what the
mathematics fixes: on tiny meshes the discrete transport operator is assembled
column by column and

  * a fixed-source solve with fission must give phi = (I - D(S + F))^-1 D Q,
  * the power iteration must converge to the dominant eigenvalue of
    (I - D S)^-1 D F and its eigenvector,

where D maps a per-cell angular source to the swept scalar flux of a group,
S is the (normalised) scattering matrix and F = norm * chi (x) nu*sigma_f.  In 1D,
D is built here from the textbook diamond-difference recurrence (independent of
the oracle); in 2D/3D from the oracle's sweep, which the reference's own compiled
solver pins (tests/golden/).  A dropped term (in-group fission, the 1/k), a
transposed chi/nu*sigma_f or a wrong normalisation fails these.  CPU only.

Reference behaviour pinned: src/TransportSolver3D.cpp:29-86, 339-456,
src/TransportSolver2D.cpp:29-86, 339-456, src/TransportSolver1D.cpp:36-90, 369-493,
include/DiamondDifference1D.h:14-46.
"""
from __future__ import annotations

import numpy as np
import pytest

from doma_amd import capi, workloads
from oracle import binding as ob


def _dd1d_sweep(p, g, q):
    """Textbook DD in a slab with vacuum boundaries: psi_c = (q + 2|mu|/h psi_in) /
    (sigma + 2|mu|/h), psi_out = 2 psi_c - psi_in; phi = sum_n w_n psi_c."""
    lx = np.asarray(p.lx, dtype=np.float64)
    mat = np.asarray(p.mat).reshape(-1)
    sig = np.asarray(p.sigt).reshape(p.num_mats, p.groups)[mat, g]
    phi = np.zeros(len(lx))
    for mu, w in zip(np.asarray(p.mu).reshape(-1), np.asarray(p.w).reshape(-1)):
        a = abs(mu)
        order = range(len(lx)) if mu > 0 else range(len(lx) - 1, -1, -1)
        psi_in = 0.0
        for i in order:
            c = (q[i] + 2.0 * a / lx[i] * psi_in) / (sig[i] + 2.0 * a / lx[i])
            phi[i] += w * c
            psi_in = 2.0 * c - psi_in
    return phi


def _operators(p, sweep, norm):
    """D (block diagonal over groups), S and F as dense [G*N, G*N] matrices."""
    N, G = p.N, p.groups
    mat = np.asarray(p.mat).reshape(-1)
    D = np.zeros((G * N, G * N))
    for g in range(G):
        for j in range(N):
            e = np.zeros(N)
            e[j] = 1.0
            D[g * N:(g + 1) * N, g * N + j] = sweep(p, g, e)
    sigs = np.asarray(p.sigs).reshape(p.num_mats, G, G)
    f = p.fission
    S = np.zeros((G * N, G * N))
    F = np.zeros((G * N, G * N))
    for g in range(G):
        for gp in range(G):
            S[g * N + np.arange(N), gp * N + np.arange(N)] = norm * sigs[mat, g, gp]
            F[g * N + np.arange(N), gp * N + np.arange(N)] = (
                norm * f.chi[mat, g] * f.prod[mat, gp] * f.has_chi[mat])
    return D, S, F


def _oracle_sweep(p, g, q):
    return ob.sweep(p, g, q)


def _setup(dims, n, groups, nu_ratio, seed=11):
    p = workloads.fissile_blocks(n, groups, 2 if dims > 1 else 4, 1, capi.SCHEME_DD,
                                 seed=seed, nu_ratio=nu_ratio, dims=dims).normalized()
    if dims == 1:
        return p, _dd1d_sweep, 0.5
    return p, _oracle_sweep, (0.25 if dims == 3 else 0.5) / np.pi


@pytest.mark.parametrize("dims, n, groups", [(1, 24, 1), (1, 20, 3), (2, 8, 2), (3, 4, 2)])
def test_fixed_source_with_fission_is_the_linear_solve(dims, n, groups):
    p, sweep, norm = _setup(dims, n, groups, (0.3, 0.6))
    D, S, F = _operators(p, sweep, norm)
    mat = np.asarray(p.mat).reshape(-1)
    src = np.asarray(p.src).reshape(p.num_mats, groups)
    Q = np.concatenate([norm * src[mat, g] * p.has_src[mat] for g in range(groups)])
    full = np.linalg.solve(np.eye(len(Q)) - D @ (S + F), D @ Q)
    nofis = np.linalg.solve(np.eye(len(Q)) - D @ S, D @ Q)
    assert np.abs(nofis - full).max() > 1e-3 * np.abs(full).max()
    phi, st = ob.solve_fixed_source(p, 1e-13, 5000, 1e-13, 2000)
    assert st["success"]
    # The multigroup loop runs once with one group (TransportSolver3D.cpp:295:
    # "... && _num_groups > 1"), so the fission source, formed from the flux
    # before the pass (zero), never enters: the answer is the fission-free one.
    want = full if groups > 1 else nofis
    assert np.abs(phi.reshape(-1) - want).max() <= 1e-9 * np.abs(want).max()


@pytest.mark.parametrize("dims, n, groups", [(1, 24, 1), (1, 20, 3), (2, 8, 2), (3, 4, 2)])
def test_power_iteration_finds_dominant_eigenvalue(dims, n, groups):
    p, sweep, norm = _setup(dims, n, groups, (1.0, 2.0))
    D, S, F = _operators(p, sweep, norm)
    A = np.linalg.solve(np.eye(D.shape[0]) - D @ S, D @ F)
    vals, vecs = np.linalg.eig(A)
    i = int(np.argmax(vals.real))
    k_exact = float(vals[i].real)
    v = vecs[:, i].real
    v = v * np.sign(v.sum())  # DD fluxes may dip below zero: fix the sign, not abs
    k, phi, st = ob.solve_eigenvalue(p, 1e-13, 5000, 1e-12, 5000, 1e-13)
    assert st["success"] and st["k_eff"] == k
    assert abs(k - k_exact) <= 1e-9 * k_exact, (k, k_exact)
    u = phi.reshape(-1)
    assert np.abs(u / np.linalg.norm(u) - v / np.linalg.norm(v)).max() < 1e-8
    assert st["k_residual"] <= 1e-13


def test_power_iteration_needs_fission():
    p = workloads.random_blocks(8, 1, 1, 1, block=4)
    with pytest.raises(ValueError):
        ob.solve_eigenvalue(p, 1e-6, 100, 1e-6, 100, 1e-6)


def test_prompt_fraction_scales_fission():
    """f = 1 - sum beta multiplies the prompt term (TransportSolver3D.cpp:380-386):
    prompt = 0.5 equals halving nu*sigma_f."""
    p = workloads.fissile_blocks(12, 2, 4, 1, dims=1, nu_ratio=(0.3, 0.6)).normalized()
    half_prompt = p.normalized()
    half_prompt.fission = capi.Fission(p.fission.chi, p.fission.prod, 0.5 * p.fission.prompt,
                                       p.fission.has_chi, p.fission.has_prod)
    half_prod = p.normalized()
    half_prod.fission = capi.Fission(p.fission.chi, 0.5 * p.fission.prod, p.fission.prompt,
                                     p.fission.has_chi, p.fission.has_prod)
    a, _ = ob.solve_fixed_source(half_prompt, 1e-12, 2000, 1e-12, 500)
    b, _ = ob.solve_fixed_source(half_prod, 1e-12, 2000, 1e-12, 500)
    assert np.abs(a - b).max() <= 1e-10 * np.abs(a).max()
