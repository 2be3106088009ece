"""Batched solves: many uneven 3D problems in one persistent launch per round.

A problem solved in a batch must give exactly (bitwise) the iterates of the
same problem solved alone with the same kernel shape; the solo path is in turn
pinned to the reference (test_gpu_parity.py), and one batch member is also
checked against the oracle here.  Training snapshots are checked against solo
source iterations stopped at the snapshot's iteration count."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rel_err

import doma_amd
from doma_amd import capi, workloads
from doma_amd.capi import MediaSpec


def _problems(sizes, groups, seed=77, c=(0.3, 0.9)):
    out = []
    for b, (n, g) in enumerate(zip(sizes, groups)):
        spec = MediaSpec(n=n, block=8, groups=g, sigt_log=0, sigt_lo=0.5, sigt_hi=2.0,
                         c_lo=c[0], c_hi=c[1], scatter_rule=capi.SCATTER_DOWN1,
                         source_rule=capi.SOURCE_BLOCKS, src_lo=0, src_hi=0,
                         source_fraction=0.1, cell_width=1.0, seed=seed * 1000 + b)
        out.append(workloads.block_media_problem(n, g, 2, 2, capi.SCHEME_TW_DD, spec,
                                                 f"b{b}_{n}c_{g}g"))
    return out


# --------------------------------------------------------------------- host --

def test_batch_rejects_mixed_problems():
    a = workloads.random_blocks(16, 1, 2, 2)
    b = workloads.random_blocks(16, 1, 2, 1)  # other quadrature
    with pytest.raises(doma_amd.DomaError) as e:
        doma_amd.Batch([a, b])
    assert e.value.code == capi.DOMA_ERR_INVALID
    c = workloads.slab_five_region()
    with pytest.raises(doma_amd.DomaError) as e:
        doma_amd.Batch([a, c])
    assert e.value.code == capi.DOMA_ERR_INVALID


# ---------------------------------------------------------------------- GPU --

@pytest.mark.gpu
def test_batch_matches_solo_bitwise():
    probs = _problems([16, 24, 32, 16, 24], [1, 2, 3, 2, 1])
    print("batch of 5", flush=True)
    with doma_amd.Batch(probs) as b:
        b.set_tuning(shape=0)
        t = b.get_tuning()
        stats, info = b.solve(1e-9, 500, 1e-9, 50)
        phis = [b.get_flux(i) for i in range(len(probs))]
    assert info["rounds"] == max(s["sweeps"] for s in stats)
    for i, p in enumerate(probs):
        print(f"solo {p.name} shape {t['shape']}", flush=True)
        with doma_amd.Context(p, tuning=dict(shape=t["shape"], schedule=1)) as ctx:
            phi, st = ctx.solve_fixed_source(1e-9, 500, 1e-9, 50)
        assert stats[i]["success"] and st["success"]
        assert stats[i]["si_iterations"] == st["si_iterations"], (i, stats[i], st)
        assert stats[i]["mg_iterations"] == st["mg_iterations"]
        assert np.array_equal(phis[i], phi), (i, rel_err(phis[i], phi))


@pytest.mark.gpu
def test_batch_member_matches_oracle():
    from oracle import binding as ob

    probs = _problems([16, 24], [2, 1], seed=78)
    with doma_amd.Batch(probs) as b:
        stats, _ = b.solve(1e-10, 500, 1e-10, 50)
        phi = b.get_flux(0)
    ophi, ost = ob.solve_fixed_source(probs[0].normalized(), 1e-10, 500, 1e-10, 50, nthreads=4)
    assert stats[0]["si_iterations"] == ost["si_iterations"]
    assert rel_err(phi, ophi).max() < 1e-10


@pytest.mark.gpu
def test_batch_autotune_and_uneven_sizes():
    probs = _problems([16, 24, 32, 48, 16, 24, 32, 16], [1, 2, 4, 1, 2, 4, 1, 2], seed=79,
                      c=(0.9, 0.99))
    with doma_amd.Batch(probs) as b:
        t, ms = b.autotune()
        assert t["shape"] > 0 and ms > 0
        ms2, units = b.time_sweeps(2)
        assert units == sum(p.N * 8 * p.dirs for p in probs)
        stats, info = b.solve(1e-6, 2000, 1e-6, 50, sample_every=5, sample_cap=8)
    assert all(s["success"] for s in stats)
    assert info["samples"] > 0 and info["units"] > 0 and info["sweep_ms"] > 0
    print("batch info", info, "tuning", t, flush=True)


@pytest.mark.gpu
def test_training_samples():
    """Snapshots: every 5 iterations, thinned to at most `cap`; residual = the
    latest increment, target = phi_converged - phi_k."""
    p = _problems([16], [1], seed=80, c=(0.9, 0.99))[0]
    cap = 4
    with doma_amd.Batch([p]) as b:
        stats, info = b.solve(1e-8, 2000, 1e-8, 10, sample_every=5, sample_cap=cap)
        its = b.sample_iterations(0, 0)
        samples = [b.get_sample(0, 0, k) for k in range(len(its))]
        phi_conv = b.get_flux(0)[0]
        shape = b.get_tuning()["shape"]
    n_it = stats[0]["si_iterations"][0]
    assert n_it >= 15
    assert 2 <= len(its) <= cap
    stride = its[0]
    assert stride % 5 == 0
    assert its == [stride * (j + 1) for j in range(len(its))]
    assert its[-1] <= n_it < its[-1] + 2 * stride
    q = np.where(p.has_src[p.mat], 0.25 * p.src[p.mat, 0] / np.pi, 0.0)
    with doma_amd.Context(p, tuning=dict(shape=shape, schedule=1)) as ctx:
        for k, it in enumerate(its):
            phi_k, n, _, _ = ctx.source_iteration(0, q, 0.0, it)
            phi_km1, _, _, _ = ctx.source_iteration(0, q, 0.0, it - 1)
            assert n == it
            res, tgt = samples[k]
            assert np.array_equal(tgt, (phi_conv - phi_k).astype(np.float32)), k
            scale = np.abs(phi_k - phi_km1).max()
            assert np.abs(res - (phi_k - phi_km1)).max() <= 1e-6 * scale + 1e-12 * np.abs(phi_k).max()


@pytest.mark.gpu
def test_uneven_batch_high_scattering_members_match_oracle():
    """The batched-data-generation media (c ~ U[0.9, 0.99] per block, 128
    directions, cubes of 16..64 cells, 1/2/4 groups): every member converges in
    one batched solve, and the two 16^3 members (one and four groups) equal the
    oracle's solves at 1e-10 with the same SI counts."""
    from oracle import binding as ob

    sizes, groups = [16, 16, 24, 32, 48, 64], [1, 4, 2, 1, 2, 4]
    probs = []
    for b, (n, g) in enumerate(zip(sizes, groups)):
        spec = MediaSpec(n=n, block=8, groups=g, sigt_log=0, sigt_lo=0.5, sigt_hi=2.0, c_lo=0.9,
                         c_hi=0.99, scatter_rule=capi.SCATTER_DOWN1,
                         source_rule=capi.SOURCE_BLOCKS, src_lo=0, src_hi=0,
                         source_fraction=0.05, cell_width=1.0, seed=workloads.SEED_BASE + 90 + b)
        probs.append(workloads.block_media_problem(n, g, 4, 4, capi.SCHEME_TW_DD, spec,
                                                   f"uneven_{b}_{n}c_{g}g"))
    print("batch of 6 (128 directions)", flush=True)
    with doma_amd.Batch(probs) as b:
        b.autotune()
        stats, info = b.solve(1e-10, 5000, 1e-10, 50, sample_every=5, sample_cap=8)
        phis = [b.get_flux(i) for i in (0, 1)]
    assert all(s["success"] for s in stats), [s["si_iterations"] for s in stats]
    assert info["samples"] > 0
    for i in (0, 1):
        print(f"oracle {probs[i].name}", flush=True)
        ophi, ost = ob.solve_fixed_source(probs[i].normalized(), 1e-10, 5000, 1e-10, 50,
                                          nthreads=16)
        print(f"  SI {stats[i]['si_iterations']} (oracle {ost['si_iterations']}), "
              f"max rel {rel_err(phis[i], ophi).max():.2e}", flush=True)
        assert stats[i]["si_iterations"] == ost["si_iterations"]
        assert rel_err(phis[i], ophi).max() < 1e-10
