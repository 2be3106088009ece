"""The device-side stopping rule of source iteration (chunks of iterations per
host round trip, kernels after the last iteration returning at entry) against
the per-iteration host loop: identical iterates (bitwise), iteration counts and
residuals, for the 3D brick sweep schedules, 1D with DSA, eigenvalue and transient
solves, and source iterations that hit the iteration limit."""
from __future__ import annotations

import numpy as np
import pytest

import doma_amd
from doma_amd import capi, workloads


def _both(problem, run, tuning=None, dsa=False):
    out = []
    for device_loop in (False, True):
        with doma_amd.Context(problem, tuning=tuning) as ctx:
            ctx.set_si_mode(device_loop)
            if dsa:
                ctx.set_dsa(True)
            out.append(run(ctx))
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("schedule", [0, 1, 3])
def test_fixed_source_identical(schedule):
    p = workloads.random_blocks(24, 3, 2, 2, block=8)
    print("device loop, schedule", schedule, flush=True)
    (phi0, st0), (phi1, st1) = _both(p, lambda c: c.solve_fixed_source(1e-9, 500, 1e-9, 50),
                                     tuning=dict(schedule=schedule))
    assert np.array_equal(phi0, phi1)
    for k in ("si_iterations", "mg_iterations", "sweeps", "si_residual", "mg_residual"):
        assert st0[k] == st1[k], k


@pytest.mark.gpu
@pytest.mark.parametrize("max_it", [1, 2, 3, 7, 9])
def test_iteration_limit_identical(max_it):
    """SI that stops at max_it (not converged) inside and at the end of a chunk."""
    p = workloads.random_blocks(16, 2, 2, 2, block=8)
    (phi0, st0), (phi1, st1) = _both(p, lambda c: c.solve_fixed_source(1e-14, max_it, 1e-14, 3))
    assert np.array_equal(phi0, phi1) and st0 == {**st1, "elapsed_ms": st0["elapsed_ms"],
                                                  "sweep_ms": st0["sweep_ms"]}
    assert not st1["success"] and st1["si_iterations"][0] == max_it


@pytest.mark.gpu
def test_dsa_eigen_transient_identical():
    slab = workloads.fissile_blocks(40, 2, 8, 1, dims=1, nu_ratio=(1.0, 2.0))
    (k0, phi0, st0), (k1, phi1, st1) = _both(
        slab, lambda c: c.solve_eigenvalue(1e-10, 3000, 1e-9, 500, 1e-10), dsa=True)
    assert k0 == k1 and np.array_equal(phi0, phi1) and st0["sweeps"] == st1["sweeps"]
    p = workloads.kinetic_blocks(8, 2, dims=3)

    def tr(c):
        c.transient_begin(0.0, 0.03, 3)
        res = []
        for _ in range(3):
            t, st = c.transient_step(1e-9, 500, 1e-9, 100)
            res.append((t, st["sweeps"], c.get_flux(), c.get_dnps()))
        return res
    a, b = _both(p, tr)
    for (ta, sa, fa, da), (tb, sb, fb, db) in zip(a, b):
        assert ta == tb and sa == sb and np.array_equal(fa, fb) and np.array_equal(da, db)


@pytest.mark.gpu
@pytest.mark.parametrize("device_loop", [False, True])
def test_sweep_time_statistic(device_loop):
    """doma_stats.sweep_ms times the first two sweeps of each group's iteration
    and counts the later ones at the second one's duration: per sweep it stays
    within a small factor of the same kernel timed on its own."""
    p = workloads.random_blocks(24, 2, 2, 2, block=8)
    with doma_amd.Context(p, tuning=dict(schedule=1)) as ctx:
        ctx.set_si_mode(device_loop)
        _, st = ctx.solve_fixed_source(1e-8, 500, 1e-8, 20)
        k_ms, _, _ = ctx.profile_sweeps(0, 5)  # kernel ms per sweep
    assert st["success"] and st["sweeps"] > 10
    per = st["sweep_ms"] / st["sweeps"]
    print(f"device loop {device_loop}: {per:.4f} ms per sweep from the statistic, "
          f"{k_ms:.4f} ms timed alone", flush=True)
    assert 0.2 * k_ms < per < 5.0 * k_ms
