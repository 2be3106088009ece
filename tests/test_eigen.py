"""Fission sources and k-eigenvalue power iteration on the GPU against the
oracle (itself pinned by tests/test_oracle_fission.py and, in 2D/3D, by the
reference's own solveFixedSource/solveEigenvalue on the fission decks in
tests/test_decks.py).

Reference behaviour: src/TransportSolver3D.cpp:29-86 (power iteration),
339-402 (prompt fission in the multigroup source), 404-456 (eigen source and
k update), and the 2D/1D counterparts.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rel_err

import doma_amd
from doma_amd import capi, workloads

# (dims, cells per axis, groups, scheme)
EIGEN_CASES = [
    (3, 8, 2, capi.SCHEME_TW_DD),
    (3, 12, 3, capi.SCHEME_DD),
    (2, 24, 3, capi.SCHEME_TW_DD),
    (2, 16, 2, capi.SCHEME_DD),
    (1, 48, 3, capi.SCHEME_DD),
    (1, 40, 1, capi.SCHEME_DD),
]


def _ids(c):
    return f"{c[0]}d_{c[1]}c_{c[2]}g_{'tw' if c[3] else 'dd'}"


def _problem(dims, n, groups, scheme, nu_ratio, seed=31):
    return workloads.fissile_blocks(n, groups, 2 if dims > 1 else 8, 2 if dims > 1 else 1,
                                    scheme, seed=seed, nu_ratio=nu_ratio, dims=dims)


# --------------------------------------------------------------------- host --

def test_fission_dataclass_layout():
    p = _problem(3, 8, 2, capi.SCHEME_DD, (1.0, 2.0)).normalized()
    f = p.fission
    assert f.chi.shape == f.prod.shape == f.prompt.shape == (p.num_mats, p.groups)
    assert f.has_chi.dtype == np.uint8 and f.has_chi.shape == (p.num_mats,)
    assert np.all(f.chi[f.has_chi == 0] == 0) and np.all(f.prod[f.has_prod == 0] == 0)


# ---------------------------------------------------------------------- GPU --

@pytest.mark.gpu
@pytest.mark.parametrize("case", EIGEN_CASES, ids=_ids)
def test_eigenvalue_matches_oracle(case):
    from oracle import binding as ob

    dims, n, groups, scheme = case
    p = _problem(dims, n, groups, scheme, (1.0, 2.0))
    tol = (1e-8, 2000, 1e-7, 500, 1e-8)
    print("eigen", _ids(case), flush=True)
    with doma_amd.Context(p) as ctx:
        k, phi, st = ctx.solve_eigenvalue(*tol)
    ok, ophi, ost = ob.solve_eigenvalue(p, *tol, nthreads=8)
    assert st["success"] and ost["success"]
    assert st["mg_iterations"] == ost["mg_iterations"]
    assert st["si_iterations"] == ost["si_iterations"]
    assert abs(k - ok) <= 1e-11 * ok, (k, ok)
    assert st["k_eff"] == k and abs(st["k_residual"] - ost["k_residual"]) <= 1e-12
    assert rel_err(phi, ophi).max() < 1e-9


@pytest.mark.gpu
@pytest.mark.parametrize("case", EIGEN_CASES[:5], ids=_ids)
def test_fixed_source_with_fission_matches_oracle(case):
    from oracle import binding as ob

    dims, n, groups, scheme = case
    p = _problem(dims, n, groups, scheme, (0.3, 0.6), seed=32)
    tol = (1e-9, 2000, 1e-9, 500)
    print("fission fixed source", _ids(case), flush=True)
    with doma_amd.Context(p) as ctx:
        phi, st = ctx.solve_fixed_source(*tol)
    ophi, ost = ob.solve_fixed_source(p, *tol, nthreads=8)
    assert st["success"] and ost["success"]
    assert st["mg_iterations"] == ost["mg_iterations"]
    assert rel_err(phi, ophi).max() < 1e-9
    # the fission source is really there: the fission-free answer is far off
    p.fission = None
    with doma_amd.Context(p) as ctx:
        phi0, _ = ctx.solve_fixed_source(*tol)
    assert rel_err(phi0, phi).max() > 1e-3


@pytest.mark.gpu
def test_eigenvalue_fp32_sweep_close():
    """FP32-sweep mode: the same eigenvalue to ~1e-6."""
    p = _problem(3, 12, 2, capi.SCHEME_TW_DD, (1.0, 2.0))
    tol = (1e-7, 2000, 1e-6, 500, 1e-7)
    with doma_amd.Context(p) as ctx:
        k64, phi64, _ = ctx.solve_eigenvalue(*tol)
    with doma_amd.Context(p, precision=capi.FP32_SWEEP) as ctx:
        k32, phi32, st = ctx.solve_eigenvalue(*tol)
    assert st["success"]
    assert abs(k32 - k64) <= 1e-5 * k64
    assert np.linalg.norm(phi32 - phi64) <= 1e-4 * np.linalg.norm(phi64)


@pytest.mark.gpu
def test_eigenvalue_resolve_and_state_errors():
    """Two solves on one context give identical results (the state is fully
    re-initialised); an eigenvalue solve without fission data, mismatched
    material counts and chi without nu*sigma_f are rejected."""
    p = _problem(3, 8, 2, capi.SCHEME_DD, (1.0, 2.0))
    with doma_amd.Context(p) as ctx:
        k1, phi1, _ = ctx.solve_eigenvalue(1e-8, 500, 1e-7, 200, 1e-8)
        ctx.solve_fixed_source(1e-6, 100, 1e-6, 10)
        k2, phi2, _ = ctx.solve_eigenvalue(1e-8, 500, 1e-7, 200, 1e-8)
        assert k1 == k2 and np.array_equal(phi1, phi2)
        f = p.fission.normalized(p.groups)
        u8 = capi._u8p
        rc = capi.lib().doma_set_fission(ctx.h, p.num_mats - 1, capi._d(f.chi), capi._d(f.prod),
                                         capi._d(f.prompt), f.has_chi.ctypes.data_as(u8),
                                         f.has_prod.ctypes.data_as(u8))
        assert rc == capi.DOMA_ERR_INVALID
        bad = f.has_prod.copy()
        bad[int(np.flatnonzero(f.has_chi)[0])] = 0
        rc = capi.lib().doma_set_fission(ctx.h, p.num_mats, capi._d(f.chi), capi._d(f.prod),
                                         capi._d(f.prompt), f.has_chi.ctypes.data_as(u8),
                                         bad.ctypes.data_as(u8))
        assert rc == capi.DOMA_ERR_INVALID
        assert "nu_sigma_f" in capi.lib().doma_last_error().decode()
    p.fission = None
    with doma_amd.Context(p) as ctx:
        with pytest.raises(doma_amd.DomaError) as e:
            ctx.solve_eigenvalue(1e-6, 100, 1e-6, 10, 1e-6)
        assert e.value.code == capi.DOMA_ERR_STATE


@pytest.mark.gpu
def test_eigenvalue_si_failure_stops():
    """A source iteration that hits its limit ends the power iteration with
    success = 0 after that group (TransportSolver3D.cpp:52-57), as the oracle."""
    from oracle import binding as ob

    p = _problem(3, 8, 2, capi.SCHEME_TW_DD, (1.0, 2.0))
    tol = (1e-12, 3, 1e-8, 100, 1e-8)
    with doma_amd.Context(p) as ctx:
        k, phi, st = ctx.solve_eigenvalue(*tol)
    ok, ophi, ost = ob.solve_eigenvalue(p, *tol, nthreads=4)
    assert not st["success"] and not ost["success"]
    assert st["mg_iterations"] == ost["mg_iterations"] == 0
    assert rel_err(phi, ophi).max() < 1e-10
