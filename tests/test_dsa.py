"""1D diffusion synthetic acceleration on the GPU (dsa1d_kernel: moments from
the sweep's direction partials, Alcouffe's tridiagonal system and its pivoted
solve in one work-group) against the oracle's restatement
(tests/test_oracle_dsa.py pins it).  Reference behaviour:
src/TransportSolver1D.cpp:495-541 (DSA after each sweep), 587-800.
"""
from __future__ import annotations

import subprocess
from pathlib import Path

import numpy as np
import pytest

from conftest import rel_err

import doma_amd
from doma_amd import capi, workloads
from doma_amd.capi import Problem


def _slab(n=120, G=1, c=(0.95, 0.6), np_=8, seed=0):
    rng = np.random.default_rng(seed)
    lx = rng.uniform(0.05, 0.3, n)
    mat = (np.arange(n) >= n // 2).astype(np.int32)
    sigt = np.array([[1.0 + 0.2 * g for g in range(G)], [2.5 - 0.3 * g for g in range(G)]])
    sigs = np.zeros((2, G, G))
    for m in range(2):
        for g in range(G):
            sigs[m, g, g] = c[m] * sigt[m, g]
            if g + 1 < G:
                sigs[m, g + 1, g] = 0.5 * (1 - c[m]) * sigt[m, g]
    src = np.zeros((2, G))
    src[0, 0], src[1, 0] = 1.0, 0.3
    mu, eta, xi, w = workloads.quadrature_1d(np_)
    return Problem(1, lx, np.ones(1), np.ones(1), G, capi.SCHEME_DD, mat, sigt, sigs, src,
                   np.ones(2, np.uint8), mu, eta, xi, w, name=f"dsa_slab_{n}c_{G}g")


@pytest.mark.gpu
@pytest.mark.parametrize("n, G", [(120, 1), (300, 3), (2000, 2)])
def test_dsa_fixed_source_matches_oracle(n, G):
    from oracle import binding as ob

    p = _slab(n, G, seed=n)
    tol = (1e-10, 5000, 1e-10, 200)
    print("dsa fixed source", n, G, flush=True)
    with doma_amd.Context(p) as ctx:
        ctx.set_dsa(True)
        phi, st = ctx.solve_fixed_source(*tol)
        ctx.set_dsa(False)
        phi_plain, st_plain = ctx.solve_fixed_source(*tol)
    ophi, ost = ob.solve_fixed_source(p, *tol, dsa=True)
    assert st["success"] and ost["success"]
    assert st["si_iterations"] == ost["si_iterations"]
    assert st["mg_iterations"] == ost["mg_iterations"]
    assert rel_err(phi, ophi).max() < 1e-9
    # acceleration: far fewer sweeps than plain SI
    assert 3 * st["sweeps"] < st_plain["sweeps"]


@pytest.mark.gpu
@pytest.mark.parametrize("G", [1, 3])
def test_dsa_eigenvalue_matches_oracle(G):
    from oracle import binding as ob

    p = workloads.fissile_blocks(48, G, 8, 1, dims=1, nu_ratio=(1.0, 2.0), seed=41)
    tol = (1e-10, 5000, 1e-9, 1000, 1e-10)
    print("dsa eigen", G, flush=True)
    with doma_amd.Context(p) as ctx:
        ctx.set_dsa(True)
        k, phi, st = ctx.solve_eigenvalue(*tol)
    ok, ophi, ost = ob.solve_eigenvalue(p, *tol, dsa=True)
    assert st["success"] == ost["success"]
    assert st["mg_iterations"] == ost["mg_iterations"]
    assert abs(k - ok) <= 1e-10 * ok, (k, ok)
    assert rel_err(phi, ophi).max() < 1e-8


@pytest.mark.gpu
def test_dsa_argument_errors():
    p3 = workloads.random_blocks(8, 1, 1, 1, block=4)
    with doma_amd.Context(p3) as ctx:
        with pytest.raises(doma_amd.DomaError, match="1D only"):
            ctx.set_dsa(True)
        ctx.set_dsa(False)  # turning it off is always fine
    void = workloads.slab_five_region()  # region 3 has sigma_t = 0
    with doma_amd.Context(void) as ctx:
        with pytest.raises(doma_amd.DomaError, match="Void region detected"):
            ctx.set_dsa(True)


@pytest.mark.gpu
def test_cli_runs_dsa_eigen_deck(tmp_path):
    """doma -i on the reference's DSA eigenvalue deck (1D_NPRE_555_mp1)."""
    decks = Path(__file__).resolve().parent / "golden" / "decks"
    cli = Path(doma_amd.__file__).resolve().parent / "doma"
    deck = tmp_path / "mp1.xml"
    deck.write_text((decks / "1D_NPRE_555_mp1.xml").read_text())
    r = subprocess.run([str(cli), "-i", str(deck)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Eigenvalue solver finished executing." in r.stdout
    assert "Converged value of k = " in r.stdout
    assert (tmp_path / "mp1_g0_flux.txt").exists()
