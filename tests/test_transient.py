"""Transient mode on the GPU: the reference's transient deck (first steps and a
step-removal variant) end to end against the reference's own run, and
synthetic kinetic problems in 1D/2D/3D against the oracle (pinned by
tests/test_oracle_transient.py).

Reference behaviour: src/TransportSolver3D.cpp:88-260 (solveTransient,
initZeroIC, updateStepFluxes, stepDNPs), :339-402 (time terms of the
multigroup source), include/DiamondDifference3D.h:25-26 (sigma_t + inv_v/dt),
src/BrickMesh2D.cpp:171-244 (per-step dumps).
"""
from __future__ import annotations

import hashlib
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

from conftest import rel_err

import doma_amd
from doma_amd import capi, workloads

GOLDEN = Path(__file__).resolve().parent / "golden"
DECKS = GOLDEN / "decks"
sys.path.insert(0, str(GOLDEN))
from deck_xml import read_deck  # noqa: E402

TRANSIENT_DECKS = ["2D_assembly_tr_short", "2D_assembly_tr_remove"]


def _sha(path):
    return hashlib.sha256(Path(path).read_bytes()).hexdigest()


def _same_text(got_path, want_lines, what):
    got = Path(got_path).read_text().splitlines()
    want = list(want_lines)
    assert len(got) == len(want), what
    for i, (a, b) in enumerate(zip(got, want)):
        if a != b:  # only where the two values straddle a rounding boundary
            assert abs(float(a) - float(b)) <= 1e-5 * max(abs(float(b)), 1e-300), (what, i, a, b)


@pytest.mark.gpu
@pytest.mark.parametrize("name", TRANSIENT_DECKS)
def test_transient_deck_matches_reference(name, tmp_path):
    ref = np.load(DECKS / f"{name}.npz")
    _, p = read_deck(DECKS / f"{name}.xml")
    base = tmp_path / name
    print(f"transient deck {name}", flush=True)
    conv, phi, st = doma_amd.run_deck(DECKS / f"{name}.xml", base)
    assert conv == bool(ref["ref_success"])
    G = p["G"]
    assert st["si_iterations"] == [int(v) for v in ref["si_iterations"][-G:]]
    assert rel_err(phi, ref["phi"]).max() < 1e-10
    # the initial-condition dump: mesh files and zero fluxes as the reference's
    for f in ("dims", "blocks", "meshx", "meshy") + tuple(f"g{g}_flux" for g in range(G)):
        assert _sha(f"{base}_{f}.txt") == str(ref[f"sha256_{f}"]), f
    nd = ref["dnps"].shape[0]
    for step in range(p["t_steps"]):
        for g in range(G):
            _same_text(f"{base}_t{step}_g{g}_flux.txt", ref[f"text_t{step}_g{g}_flux"],
                       (step, "g", g))
        for d in range(nd):
            _same_text(f"{base}_t{step}_d{d}_dnps.txt", ref[f"text_t{step}_d{d}_dnps"],
                       (step, "d", d))
    assert not Path(f"{base}_t{p['t_steps']}_g0_flux.txt").exists()


@pytest.mark.gpu
@pytest.mark.parametrize("dims, n, scheme, dsa", [
    (3, 8, capi.SCHEME_TW_DD, False), (3, 12, capi.SCHEME_DD, False),
    (2, 16, capi.SCHEME_TW_DD, False), (1, 40, capi.SCHEME_DD, False),
    (1, 40, capi.SCHEME_DD, True)])
def test_transient_matches_oracle(dims, n, scheme, dsa):
    from oracle import binding as ob

    p = workloads.kinetic_blocks(n, 2, scheme=scheme, dims=dims)
    tol, steps = (1e-10, 2000, 1e-10, 200), 4
    print("transient", dims, n, scheme, dsa, flush=True)
    ok, ophi, odnp, ost = ob.solve_transient(p, *tol, 0.0, 0.04, steps, nthreads=4, dsa=dsa)
    assert ok
    with doma_amd.Context(p) as ctx:
        if dsa:
            ctx.set_dsa(True)
        ctx.transient_begin(0.0, 0.04, steps)
        for k in range(steps):
            t, st = ctx.transient_step(*tol)
            assert st["success"]
            assert abs(t - 0.01 * k) < 1e-15
            phi = ctx.get_flux()
            dnp = ctx.get_dnps()
            assert rel_err(phi, ophi[k]).max() < 1e-9, k
            assert np.abs(dnp - odnp[k, :dnp.shape[0]]).max() <= 1e-9 * np.abs(odnp[k]).max()
        assert st["si_iterations"] == ost["si_iterations"]
        with pytest.raises(doma_amd.DomaError, match="steps have been solved"):
            ctx.transient_step(*tol)


@pytest.mark.gpu
def test_transient_state_errors():
    p = workloads.kinetic_blocks(8, 2, dims=3)
    kin = p.kinetics
    p.kinetics = None
    with doma_amd.Context(p) as ctx:
        with pytest.raises(doma_amd.DomaError) as e:
            ctx.transient_begin(0.0, 1.0, 2)
        assert e.value.code == capi.DOMA_ERR_STATE
        with pytest.raises(doma_amd.DomaError) as e:
            ctx.transient_step(1e-6, 10, 1e-6, 10)
        assert e.value.code == capi.DOMA_ERR_STATE
    p.kinetics = kin
    with doma_amd.Context(p) as ctx:
        with pytest.raises(doma_amd.DomaError):
            ctx.transient_begin(1.0, 0.5, 2)  # t1 < t0
        ctx.transient_begin(0.0, 0.02, 2)
        ctx.transient_step(1e-8, 500, 1e-8, 100)
        # a steady-state solve in between ends the transient
        ctx.solve_fixed_source(1e-6, 100, 1e-6, 10)
        with pytest.raises(doma_amd.DomaError) as e:
            ctx.transient_step(1e-8, 500, 1e-8, 100)
        assert e.value.code == capi.DOMA_ERR_STATE


@pytest.mark.gpu
def test_cli_runs_transient_deck(tmp_path):
    cli = Path(doma_amd.__file__).resolve().parent / "doma"
    deck = tmp_path / "tr.xml"
    deck.write_text((DECKS / "2D_assembly_tr_remove.xml").read_text())
    r = subprocess.run([str(cli), "-i", str(deck)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Transient solver finished executing." in r.stdout
    assert r.stdout.count("Solving timestep") == 4
    assert (tmp_path / "tr_t3_g2_flux.txt").exists() and (tmp_path / "tr_t3_d5_dnps.txt").exists()
