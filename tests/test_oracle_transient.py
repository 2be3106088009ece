"""Pins the oracle's transient (zero initial condition, implicit Euler with
delayed-neutron precursors and step sources) against the reference's own
compiled solveTransient on the reference's transient deck (its first steps, on
the unrefined mesh, and a variant whose step source is removed mid-run;
fixtures from tests/golden/make_deck_golden.py).  CPU only.

Reference behaviour: src/TransportSolver2D.cpp:88-260 (solveTransient,
initZeroIC, updateStepFluxes, stepDNPs), :339-402 (the time terms of the
multigroup source), include/TWDiamondDifference2D.h:25-26 (sigma_t + inv_v/dt).
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

from oracle import binding as ob

GOLDEN = Path(__file__).resolve().parent / "golden"
sys.path.insert(0, str(GOLDEN))
from deck_xml import problem_from_deck, read_deck  # noqa: E402

TRANSIENT_DECKS = ["2D_assembly_tr_short", "2D_assembly_tr_remove"]


@pytest.mark.parametrize("name", TRANSIENT_DECKS)
def test_oracle_transient_matches_reference(name):
    _, p = read_deck(GOLDEN / "decks" / f"{name}.xml")
    ref = np.load(GOLDEN / "decks" / f"{name}.npz")
    prob = problem_from_deck(p)
    ok, phi, dnp, st = ob.solve_transient(prob, p["fs_tol"], p["fs_max_it"], p["mg_tol"],
                                          p["mg_max_it"], p["t0"], p["t1"], p["t_steps"],
                                          nthreads=8)
    assert ok == bool(ref["ref_success"])
    G = p["G"]
    last = [int(v) for v in ref["si_iterations"][-G:]]
    assert st["si_iterations"] == last
    scale = np.abs(ref["phi"]).max(axis=1, keepdims=True)
    assert (np.abs(phi[-1] - ref["phi"]) / scale).max() < 1e-12
    nd = ref["dnps"].shape[0]
    assert (np.abs(dnp[-1, :nd] - ref["dnps"]).max() <= 1e-12 * np.abs(ref["dnps"]).max())
    # every step's flux text matches the reference's at 6 digits
    for step in range(p["t_steps"]):
        for g in range(G):
            want = list(ref[f"text_t{step}_g{g}_flux"])
            got = ["%.6g" % v for v in phi[step, g]]
            bad = [i for i, (a, b) in enumerate(zip(got, want)) if a != b]
            for i in bad:  # only where a value straddles a rounding boundary
                assert abs(float(got[i]) - float(want[i])) <= 1e-5 * abs(float(want[i])), (step, g, i)


def test_step_source_removal_changes_the_flux():
    """The remove_at variant keeps the source through t = 0.02 (t <= remove_at,
    TransportSolver3D.cpp:356) and drops it at t = 0.03."""
    a = np.load(GOLDEN / "decks" / "2D_assembly_tr_short.npz")
    b = np.load(GOLDEN / "decks" / "2D_assembly_tr_remove.npz")
    for step in (0, 1, 2):
        assert list(a[f"text_t{step}_g0_flux"]) == list(b[f"text_t{step}_g0_flux"])
    fa = np.array(a["text_t3_g0_flux"], dtype=float)
    fb = np.array(b["text_t3_g0_flux"], dtype=float)
    assert fb.max() < 1e-2 * fa.max()
