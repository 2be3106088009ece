"""The oracle against the reference's own recorded outputs: DOMA's diamond-
difference values on Kobayashi problem 1, lines 1a, 1b and 1c
(results/line_plot_results.py:23-26, 127-130, 231-234; deck
cases/3D_kobayashi_1.xml).  A solve to a 1e-7 scattering residual reproduces
every recorded value to the 6 digits it was printed with (kobayashi_lines), so
these rows are a known-answer test of the whole fixed-source path -- product
quadrature, mesh expansion, DD cell solve, sweep order and source iteration.
The deck's own tolerance (1e-3) or its scheme (TW-DD) miss them by percents,
which fixes the scheme and the stopping point the recorded run used.  CPU only;
the GPU counterpart in test_kobayashi.py covers every recorded mesh and
quadrature.
"""
from __future__ import annotations

import pytest

from doma_amd import capi, workloads
from kobayashi_lines import worst
from oracle import binding as ob


@pytest.mark.parametrize("refinement", [0, 1])
def test_oracle_reproduces_recorded_dd_lines(refinement):
    p = workloads.kobayashi1(refinement, 10, 10).normalized()
    phi, st = ob.solve_fixed_source(p, 1e-7, 1000, 1e-7, 10, nthreads=8)
    assert st["success"]
    dev = worst(p, phi[0], 10, refinement)
    print(f"refinement {refinement}: SI {st['si_iterations']}, worst deviation {dev}", flush=True)
    assert max(dev.values()) <= 1.0, dev


def test_loose_tolerance_or_twdd_miss_the_recorded_lines():
    p = workloads.kobayashi1(0, 10, 10).normalized()
    loose, _ = ob.solve_fixed_source(p, 1e-3, 100, 1e-3, 10, nthreads=8)
    assert worst(p, loose[0], 10, 0)["1a"] > 100.0
    p.scheme = capi.SCHEME_TW_DD
    tw, _ = ob.solve_fixed_source(p, 1e-7, 1000, 1e-7, 10, nthreads=8)
    assert worst(p, tw[0], 10, 0)["1a"] > 1000.0
