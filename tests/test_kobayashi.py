"""The HIP path against the reference's recorded outputs: DOMA's diamond-
difference line values on Kobayashi problem 1 (results/line_plot_results.py;
cases/3D_kobayashi_1.xml), lines 1a, 1b and 1c, for every recorded mesh (20^3,
40^3, 60^3, 80^3 cells with np = na = 10, 800 directions) and quadrature
(np = na = 2, 4, 6, 8 on 80^3).  A solve to a 1e-7 scattering residual must
reproduce each recorded value to its 6 printed digits (kobayashi_lines); the oracle
does the same on the CPU (test_oracle_kobayashi.py)."""
from __future__ import annotations

import pytest

from doma_amd import workloads
from doma_amd.capi import Context
from kobayashi_lines import worst

pytestmark = pytest.mark.gpu

CASES = [(r, 10) for r in range(4)] + [(3, a) for a in (2, 4, 6, 8)]


@pytest.mark.parametrize("refinement, angle", CASES)
def test_recorded_dd_lines(refinement, angle):
    p = workloads.kobayashi1(refinement, angle, angle)
    print(f"case kobayashi1 refinement {refinement} np=na={angle} N={p.N} "
          f"dirs={p.noct * p.dirs}", flush=True)
    with Context(p) as ctx:
        phi, st = ctx.solve_fixed_source(1e-7, 1000, 1e-7, 10)
    assert st["success"]
    dev = worst(p, phi[0], angle, refinement)
    print(f"  SI {st['si_iterations']}, worst deviation per line in units of the bar {dev}",
          flush=True)
    assert max(dev.values()) <= 1.0, dev
