"""Writes tests/golden/percell_spread.json: for every golden fixture, the
reference's own per-cell reproducibility -- the largest relative difference,
over cells with |phi| > 1e-10 of the group's largest, between the fixture (the
reference run with its recorded OpenMP thread count) and the same solve with 1,
2, 3 and 8 threads.  The reference sums each cell's scalar flux over per-thread
partials in thread order (src/TransportSolver3D.cpp:609-621), so its result
depends on the thread count at the rounding level; cells whose signed angular
fluxes cancel (diamond difference) or whose theta-weighted branches sit at their
1e-8 threshold move far more than one rounding unit.  The oracle reproduces the
reference bit for bit at any thread count (test_oracle_golden.py), so it stands
in for those runs.  Run in the build container:

    python tests/golden/make_percell_spread.py
"""
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))
sys.path.insert(0, str(HERE.parent))

from conftest import golden_names, load_golden, solve_args  # noqa: E402
from oracle import binding as ob  # noqa: E402

CUT = 1e-10


def spread(phi, ref):
    worst = 0.0
    for g in range(ref.shape[0]):
        big = np.abs(ref[g]) > CUT * np.abs(ref[g]).max()
        if big.any():
            worst = max(worst, float((np.abs(phi[g][big] - ref[g][big]) / np.abs(ref[g][big])).max()))
    return worst


out = {}
for name in golden_names():
    p, z, _ = load_golden(name)
    idx = z["sample_idx"] if "sample_idx" in z.files else None
    ref = z["phi_sample"] if idx is not None else z["phi"].reshape(p.groups, -1)
    args = solve_args(z)
    worst = 0.0
    for t in (1, 2, 3, 8):
        phi, _ = ob.solve_fixed_source(p, *args, nthreads=t)
        worst = max(worst, spread(phi[:, idx] if idx is not None else phi, ref))
    out[name] = worst
    print(f"{name}: {worst:.2e}", flush=True)
(HERE / "percell_spread.json").write_text(json.dumps(out, indent=1) + "\n")
