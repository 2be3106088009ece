import json
import os
import sys
from pathlib import Path

import numpy as np
import pytest

# The oracle's OpenMP threads sleep rather than spin while they wait: several
# test processes sharing the host's cores (pytest -n) otherwise spin against
# each other.  Read when the oracle library first loads, so it is set here.
os.environ.setdefault("OMP_WAIT_POLICY", "PASSIVE")

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = ROOT / "tests" / "golden"
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: long-running case")


def load_golden(name):
    """(Problem, npz, manifest entry) for a golden fixture written by make_golden.py."""
    from doma_amd import workloads
    from doma_amd.capi import Problem

    z = np.load(GOLDEN / f"{name}.npz", allow_pickle=False)
    man = json.loads((GOLDEN / "manifest.json").read_text())[name]
    if "mat" in z.files:
        p = Problem(int(z["dims"]), z["lx"], z["ly"], z["lz"], int(z["groups"]), int(z["scheme"]),
                    z["mat"], z["sigt"], z["sigs"], z["src"], z["has_src"], z["mu"], z["eta"],
                    z["xi"], z["w"], name=name)
    else:
        n = len(z["lx"])
        p = workloads.random_blocks(n, int(z["groups"]), int(z["np_"]), int(z["na"]),
                                    int(z["scheme"]), seed=int(z["seed"]))
        assert np.array_equal(p.normalized().sigt, z["sigt"]), "seeded media drifted"
    return p, z, man


def golden_names():
    man = json.loads((GOLDEN / "manifest.json").read_text())
    return sorted(man)


def solve_args(z):
    return (float(z["fs_tol"]), int(z["fs_max_it"]), float(z["mg_tol"]), int(z["mg_max_it"]))


def rel_err(a, b):
    """max |a - b| / max |b| per leading index (group), as an array."""
    a = np.asarray(a).reshape(np.asarray(b).shape[0], -1)
    b = np.asarray(b).reshape(a.shape)
    scale = np.maximum(np.abs(b).max(axis=1), 1e-300)
    return np.abs(a - b).max(axis=1) / scale
