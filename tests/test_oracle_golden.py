"""The CPU oracle against the golden vectors written by the reference's own
compiled TransportSolver2D/3D (tests/golden/make_golden.py).  The oracle keeps
the reference's expression order and OpenMP partition, so at the fixture's
thread count it must agree bit for bit, iteration counts included."""
import numpy as np
import pytest

from conftest import golden_names, load_golden, rel_err, solve_args
from oracle import binding as ob


@pytest.mark.parametrize("name", golden_names())
def test_oracle_matches_reference(name):
    p, z, man = load_golden(name)
    fs_tol, fs_max, mg_tol, mg_max = solve_args(z)
    phi, st = ob.solve_fixed_source(p, fs_tol, fs_max, mg_tol, mg_max,
                                    nthreads=int(z["nthreads"]))
    assert st["success"] == bool(z["ref_success"])
    if "phi" in z.files:
        np.testing.assert_array_equal(phi, z["phi"])
    else:
        idx = z["sample_idx"]
        assert np.all(rel_err(phi[:, idx], z["phi_sample"]) <= 1e-13)
        np.testing.assert_allclose(np.sqrt((phi ** 2).sum(axis=1)), z["phi_group_l2"], rtol=1e-13)
    # The reference logs "SI converged after k iterations ... for Gg" per group.
    log = man["log"]["si"]
    if log and st["success"]:
        G = p.groups
        last = {g: k for k, _, g in log[-G:]}
        assert [last[g] for g in range(G)] == st["si_iterations"]
    if man["log"]["mg"] and st["success"]:
        assert man["log"]["mg"][-1][0] == st["mg_iterations"]


def test_oracle_thread_count_only_reorders_sum():
    """Thread count changes only the per-cell partial-sum order (1e-15 level)."""
    p, z, _ = load_golden("twdd3d_small")
    fs_tol, fs_max, mg_tol, mg_max = solve_args(z)
    a, sa = ob.solve_fixed_source(p, fs_tol, fs_max, mg_tol, mg_max, nthreads=1)
    b, sb = ob.solve_fixed_source(p, fs_tol, fs_max, mg_tol, mg_max, nthreads=4)
    assert sa["si_iterations"] == sb["si_iterations"]
    assert np.all(rel_err(a, b) < 1e-13)


def test_planted_transpose_is_caught():
    """A transposed scattering matrix (a plausible mistake) must miss the golden
    vector: the fixture has up-scatter, so the transpose changes the answer."""
    p, z, _ = load_golden("twdd3d_small")
    q = p.normalized()
    q.sigs = np.ascontiguousarray(np.transpose(q.sigs, (0, 2, 1)))
    fs_tol, fs_max, mg_tol, mg_max = solve_args(z)
    phi, _ = ob.solve_fixed_source(q, fs_tol, fs_max, mg_tol, mg_max)
    assert np.max(rel_err(phi, z["phi"])) > 1e-3
