"""ctypes binding of the CPU oracle (oracle/libdoma_oracle.so).

TEST INFRASTRUCTURE: only tests, the smoke check and the benchmark's CPU
baseline use this module; the product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import struct
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libdoma_oracle.so"
REF_HARNESS = HERE / "_ref" / "ref_harness"

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_u8p = C.POINTER(C.c_uint8)


class OracleProblem(C.Structure):
    _fields_ = [
        ("dims", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("lx", _dp), ("ly", _dp), ("lz", _dp), ("groups", C.c_int), ("scheme", C.c_int),
        ("num_mats", C.c_int), ("mat", _ip), ("sigt", _dp), ("sigs", _dp), ("src", _dp),
        ("has_src", _u8p), ("dirs", C.c_int), ("mu", _dp), ("eta", _dp), ("xi", _dp),
        ("w", _dp), ("chi", _dp), ("prod", _dp), ("prompt", _dp), ("has_chi", _u8p),
        ("has_prod", _u8p), ("dsa", C.c_int), ("inv_v", _dp), ("nd", _ip), ("chi_d", _dp),
        ("beta", _dp), ("lambda_", _dp), ("step_type", _ip), ("step_in", _dp), ("step_out", _dp),
        ("step_src", _dp),
    ]


class OracleStats(C.Structure):
    _fields_ = [
        ("success", C.c_int), ("mg_iterations", C.c_int), ("mg_residual", C.c_double),
        ("sweeps", C.c_int), ("si_iterations_last", C.c_int * 64),
        ("si_residual_last", C.c_double * 64), ("k_eff", C.c_double),
        ("k_residual", C.c_double),
    ]


_LIB = None


def build():
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib():
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            build()
        h = C.CDLL(str(LIB_PATH))
        h.oracle_legendre.argtypes = [C.c_int, _dp, _dp]
        h.oracle_chebyshev.argtypes = [C.c_int, _dp, _dp, _dp]
        h.oracle_gc_quadrature.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp]
        h.oracle_gl_quadrature.argtypes = [C.c_int, _dp, _dp]
        for nm in ("oracle_twdd3d_cell", "oracle_dd3d_cell"):
            f = getattr(h, nm)
            f.restype = C.c_double
            f.argtypes = [C.c_double] * 8 + [_dp, _dp]
        h.oracle_twdd2d_cell.restype = C.c_double
        h.oracle_twdd2d_cell.argtypes = [C.c_double] * 7 + [_dp, _dp]
        h.oracle_dd2d_cell.restype = C.c_double
        h.oracle_dd2d_cell.argtypes = [C.c_double] * 6 + [_dp, _dp]
        h.oracle_dd1d_cell.restype = C.c_double
        h.oracle_dd1d_cell.argtypes = [C.c_double] * 5 + [_dp]
        h.oracle_sweep.argtypes = [C.POINTER(OracleProblem), C.c_int, _dp, _dp, C.c_int]
        h.oracle_solve_fixed_source.argtypes = [C.POINTER(OracleProblem), C.c_double, C.c_int,
                                                C.c_double, C.c_int, C.c_int, _dp,
                                                C.POINTER(OracleStats)]
        h.oracle_solve_eigenvalue.argtypes = [C.POINTER(OracleProblem), C.c_double, C.c_int,
                                              C.c_double, C.c_int, C.c_double, C.c_int, _dp,
                                              _dp, C.POINTER(OracleStats)]
        h.oracle_solve_transient.argtypes = [C.POINTER(OracleProblem), C.c_double, C.c_int,
                                             C.c_double, C.c_int, C.c_double, C.c_double,
                                             C.c_int, C.c_int, _dp, _dp, C.POINTER(OracleStats)]
        h.oracle_tridiag_solve.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp]
        h.oracle_dsa_system.argtypes = [C.POINTER(OracleProblem), C.c_int, C.c_int] + [_dp] * 9
        h.oracle_source_iteration.argtypes = [C.POINTER(OracleProblem), C.c_int, _dp, C.c_double,
                                              C.c_int, C.c_int, _dp, _ip, _dp, _dp]
        _LIB = h
    return _LIB


def _d(a):
    return a.ctypes.data_as(_dp)


class _Bound:
    """Keeps the numpy arrays alive while the C struct points at them."""

    def __init__(self, p, dsa=False):
        p = p.normalized()
        self.p = p
        self.s = OracleProblem(p.dims, p.nx, p.ny, p.nz, _d(p.lx), _d(p.ly), _d(p.lz), p.groups,
                               p.scheme, p.num_mats, p.mat.ctypes.data_as(_ip), _d(p.sigt),
                               _d(p.sigs), _d(p.src), p.has_src.ctypes.data_as(_u8p), p.dirs,
                               _d(p.mu), _d(p.eta), _d(p.xi), _d(p.w))
        f = p.fission
        if f is not None:
            self.s.chi, self.s.prod, self.s.prompt = _d(f.chi), _d(f.prod), _d(f.prompt)
            self.s.has_chi = f.has_chi.ctypes.data_as(_u8p)
            self.s.has_prod = f.has_prod.ctypes.data_as(_u8p)
        self.s.dsa = int(dsa)
        kn = p.kinetics
        if kn is not None:
            self.s.inv_v, self.s.nd = _d(kn.inv_v), kn.nd.ctypes.data_as(_ip)
            self.s.chi_d, self.s.beta, self.s.lambda_ = _d(kn.chi_d), _d(kn.beta), _d(kn.lambda_)
        st = p.steps
        if st is not None:
            self.s.step_type = st.type.ctypes.data_as(_ip)
            self.s.step_in, self.s.step_out, self.s.step_src = (_d(st.insert_at),
                                                                _d(st.remove_at), _d(st.src))


def legendre(n):
    r, w = np.zeros(n), np.zeros(n)
    lib().oracle_legendre(n, _d(r), _d(w))
    return r, w


def gc_quadrature(n_c, n_l, dims):
    noct = 8 if dims == 3 else 4
    per = n_c * n_l // 4
    a = [np.zeros(noct * max(per, 1)) for _ in range(4)]
    r = lib().oracle_gc_quadrature(n_c, n_l, dims, *[_d(x) for x in a])
    if r < 0:
        raise ValueError("zero direction component")
    return tuple(x.reshape(noct, per) for x in a)


def gl_quadrature(n):
    mu, w = np.zeros(n), np.zeros(n)
    lib().oracle_gl_quadrature(n, _d(mu), _d(w))
    return mu.reshape(2, n // 2), w.reshape(2, n // 2)


def sweep(problem, g, q, nthreads=1):
    b = _Bound(problem)
    q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1)
    out = np.zeros(b.p.N)
    rc = lib().oracle_sweep(C.byref(b.s), g, _d(q), _d(out), nthreads)
    if rc:
        raise RuntimeError("oracle_sweep failed")
    return out


def solve_fixed_source(problem, fs_tol, fs_max_it, mg_tol, mg_max_it, nthreads=1, dsa=False):
    b = _Bound(problem, dsa)
    phi = np.zeros(b.p.groups * b.p.N)
    st = OracleStats()
    ok = lib().oracle_solve_fixed_source(C.byref(b.s), fs_tol, fs_max_it, mg_tol, mg_max_it,
                                         nthreads, _d(phi), C.byref(st))
    if ok < 0:
        raise RuntimeError("oracle_solve_fixed_source failed")
    stats = {"success": bool(st.success), "mg_iterations": st.mg_iterations,
             "mg_residual": st.mg_residual, "sweeps": st.sweeps,
             "si_iterations": list(st.si_iterations_last[: b.p.groups]),
             "si_residual": list(st.si_residual_last[: b.p.groups])}
    return phi.reshape(b.p.groups, b.p.N), stats


def solve_eigenvalue(problem, fs_tol, fs_max_it, mg_tol, mg_max_it, k_tol, nthreads=1,
                     dsa=False):
    """Power iteration: (k, phi[G, N], stats)."""
    b = _Bound(problem, dsa)
    if b.p.fission is None:
        raise ValueError("an eigenvalue problem needs fission data")
    phi = np.zeros(b.p.groups * b.p.N)
    st, k = OracleStats(), C.c_double()
    ok = lib().oracle_solve_eigenvalue(C.byref(b.s), fs_tol, fs_max_it, mg_tol, mg_max_it, k_tol,
                                       nthreads, _d(phi), C.byref(k), C.byref(st))
    if ok < 0:
        raise RuntimeError("oracle_solve_eigenvalue failed")
    stats = {"success": bool(st.success), "mg_iterations": st.mg_iterations,
             "mg_residual": st.mg_residual, "sweeps": st.sweeps,
             "si_iterations": list(st.si_iterations_last[: b.p.groups]),
             "si_residual": list(st.si_residual_last[: b.p.groups]),
             "k_eff": st.k_eff, "k_residual": st.k_residual}
    return k.value, phi.reshape(b.p.groups, b.p.N), stats


def solve_transient(problem, fs_tol, fs_max_it, mg_tol, mg_max_it, t0, t1, steps, nthreads=1,
                    dsa=False):
    """Zero-IC transient: (ok, phi[steps, G, N], dnps[steps, 8, N], stats of the last step)."""
    b = _Bound(problem, dsa)
    if b.p.kinetics is None:
        raise ValueError("a transient needs kinetics data")
    G, N = b.p.groups, b.p.N
    phi = np.zeros(steps * G * N)
    dnp = np.zeros(steps * 8 * N)
    st = OracleStats()
    ok = lib().oracle_solve_transient(C.byref(b.s), fs_tol, fs_max_it, mg_tol, mg_max_it, t0, t1,
                                      steps, nthreads, _d(phi), _d(dnp), C.byref(st))
    if ok < 0:
        raise RuntimeError("oracle_solve_transient failed")
    stats = {"success": bool(st.success), "mg_iterations": st.mg_iterations,
             "mg_residual": st.mg_residual, "sweeps": st.sweeps,
             "si_iterations": list(st.si_iterations_last[:G]),
             "si_residual": list(st.si_residual_last[:G])}
    return bool(ok), phi.reshape(steps, G, N), dnp.reshape(steps, 8, N), stats


def source_iteration(problem, g, qmg, fs_tol, fs_max_it, nthreads=1, trace=False):
    b = _Bound(problem)
    qmg = np.ascontiguousarray(qmg, dtype=np.float64).reshape(-1)
    phi = np.zeros(b.p.N)
    its, res = C.c_int32(), C.c_double()
    tr = np.zeros(fs_max_it) if trace else None
    conv = lib().oracle_source_iteration(C.byref(b.s), g, _d(qmg), fs_tol, fs_max_it, nthreads,
                                         _d(phi), C.byref(its), C.byref(res),
                                         _d(tr) if trace else None)
    if conv < 0:
        raise RuntimeError("oracle_source_iteration failed")
    out = (phi, its.value, res.value, bool(conv))
    return out + ((tr[: its.value],) if trace else ())


def tridiag_solve(lo, di, up, rhs):
    """x with A x = rhs, A tridiagonal (lo[i] = A[i][i-1], up[i] = A[i][i+1])."""
    a = [np.array(v, dtype=np.float64) for v in (lo, di, up, rhs)]
    x = np.zeros(len(a[1]))
    lib().oracle_tridiag_solve(len(x), *[_d(v) for v in a], _d(x))
    return x


def dsa_system(problem, g, eigen, phi, cur, sfr, sfl, qmg):
    """The 1D DSA tridiagonal rows (lo, di, up, rhs) for one sweep's moments."""
    b = _Bound(problem)
    n = b.p.N + 1
    ins = [np.ascontiguousarray(v, dtype=np.float64) for v in (phi, cur, sfr, sfl, qmg)]
    outs = [np.zeros(n) for _ in range(4)]
    lib().oracle_dsa_system(C.byref(b.s), g, int(eigen), *[_d(v) for v in ins],
                            *[_d(v) for v in outs])
    return tuple(outs)


def twdd3d_cell(q, sig, lx, ly, lz, amu, aeta, axi, inflow):
    i = np.ascontiguousarray(inflow, dtype=np.float64)
    o = np.zeros(3)
    psi = lib().oracle_twdd3d_cell(q, sig, lx, ly, lz, amu, aeta, axi, _d(i), _d(o))
    return psi, o


def dd3d_cell(q, sig, lx, ly, lz, amu, aeta, axi, inflow):
    i = np.ascontiguousarray(inflow, dtype=np.float64)
    o = np.zeros(3)
    psi = lib().oracle_dd3d_cell(q, sig, lx, ly, lz, amu, aeta, axi, _d(i), _d(o))
    return psi, o


def twdd2d_cell(q, sig, lx, ly, amu, aeta, axi, inflow):
    i = np.ascontiguousarray(inflow, dtype=np.float64)
    o = np.zeros(2)
    psi = lib().oracle_twdd2d_cell(q, sig, lx, ly, amu, aeta, axi, _d(i), _d(o))
    return psi, o


def dd2d_cell(q, sig, lx, ly, amu, aeta, inflow):
    i = np.ascontiguousarray(inflow, dtype=np.float64)
    o = np.zeros(2)
    psi = lib().oracle_dd2d_cell(q, sig, lx, ly, amu, aeta, _d(i), _d(o))
    return psi, o


def dd1d_cell(q, sig, lx, mu, inflow):
    o = np.zeros(1)
    psi = lib().oracle_dd1d_cell(q, sig, lx, mu, inflow, _d(o))
    return psi, float(o[0])


# --------------------------------------------------------------------------
# Reference harness (oracle/_ref, built only where /root/reference exists).
# --------------------------------------------------------------------------

def write_problem_file(problem, path, np_, na, fs_tol, fs_max_it, mg_tol, mg_max_it, nthreads=1):
    p = problem.normalized()
    with open(path, "wb") as f:
        f.write(struct.pack("<12i", p.dims, p.nx, p.ny, p.nz, p.groups, p.scheme, p.num_mats,
                            np_, na, fs_max_it, mg_max_it, nthreads))
        f.write(struct.pack("<2d", fs_tol, mg_tol))
        for a in (p.lx, p.ly, p.lz):
            f.write(a.astype("<f8").tobytes())
        f.write(p.mat.astype("<i4").tobytes())
        f.write(p.sigt.astype("<f8").tobytes())
        f.write(p.sigs.astype("<f8").tobytes())
        f.write(p.src.astype("<f8").tobytes())
        f.write(p.has_src.astype("u1").tobytes())


def run_reference(problem, np_, na, fs_tol, fs_max_it, mg_tol, mg_max_it, workdir, nthreads=1,
                  tag="case"):
    """Solve with the reference's own TransportSolver2D/3D; returns (ok, phi[G,N], log)."""
    if not REF_HARNESS.exists():
        raise RuntimeError("reference harness not built (make -C oracle ref)")
    workdir = Path(workdir)
    workdir.mkdir(parents=True, exist_ok=True)
    pin = workdir / f"{tag}.bin"
    pout = workdir / f"{tag}.out"
    write_problem_file(problem, pin, np_, na, fs_tol, fs_max_it, mg_tol, mg_max_it, nthreads)
    env = dict(os.environ, OMP_NUM_THREADS=str(nthreads))
    subprocess.run([str(REF_HARNESS), str(pin), str(pout)], check=True, env=env)
    raw = pout.read_bytes()
    ok = struct.unpack("<i", raw[:4])[0]
    p = problem.normalized()
    phi = np.frombuffer(raw[4:], dtype="<f8").reshape(p.groups, p.N).copy()
    log = Path(str(pout) + ".log").read_text()
    return bool(ok), phi, log
