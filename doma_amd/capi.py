"""ctypes binding of the doma_amd C ABI (include/doma_c.h).

The native library ``libdoma_amd.so`` is built in tree by ``doma_amd.build``.
Importing this module never falls back to a CPU path: if the library is missing
or cannot be loaded, :func:`lib` raises.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libdoma_amd.so"

DOMA_OK = 0
DOMA_ERR_INVALID = 1
DOMA_ERR_DEVICE = 2
DOMA_ERR_STATE = 3

SCHEME_DD = 0
SCHEME_TW_DD = 1
FP64 = 0
FP32_SWEEP = 1
MAX_GROUPS = 64
MAX_DELAYED = 8
STEP_NONE, STEP_BOTH, STEP_INSERT, STEP_REMOVE = -1, 0, 1, 2

SCATTER_DOWN1 = 0
SCATTER_FULL = 1
SCATTER_INGROUP = 2
STOP_RESIDUAL = 0
NET_CIN = 4  # input channels of the correction network
NET_NOALLOC_MAX_VOXELS = 32768
STOP_ERROR_ESTIMATE = 1
SOURCE_CENTRAL = 0
SOURCE_BLOCKS = 1

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_u8p = C.POINTER(C.c_uint8)


class ProblemDesc(C.Structure):
    _fields_ = [
        ("dims", C.c_int32), ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("lx", _dp), ("ly", _dp), ("lz", _dp),
        ("groups", C.c_int32), ("scheme", C.c_int32), ("precision", C.c_int32),
        ("dirs", C.c_int32),
        ("mu", _dp), ("eta", _dp), ("xi", _dp), ("w", _dp),
    ]


class Tuning(C.Structure):
    _fields_ = [("brick", C.c_int32), ("threads", C.c_int32), ("exact_arith", C.c_int32),
                ("schedule", C.c_int32), ("octants", C.c_int32), ("dirs_per_lane", C.c_int32),
                ("shape", C.c_int32)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class Stats(C.Structure):
    _fields_ = [
        ("success", C.c_int32), ("mg_iterations", C.c_int32), ("mg_residual", C.c_double),
        ("sweeps", C.c_int32), ("si_iterations", C.c_int32 * MAX_GROUPS),
        ("si_residual", C.c_double * MAX_GROUPS), ("elapsed_ms", C.c_double),
        ("sweep_ms", C.c_double), ("corrections_accepted", C.c_int32),
        ("corrections_rejected", C.c_int32), ("correction_ms", C.c_double),
        ("k_eff", C.c_double), ("k_residual", C.c_double),
    ]


class MediaSpec(C.Structure):
    _fields_ = [
        ("n", C.c_int32), ("block", C.c_int32), ("groups", C.c_int32), ("sigt_log", C.c_int32),
        ("sigt_lo", C.c_double), ("sigt_hi", C.c_double), ("c_lo", C.c_double),
        ("c_hi", C.c_double), ("scatter_rule", C.c_int32), ("source_rule", C.c_int32),
        ("src_lo", C.c_int32), ("src_hi", C.c_int32), ("source_fraction", C.c_double),
        ("cell_width", C.c_double), ("seed", C.c_uint64),
    ]


class DeckInfo(C.Structure):
    _fields_ = [(n, C.c_int32) for n in (
        "dims", "nx", "ny", "nz", "groups", "scheme", "mode", "np", "na", "num_materials", "dsa",
        "refinement", "fs_max_it", "mg_max_it")] + [
        ("fs_tol", C.c_double), ("mg_tol", C.c_double), ("total_volume", C.c_double)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class BatchInfo(C.Structure):
    _fields_ = [("rounds", C.c_int32), ("units", C.c_int64), ("elapsed_ms", C.c_double),
                ("sweep_ms", C.c_double), ("samples", C.c_int64)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


_LIB = None

# (name, restype, argtypes) of every exported entry point.
SIGNATURES = [
    ("doma_last_error", C.c_char_p, []),
    ("doma_version", C.c_char_p, []),
    ("doma_create", C.c_int, [C.POINTER(ProblemDesc), C.POINTER(C.c_void_p)]),
    ("doma_destroy", C.c_int, [C.c_void_p]),
    ("doma_set_materials", C.c_int, [C.c_void_p, _ip, C.c_int32, _dp, _dp, _dp, _u8p]),
    ("doma_set_fission", C.c_int, [C.c_void_p, C.c_int32, _dp, _dp, _dp, _u8p, _u8p]),
    ("doma_set_dsa", C.c_int, [C.c_void_p, C.c_int32]),
    ("doma_set_si_mode", C.c_int, [C.c_void_p, C.c_int32]),
    ("doma_set_tuning", C.c_int, [C.c_void_p, C.POINTER(Tuning)]),
    ("doma_get_tuning", C.c_int, [C.c_void_p, C.POINTER(Tuning)]),
    ("doma_sweep", C.c_int, [C.c_void_p, C.c_int32, _dp, _dp]),
    ("doma_sweep_device", C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]),
    ("doma_source_iteration", C.c_int,
     [C.c_void_p, C.c_int32, _dp, C.c_double, C.c_int32, _dp, _ip, _dp, _ip]),
    ("doma_solve_fixed_source", C.c_int,
     [C.c_void_p, C.c_double, C.c_int32, C.c_double, C.c_int32, _dp, C.POINTER(Stats)]),
    ("doma_solve_eigenvalue", C.c_int,
     [C.c_void_p, C.c_double, C.c_int32, C.c_double, C.c_int32, C.c_double, _dp, _dp,
      C.POINTER(Stats)]),
    ("doma_set_kinetics", C.c_int, [C.c_void_p, C.c_int32, _dp, _ip, _dp, _dp, _dp]),
    ("doma_set_step_sources", C.c_int, [C.c_void_p, C.c_int32, _ip, _dp, _dp, _dp]),
    ("doma_transient_begin", C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_int32]),
    ("doma_transient_step", C.c_int, [C.c_void_p, C.c_double, C.c_int32, C.c_double, C.c_int32,
                                      _dp, C.POINTER(Stats)]),
    ("doma_get_dnps", C.c_int, [C.c_void_p, _dp, _ip]),
    ("doma_get_flux", C.c_int, [C.c_void_p, _dp]),
    ("doma_time_sweeps", C.c_int, [C.c_void_p, C.c_int32, C.c_int32, _dp, _ip]),
    ("doma_profile_sweeps", C.c_int, [C.c_void_p, C.c_int32, C.c_int32, _dp, _dp, _ip]),
    ("doma_units_per_sweep", C.c_int64, [C.c_void_p]),
    ("doma_autotune", C.c_int, [C.c_void_p, C.c_int32, C.POINTER(Tuning), _dp]),
    ("doma_gc_quadrature", C.c_int, [C.c_int32, C.c_int32, C.c_int32, _dp, _dp, _dp, _dp]),
    ("doma_gl_quadrature", C.c_int, [C.c_int32, _dp, _dp]),
    ("doma_media_num_materials", C.c_int32, [C.POINTER(MediaSpec)]),
    ("doma_media_generate", C.c_int, [C.POINTER(MediaSpec), _ip, _dp, _dp, _dp, _u8p]),
    ("doma_net_param_count", C.c_int64, []),
    ("doma_set_learned_correction", C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.c_int64,
                                              C.c_int32]),
    ("doma_set_warm_start", C.c_int, [C.c_void_p, C.c_int32]),
    ("doma_set_stopping_rule", C.c_int, [C.c_void_p, C.c_int32]),
    ("doma_net_forward_noalloc", C.c_int, [C.c_void_p, C.c_int64, C.c_int32, C.c_int32,
                                           C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("doma_net_forward", C.c_int, [C.POINTER(C.c_float), C.c_int64, C.c_int32, C.c_int32,
                                   C.c_int32, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    ("doma_net_trainer_create", C.c_int, [C.c_int32, C.c_int32, C.POINTER(C.c_float), C.c_int64,
                                          C.POINTER(C.c_void_p)]),
    ("doma_net_trainer_destroy", C.c_int, [C.c_void_p]),
    ("doma_net_trainer_step", C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                        C.c_float, C.POINTER(C.c_float)]),
    ("doma_net_trainer_gradient", C.c_int, [C.c_void_p, C.POINTER(C.c_float),
                                            C.POINTER(C.c_float), C.POINTER(C.c_float),
                                            C.POINTER(C.c_float)]),
    ("doma_net_trainer_get_params", C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    ("doma_batch_create", C.c_int, [C.c_int32, C.POINTER(ProblemDesc), C.POINTER(C.c_void_p)]),
    ("doma_batch_destroy", C.c_int, [C.c_void_p]),
    ("doma_batch_set_materials", C.c_int, [C.c_void_p, C.c_int32, _ip, C.c_int32, _dp, _dp, _dp,
                                           _u8p]),
    ("doma_batch_set_tuning", C.c_int, [C.c_void_p, C.POINTER(Tuning)]),
    ("doma_batch_get_tuning", C.c_int, [C.c_void_p, C.POINTER(Tuning)]),
    ("doma_batch_autotune", C.c_int, [C.c_void_p, C.POINTER(Tuning), _dp]),
    ("doma_batch_time_sweeps", C.c_int, [C.c_void_p, C.c_int32, _dp, C.POINTER(C.c_int64)]),
    ("doma_batch_solve", C.c_int, [C.c_void_p, C.c_double, C.c_int32, C.c_double, C.c_int32,
                                   C.c_int32, C.c_int32, C.POINTER(Stats), C.POINTER(BatchInfo)]),
    ("doma_batch_get_flux", C.c_int, [C.c_void_p, C.c_int32, _dp]),
    ("doma_batch_sample_count", C.c_int, [C.c_void_p, C.c_int32, C.c_int32, _ip, _ip]),
    ("doma_batch_get_sample", C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32,
                                        C.POINTER(C.c_float), C.POINTER(C.c_float),
                                        C.POINTER(C.c_float)]),
    ("doma_deck_read", C.c_int, [C.c_char_p, C.POINTER(DeckInfo)]),
    ("doma_run_deck", C.c_int,
     [C.c_char_p, C.c_char_p, C.c_int32, C.POINTER(Tuning), C.c_int32, _dp, C.c_int64,
      C.POINTER(Stats), _ip]),
]


def lib() -> C.CDLL:
    """Load libdoma_amd.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m doma_amd.build`")
        h = C.CDLL(str(LIB_PATH))
        for name, res, args in SIGNATURES:
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = h
    return _LIB


class DomaError(RuntimeError):
    def __init__(self, code: int, where: str):
        msg = lib().doma_last_error().decode(errors="replace")
        super().__init__(f"{where} failed ({code}): {msg}")
        self.code = code


def _check(code: int, where: str):
    if code != DOMA_OK:
        raise DomaError(code, where)


def _d(a):
    return a.ctypes.data_as(_dp)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def gc_quadrature(n_c: int, n_l: int, dims: int):
    """Directions/weights of GCAngularQuadrature(n_c, n_l, dims) as [noct, per] arrays."""
    noct = 8 if dims == 3 else 4
    per = (n_c * n_l) // 4
    arrs = [np.zeros(noct * max(per, 1)) for _ in range(4)]
    r = lib().doma_gc_quadrature(n_c, n_l, dims, *[_d(a) for a in arrs])
    if r < 0:
        raise ValueError(f"GC quadrature ({n_c}, {n_l}, {dims}) has a zero direction component")
    return tuple(a.reshape(noct, per) for a in arrs)


def gl_quadrature(n: int):
    """GLAngularQuadrature(n) nodes (descending) and weights, as [2, n/2] half ranges."""
    mu, w = np.zeros(n), np.zeros(n)
    r = lib().doma_gl_quadrature(n, _d(mu), _d(w))
    if r < 0:
        raise ValueError("GL quadrature needs an even order >= 2")
    return mu.reshape(2, n // 2), w.reshape(2, n // 2)


@dataclass
class Fission:
    """Fission data of the materials (MaterialProps _g_chi_p, _g_prod, _g_n_beta):
    chi [M, G] prompt spectrum, prod [M, G] nu*sigma_f, prompt [M, G] prompt
    fraction 1 - sum_d beta (1 without delayed groups), has_chi / has_prod [M]."""
    chi: np.ndarray
    prod: np.ndarray
    prompt: np.ndarray
    has_chi: np.ndarray
    has_prod: np.ndarray

    def normalized(self, G: int) -> "Fission":
        return Fission(_f64(self.chi).reshape(-1, G), _f64(self.prod).reshape(-1, G),
                       _f64(self.prompt).reshape(-1, G),
                       np.ascontiguousarray(self.has_chi, dtype=np.uint8).reshape(-1),
                       np.ascontiguousarray(self.has_prod, dtype=np.uint8).reshape(-1))


@dataclass
class Kinetics:
    """Transient data: inv_v [M, G]; nd [M] delayed groups; chi_d, beta [M, G, 8];
    lambda_ [M, 8] (slots past a material's count are ignored)."""
    inv_v: np.ndarray
    nd: np.ndarray
    chi_d: np.ndarray
    beta: np.ndarray
    lambda_: np.ndarray

    def normalized(self, G: int) -> "Kinetics":
        return Kinetics(_f64(self.inv_v).reshape(-1, G),
                        np.ascontiguousarray(self.nd, dtype=np.int32).reshape(-1),
                        _f64(self.chi_d).reshape(-1, G, MAX_DELAYED),
                        _f64(self.beta).reshape(-1, G, MAX_DELAYED),
                        _f64(self.lambda_).reshape(-1, MAX_DELAYED))


@dataclass
class StepSources:
    """Per material: type [M] (STEP_*), insert_at, remove_at [M], src [M, G]."""
    type: np.ndarray
    insert_at: np.ndarray
    remove_at: np.ndarray
    src: np.ndarray

    def normalized(self, G: int) -> "StepSources":
        return StepSources(np.ascontiguousarray(self.type, dtype=np.int32).reshape(-1),
                           _f64(self.insert_at).reshape(-1), _f64(self.remove_at).reshape(-1),
                           _f64(self.src).reshape(-1, G))


@dataclass
class Problem:
    """A problem in the C-ABI layout (numpy, host); `fission` is optional."""
    dims: int
    lx: np.ndarray
    ly: np.ndarray
    lz: np.ndarray
    groups: int
    scheme: int
    mat: np.ndarray          # int32 [nz, ny, nx] or flat
    sigt: np.ndarray         # [M, G]
    sigs: np.ndarray         # [M, G, G] dest-major
    src: np.ndarray          # [M, G]
    has_src: np.ndarray      # uint8 [M]
    mu: np.ndarray           # [noct, dirs]
    eta: np.ndarray
    xi: np.ndarray
    w: np.ndarray
    name: str = ""
    meta: dict = field(default_factory=dict)
    fission: Fission | None = None
    kinetics: Kinetics | None = None
    steps: StepSources | None = None

    @property
    def nx(self):
        return len(self.lx)

    @property
    def ny(self):
        return len(self.ly)

    @property
    def nz(self):
        return len(self.lz)

    @property
    def N(self):
        return self.nx * self.ny * self.nz

    @property
    def dirs(self):
        return self.mu.shape[1]

    @property
    def noct(self):
        return self.mu.shape[0]

    @property
    def num_mats(self):
        return self.sigt.shape[0]

    def normalized(self) -> "Problem":
        """Contiguous arrays with the exact dtypes the C ABI expects."""
        return Problem(
            self.dims, _f64(self.lx), _f64(self.ly), _f64(self.lz), int(self.groups),
            int(self.scheme), np.ascontiguousarray(self.mat, dtype=np.int32).reshape(-1),
            _f64(self.sigt).reshape(-1, self.groups),
            _f64(self.sigs).reshape(-1, self.groups, self.groups),
            _f64(self.src).reshape(-1, self.groups),
            np.ascontiguousarray(self.has_src, dtype=np.uint8).reshape(-1),
            _f64(self.mu), _f64(self.eta), _f64(self.xi), _f64(self.w), self.name, dict(self.meta),
            None if self.fission is None else self.fission.normalized(int(self.groups)),
            None if self.kinetics is None else self.kinetics.normalized(int(self.groups)),
            None if self.steps is None else self.steps.normalized(int(self.groups)))


class Context:
    """Owns a doma_ctx: device-resident problem, sweep and iteration drivers."""

    def __init__(self, problem: Problem, precision: int = FP64, tuning: dict | None = None):
        p = problem.normalized()
        self.problem = p
        self._keep = p  # arrays must outlive the create call
        d = ProblemDesc(p.dims, p.nx, p.ny, p.nz, _d(p.lx), _d(p.ly), _d(p.lz), p.groups,
                        p.scheme, precision, p.dirs, _d(p.mu), _d(p.eta), _d(p.xi), _d(p.w))
        h = C.c_void_p()
        _check(lib().doma_create(C.byref(d), C.byref(h)), "doma_create")
        self.h = h
        _check(lib().doma_set_materials(h, p.mat.ctypes.data_as(_ip), p.num_mats, _d(p.sigt),
                                        _d(p.sigs), _d(p.src), p.has_src.ctypes.data_as(_u8p)),
               "doma_set_materials")
        if p.fission is not None:
            f = p.fission
            _check(lib().doma_set_fission(h, p.num_mats, _d(f.chi), _d(f.prod), _d(f.prompt),
                                          f.has_chi.ctypes.data_as(_u8p),
                                          f.has_prod.ctypes.data_as(_u8p)), "doma_set_fission")
        if p.kinetics is not None:
            kn = p.kinetics
            _check(lib().doma_set_kinetics(h, p.num_mats, _d(kn.inv_v), kn.nd.ctypes.data_as(_ip),
                                           _d(kn.chi_d), _d(kn.beta), _d(kn.lambda_)),
                   "doma_set_kinetics")
        if p.steps is not None:
            st = p.steps
            _check(lib().doma_set_step_sources(h, p.num_mats, st.type.ctypes.data_as(_ip),
                                               _d(st.insert_at), _d(st.remove_at), _d(st.src)),
                   "doma_set_step_sources")
        if tuning:
            self.set_tuning(**tuning)

    def close(self):
        if getattr(self, "h", None):
            lib().doma_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_tuning(self, brick=0, threads=0, exact_arith=0, schedule=0, octants=0,
                   dirs_per_lane=0, shape=0):
        """Restrict the sweep kernel shape (0 = any); shape is the kernel id + 1."""
        t = Tuning(brick, threads, exact_arith, schedule, octants, dirs_per_lane, shape)
        _check(lib().doma_set_tuning(self.h, C.byref(t)), "doma_set_tuning")

    def get_tuning(self) -> dict:
        t = Tuning()
        _check(lib().doma_get_tuning(self.h, C.byref(t)), "doma_get_tuning")
        return t.as_dict()

    def sweep(self, g: int, q: np.ndarray) -> np.ndarray:
        q = _f64(q).reshape(-1)
        out = np.zeros(self.problem.N)
        _check(lib().doma_sweep(self.h, g, _d(q), _d(out)), "doma_sweep")
        return out

    def source_iteration(self, g: int, qmg: np.ndarray, tol: float, max_it: int):
        qmg = _f64(qmg).reshape(-1)
        phi = np.zeros(self.problem.N)
        its, res, conv = C.c_int32(), C.c_double(), C.c_int32()
        _check(lib().doma_source_iteration(self.h, g, _d(qmg), tol, max_it, _d(phi), C.byref(its),
                                           C.byref(res), C.byref(conv)), "doma_source_iteration")
        return phi, its.value, res.value, bool(conv.value)

    def solve_fixed_source(self, fs_tol: float, fs_max_it: int, mg_tol: float, mg_max_it: int,
                           want_flux: bool = True):
        phi = np.zeros(self.problem.groups * self.problem.N) if want_flux else None
        st = Stats()
        _check(lib().doma_solve_fixed_source(self.h, fs_tol, fs_max_it, mg_tol, mg_max_it,
                                             _d(phi) if want_flux else None, C.byref(st)),
               "doma_solve_fixed_source")
        stats = _stats_dict(st, self.problem.groups)
        if want_flux:
            phi = phi.reshape(self.problem.groups, self.problem.N)
        return phi, stats

    def solve_eigenvalue(self, fs_tol: float, fs_max_it: int, mg_tol: float, mg_max_it: int,
                         k_tol: float, want_flux: bool = True):
        """Power iteration: (k, phi[G, N] or None, stats)."""
        phi = np.zeros(self.problem.groups * self.problem.N) if want_flux else None
        st, k = Stats(), C.c_double()
        _check(lib().doma_solve_eigenvalue(self.h, fs_tol, fs_max_it, mg_tol, mg_max_it, k_tol,
                                           _d(phi) if want_flux else None, C.byref(k),
                                           C.byref(st)), "doma_solve_eigenvalue")
        stats = _stats_dict(st, self.problem.groups)
        if want_flux:
            phi = phi.reshape(self.problem.groups, self.problem.N)
        return k.value, phi, stats

    def transient_begin(self, t0: float, t1: float, steps: int):
        """Zero initial condition; dt = (t1 - t0) / steps; the step clock starts at 0."""
        _check(lib().doma_transient_begin(self.h, t0, t1, steps), "doma_transient_begin")

    def transient_step(self, fs_tol: float, fs_max_it: int, mg_tol: float, mg_max_it: int):
        """Solve the next step: (t of the step, stats)."""
        t, st = C.c_double(), Stats()
        _check(lib().doma_transient_step(self.h, fs_tol, fs_max_it, mg_tol, mg_max_it,
                                         C.byref(t), C.byref(st)), "doma_transient_step")
        return t.value, _stats_dict(st, self.problem.groups)

    def get_dnps(self) -> np.ndarray:
        """Precursors [nd_max, N]."""
        nd = C.c_int32()
        _check(lib().doma_get_dnps(self.h, None, C.byref(nd)), "doma_get_dnps")
        out = np.zeros((nd.value, self.problem.N))
        if nd.value:
            _check(lib().doma_get_dnps(self.h, _d(out), C.byref(nd)), "doma_get_dnps")
        return out

    def get_flux(self) -> np.ndarray:
        phi = np.zeros(self.problem.groups * self.problem.N)
        _check(lib().doma_get_flux(self.h, _d(phi)), "doma_get_flux")
        return phi.reshape(self.problem.groups, self.problem.N)

    def set_si_mode(self, device_loop: bool = True):
        """Stopping rule of source iteration on the device (default) or per sweep on the host."""
        _check(lib().doma_set_si_mode(self.h, int(device_loop)), "doma_set_si_mode")

    def set_dsa(self, enable: bool = True):
        """1D diffusion synthetic acceleration after every sweep."""
        _check(lib().doma_set_dsa(self.h, int(enable)), "doma_set_dsa")

    def set_warm_start(self, enable: bool = True):
        """Start each group's iteration from the previous multigroup iterate."""
        _check(lib().doma_set_warm_start(self.h, int(enable)), "doma_set_warm_start")

    def set_stopping_rule(self, rule: int = STOP_RESIDUAL):
        """STOP_RESIDUAL (the reference's) or STOP_ERROR_ESTIMATE (r rho / (1 - rho))."""
        _check(lib().doma_set_stopping_rule(self.h, int(rule)), "doma_set_stopping_rule")

    def set_learned_correction(self, params, every: int = 5):
        """Enable (every > 0) or disable (every = 0) the learned SI correction."""
        fp = C.POINTER(C.c_float)
        if every > 0:
            params = np.ascontiguousarray(params, dtype=np.float32)
            _check(lib().doma_set_learned_correction(self.h, params.ctypes.data_as(fp),
                                                     params.size, every),
                   "doma_set_learned_correction")
        else:
            _check(lib().doma_set_learned_correction(self.h, None, 0, 0),
                   "doma_set_learned_correction")

    def time_sweeps(self, g: int, n: int):
        ms, launches = C.c_double(), C.c_int32()
        _check(lib().doma_time_sweeps(self.h, g, n, C.byref(ms), C.byref(launches)),
               "doma_time_sweeps")
        return ms.value, launches.value

    def profile_sweeps(self, g: int, n: int):
        """(kernel ms per sweep, wall ms per sweep, launches per sweep) from HIP events
        recorded around every launch on the library's stream."""
        k, w, L = C.c_double(), C.c_double(), C.c_int32()
        _check(lib().doma_profile_sweeps(self.h, g, n, C.byref(k), C.byref(w), C.byref(L)),
               "doma_profile_sweeps")
        return k.value, w.value, L.value

    def units_per_sweep(self) -> int:
        return int(lib().doma_units_per_sweep(self.h))

    def autotune(self, g: int = 0):
        t, ms = Tuning(), C.c_double()
        _check(lib().doma_autotune(self.h, g, C.byref(t), C.byref(ms)), "doma_autotune")
        return t.as_dict(), ms.value


def _stats_dict(st: Stats, G: int) -> dict:
    return {"success": bool(st.success), "mg_iterations": st.mg_iterations,
            "mg_residual": st.mg_residual, "sweeps": st.sweeps,
            "si_iterations": list(st.si_iterations[:G]), "si_residual": list(st.si_residual[:G]),
            "elapsed_ms": st.elapsed_ms, "sweep_ms": st.sweep_ms,
            "corrections_accepted": st.corrections_accepted,
            "corrections_rejected": st.corrections_rejected, "correction_ms": st.correction_ms,
            "k_eff": st.k_eff, "k_residual": st.k_residual}


def net_param_count() -> int:
    return int(lib().doma_net_param_count())


def net_forward(params: np.ndarray, features: np.ndarray) -> np.ndarray:
    """Stand-alone GPU forward pass: features [D, H, W, 4] (FP32) -> [D, H, W]."""
    params = np.ascontiguousarray(params, dtype=np.float32)
    f = np.ascontiguousarray(features, dtype=np.float32)
    D, H, W, cin = f.shape
    if cin != NET_CIN:
        raise ValueError(f"features must have {NET_CIN} channels")
    out = np.zeros((D, H, W), dtype=np.float32)
    fp = C.POINTER(C.c_float)
    _check(lib().doma_net_forward(params.ctypes.data_as(fp), params.size, D, H, W,
                                  f.ctypes.data_as(fp), out.ctypes.data_as(fp)),
           "doma_net_forward")
    return out


class Batch:
    """Owns a doma_batch: many independent 3D problems (same scheme, precision and
    quadrature) swept together, one persistent launch per source-iteration round."""

    def __init__(self, problems, precision: int = FP64):
        ps = [p.normalized() for p in problems]
        self.problems = ps
        descs = (ProblemDesc * len(ps))()
        for i, p in enumerate(ps):
            descs[i] = ProblemDesc(p.dims, p.nx, p.ny, p.nz, _d(p.lx), _d(p.ly), _d(p.lz),
                                   p.groups, p.scheme, precision, p.dirs, _d(p.mu), _d(p.eta),
                                   _d(p.xi), _d(p.w))
        h = C.c_void_p()
        _check(lib().doma_batch_create(len(ps), descs, C.byref(h)), "doma_batch_create")
        self.h = h
        for i, p in enumerate(ps):
            _check(lib().doma_batch_set_materials(h, i, p.mat.ctypes.data_as(_ip), p.num_mats,
                                                  _d(p.sigt), _d(p.sigs), _d(p.src),
                                                  p.has_src.ctypes.data_as(_u8p)),
                   "doma_batch_set_materials")

    def close(self):
        if getattr(self, "h", None):
            lib().doma_batch_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_tuning(self, **kw):
        t = Tuning(**kw)
        _check(lib().doma_batch_set_tuning(self.h, C.byref(t)), "doma_batch_set_tuning")

    def get_tuning(self) -> dict:
        t = Tuning()
        _check(lib().doma_batch_get_tuning(self.h, C.byref(t)), "doma_batch_get_tuning")
        return t.as_dict()

    def autotune(self):
        t, ms = Tuning(), C.c_double()
        _check(lib().doma_batch_autotune(self.h, C.byref(t), C.byref(ms)), "doma_batch_autotune")
        return t.as_dict(), ms.value

    def time_sweeps(self, n: int):
        ms, units = C.c_double(), C.c_int64()
        _check(lib().doma_batch_time_sweeps(self.h, n, C.byref(ms), C.byref(units)),
               "doma_batch_time_sweeps")
        return ms.value, units.value

    def solve(self, fs_tol: float, fs_max_it: int, mg_tol: float, mg_max_it: int,
              sample_every: int = 0, sample_cap: int = 0):
        st = (Stats * len(self.problems))()
        info = BatchInfo()
        _check(lib().doma_batch_solve(self.h, fs_tol, fs_max_it, mg_tol, mg_max_it, sample_every,
                                      sample_cap, st, C.byref(info)), "doma_batch_solve")
        return [_stats_dict(st[i], p.groups) for i, p in enumerate(self.problems)], info.as_dict()

    def get_flux(self, p: int) -> np.ndarray:
        pr = self.problems[p]
        phi = np.zeros(pr.groups * pr.N)
        _check(lib().doma_batch_get_flux(self.h, p, _d(phi)), "doma_batch_get_flux")
        return phi.reshape(pr.groups, pr.N)

    def sample_iterations(self, p: int, g: int):
        n = C.c_int32()
        _check(lib().doma_batch_sample_count(self.h, p, g, C.byref(n), None),
               "doma_batch_sample_count")
        its = np.zeros(max(n.value, 1), dtype=np.int32)
        _check(lib().doma_batch_sample_count(self.h, p, g, C.byref(n), its.ctypes.data_as(_ip)),
               "doma_batch_sample_count")
        return its[: n.value].tolist()

    def get_sample(self, p: int, g: int, k: int, with_prev: bool = False):
        """(d_k, target) or, with_prev, (d_k, d_(k-1), target) of snapshot k."""
        N = self.problems[p].N
        r, rp, t = (np.zeros(N, dtype=np.float32) for _ in range(3))
        fp = C.POINTER(C.c_float)
        _check(lib().doma_batch_get_sample(self.h, p, g, k, r.ctypes.data_as(fp),
                                           rp.ctypes.data_as(fp), t.ctypes.data_as(fp)),
               "doma_batch_get_sample")
        return (r, rp, t) if with_prev else (r, t)


class NetTrainer:
    """Hand-written GPU training of the correction network (native kernels)."""

    def __init__(self, params, crop: int = 32, batch: int = 8):
        self.crop, self.batch = crop, batch
        params = np.ascontiguousarray(params, dtype=np.float32)
        h = C.c_void_p()
        fp = C.POINTER(C.c_float)
        _check(lib().doma_net_trainer_create(crop, batch, params.ctypes.data_as(fp), params.size,
                                             C.byref(h)), "doma_net_trainer_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().doma_net_trainer_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _io(self, features, targets):
        f = np.ascontiguousarray(features, dtype=np.float32)
        t = np.ascontiguousarray(targets, dtype=np.float32)
        n = self.batch * self.crop ** 3
        if f.size != NET_CIN * n or t.size != n:
            raise ValueError(f"features must be [batch, crop, crop, crop, {NET_CIN}], "
                             "targets [batch, crop^3]")
        return f, t

    def step(self, features, targets, lr: float) -> float:
        f, t = self._io(features, targets)
        fp = C.POINTER(C.c_float)
        loss = C.c_float()
        _check(lib().doma_net_trainer_step(self.h, f.ctypes.data_as(fp), t.ctypes.data_as(fp), lr,
                                           C.byref(loss)), "doma_net_trainer_step")
        return loss.value

    def gradient(self, features, targets):
        f, t = self._io(features, targets)
        fp = C.POINTER(C.c_float)
        loss = C.c_float()
        g = np.zeros(net_param_count(), dtype=np.float32)
        _check(lib().doma_net_trainer_gradient(self.h, f.ctypes.data_as(fp), t.ctypes.data_as(fp),
                                               C.byref(loss), g.ctypes.data_as(fp)),
               "doma_net_trainer_gradient")
        return loss.value, g

    def params(self) -> np.ndarray:
        p = np.zeros(net_param_count(), dtype=np.float32)
        _check(lib().doma_net_trainer_get_params(self.h, p.ctypes.data_as(C.POINTER(C.c_float))),
               "doma_net_trainer_get_params")
        return p


def media_generate(spec: MediaSpec):
    """Seeded block media: (mat[n^3], sigt[M,G], sigs[M,G,G], src[M,G], has_src[M])."""
    M = lib().doma_media_num_materials(C.byref(spec))
    if M < 0:
        raise ValueError("invalid media spec")
    n, G = spec.n, spec.groups
    mat = np.zeros(n * n * n, dtype=np.int32)
    sigt, sigs, src = np.zeros((M, G)), np.zeros((M, G, G)), np.zeros((M, G))
    has = np.zeros(M, dtype=np.uint8)
    _check(lib().doma_media_generate(C.byref(spec), mat.ctypes.data_as(_ip), _d(sigt), _d(sigs),
                                     _d(src), has.ctypes.data_as(_u8p)), "doma_media_generate")
    return mat, sigt, sigs, src, has


def deck_read(path) -> dict:
    """Parse a <doma_inp> deck and expand its mesh (host only, no GPU needed)."""
    info = DeckInfo()
    _check(lib().doma_deck_read(str(path).encode(), C.byref(info)), "doma_deck_read")
    return info.as_dict()


def run_deck(path, output_base=None, precision: int = FP64, tuning: dict | None = None,
             autotune: bool = False):
    """Solve a fixed-source or eigenvalue deck on the GPU like the reference's
    main(): returns (converged, phi[G, N], stats dict; k in stats["k_eff"]).
    With output_base the reference's text files <base>_*.txt are written."""
    info = deck_read(path)
    N = info["nx"] * info["ny"] * info["nz"]
    phi = np.zeros((info["groups"], N))
    st, conv = Stats(), C.c_int32()
    t = None
    if tuning:
        t = Tuning(**tuning)
    _check(lib().doma_run_deck(str(path).encode(),
                               None if output_base is None else str(output_base).encode(),
                               precision, None if t is None else C.byref(t), int(autotune),
                               _d(phi), phi.size, C.byref(st), C.byref(conv)), "doma_run_deck")
    G = info["groups"]
    stats = {"success": st.success, "mg_iterations": st.mg_iterations,
             "mg_residual": st.mg_residual, "sweeps": st.sweeps,
             "si_iterations": list(st.si_iterations[:G]), "si_residual": list(st.si_residual[:G]),
             "elapsed_ms": st.elapsed_ms, "sweep_ms": st.sweep_ms, "k_eff": st.k_eff,
             "k_residual": st.k_residual}
    return bool(conv.value), phi, stats
