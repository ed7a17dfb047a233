"""simplexprojects_amd -- batched simplex LP solver for AMD Instinct MI355X (gfx950).

Python view of the C ABI in include/blpg/blpg.h (libblpg.so, built in-tree by
simplexprojects_amd.build).  The reference's entry point is the C++ class
`Simplex` (BLPG-Solver/src/simplex.cuh:37-83); its C++ drop-in is
include/blpg/simplex.hpp.  This module is the harness the tests and the benchmark use:
it marshals LPs into the ABI and results back into numpy arrays.  There is no CPU
fallback: every solve runs the HIP kernels, and a missing library or device raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from pathlib import Path
from typing import Iterable, Optional, Sequence

import numpy as np

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "libblpg.so"

# ---- constants (blpg.h) ----
OK = 0
E_INVALID, E_HIP, E_NOMEM, E_NODEVICE, E_BUDGET = -1, -2, -3, -4, -5
UNDECIDED, INFEASIBLE, OPTIMAL, UNBOUNDED, ITER_LIMIT, NUMERIC = 1, 4, 5, 6, 100, 101
FP64, FP32 = 0, 1
KERNEL_AUTO, KERNEL_STREAM, KERNEL_LDS, KERNEL_TILE = 0, 1, 2, 3
_KERNELS = {"auto": 0, "stream": 1, "lds": 2, "tile": 3}
STATUS_NAMES = {1: "undecided", 4: "infeasible", 5: "optimal", 6: "unbounded",
                100: "iter_limit", 101: "numeric"}


class BlpgError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"blpg error {code}: {msg}")
        self.code = code


class _LP(C.Structure):
    _fields_ = [("m", C.c_int32), ("n", C.c_int32), ("A", C.POINTER(C.c_double)),
                ("b", C.POINTER(C.c_double)), ("c", C.POINTER(C.c_double)),
                ("sense", C.POINTER(C.c_int8)), ("family", C.c_int32),
                ("warm_basis", C.POINTER(C.c_int32))]


class _Result(C.Structure):
    _fields_ = [("status", C.c_int32), ("pivots", C.c_int32), ("phase1_pivots", C.c_int32),
                ("fallback_used", C.c_int32), ("objective", C.c_double),
                ("kkt_primal", C.c_double), ("kkt_dual", C.c_double),
                ("x", C.POINTER(C.c_double)), ("basis", C.POINTER(C.c_int32))]


class _Options(C.Structure):
    _fields_ = [("precision", C.c_int32), ("max_iter", C.c_int32), ("bland_after", C.c_int32),
                ("verify", C.c_int32), ("eps_d", C.c_double), ("eps_p", C.c_double),
                ("tol_feas", C.c_double), ("kkt_tol", C.c_double), ("hbm_budget", C.c_int64),
                ("kernel", C.c_int32), ("device", C.c_int32), ("use_graph", C.c_int32),
                ("reserved", C.c_int32)]


class _Info(C.Structure):
    _fields_ = [("nbins", C.c_int32), ("kernel", C.c_int32 * 16), ("rcap", C.c_int32 * 16),
                ("ccap", C.c_int32 * 16), ("count", C.c_int64 * 16),
                ("kernel_ms", C.c_double * 16), ("verify_ms", C.c_double),
                ("fallback_ms", C.c_double), ("fallback_count", C.c_int64),
                ("total_pivots", C.c_int64), ("total_ms", C.c_double)]


# every symbol include/blpg/blpg.h declares (checked by the CPU test suite)
ABI_SYMBOLS = [
    "blpg_default_options", "blpg_create", "blpg_destroy", "blpg_last_error",
    "blpg_set_options", "blpg_solve_batch", "blpg_solve_shared_A", "blpg_batch_upload",
    "blpg_batch_solve", "blpg_batch_sync", "blpg_batch_download", "blpg_batch_free",
    "blpg_batch_info_get", "blpg_gen_config", "blpg_version",
]

_lib: Optional[C.CDLL] = None


def load_library(path: Optional[os.PathLike] = None) -> C.CDLL:
    """Load libblpg.so (raises if it has not been built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise FileNotFoundError(f"{p} not built; run `python -m simplexprojects_amd.build`")
    lib = C.CDLL(str(p))
    P = C.POINTER
    lib.blpg_default_options.argtypes = [P(_Options)]
    lib.blpg_create.argtypes = [P(C.c_void_p), P(_Options)]
    lib.blpg_destroy.argtypes = [C.c_void_p]
    lib.blpg_last_error.argtypes = [C.c_void_p]
    lib.blpg_last_error.restype = C.c_char_p
    lib.blpg_set_options.argtypes = [C.c_void_p, P(_Options)]
    lib.blpg_solve_batch.argtypes = [C.c_void_p, P(_LP), C.c_int64, P(_Result)]
    lib.blpg_solve_shared_A.argtypes = [C.c_void_p, C.c_int32, C.c_int32, P(C.c_double),
                                        P(C.c_double), P(C.c_int8), P(C.c_double), C.c_int64,
                                        P(_Result)]
    lib.blpg_batch_upload.argtypes = [C.c_void_p, P(_LP), C.c_int64, P(C.c_void_p)]
    lib.blpg_batch_solve.argtypes = [C.c_void_p, C.c_void_p]
    lib.blpg_batch_sync.argtypes = [C.c_void_p]
    lib.blpg_batch_download.argtypes = [C.c_void_p, C.c_void_p, P(_Result)]
    lib.blpg_batch_free.argtypes = [C.c_void_p, C.c_void_p]
    lib.blpg_batch_info_get.argtypes = [C.c_void_p, C.c_void_p, P(_Info)]
    lib.blpg_gen_config.argtypes = [C.c_int32, C.c_uint64, C.c_int32, C.c_int64, C.c_int64,
                                    C.c_int64, C.c_int32, C.c_int32, P(C.c_double),
                                    P(C.c_double), P(C.c_double), P(C.c_int8), P(C.c_int32)]
    lib.blpg_version.restype = C.c_char_p
    if path is None:
        _lib = lib
    return lib


# ---------------------------------------------------------------------------------
@dataclass
class LP:
    """maximise c^T x  s.t.  A x (<=,=,>=) b, x >= 0   (sense: 0 '<=', 1 '=', 2 '>=')."""
    A: np.ndarray
    b: np.ndarray
    c: np.ndarray
    sense: Optional[np.ndarray] = None
    family: int = -1
    flags: int = 0
    warm_basis: Optional[np.ndarray] = None

    def __post_init__(self):
        self.A = np.ascontiguousarray(self.A, dtype=np.float64)
        self.b = np.ascontiguousarray(self.b, dtype=np.float64).reshape(-1)
        self.c = np.ascontiguousarray(self.c, dtype=np.float64).reshape(-1)
        if self.A.ndim != 2 or self.A.shape != (self.b.size, self.c.size):
            raise ValueError(f"inconsistent shapes A{self.A.shape} b{self.b.shape} c{self.c.shape}")
        if self.sense is not None:
            self.sense = np.ascontiguousarray(self.sense, dtype=np.int8).reshape(-1)
            if self.sense.size != self.b.size:
                raise ValueError("sense must have m entries")
        if self.warm_basis is not None:
            self.warm_basis = np.ascontiguousarray(self.warm_basis, dtype=np.int32).reshape(-1)
            if self.warm_basis.size != self.b.size:
                raise ValueError("warm_basis must have m entries")

    @property
    def m(self) -> int:
        return self.A.shape[0]

    @property
    def n(self) -> int:
        return self.A.shape[1]


def _dp(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


def _lp_struct(lp: LP) -> _LP:
    s = _LP()
    s.m, s.n = lp.m, lp.n
    s.A, s.b, s.c = _dp(lp.A), _dp(lp.b), _dp(lp.c)
    s.sense = lp.sense.ctypes.data_as(C.POINTER(C.c_int8)) if lp.sense is not None else None
    s.family = lp.family
    s.warm_basis = (lp.warm_basis.ctypes.data_as(C.POINTER(C.c_int32))
                    if lp.warm_basis is not None else None)
    return s


def gen_config(config: int, seed: int, m: int, n: int, count: int, shape_idx: int = 0,
               first: int = 0, batch_count: Optional[int] = None) -> list[LP]:
    """Seeded synthetic LPs (host generator in libblpg; see csrc/gen.cpp)."""
    lib = load_library()
    bc = batch_count if batch_count is not None else first + count
    A = np.empty((count, m, n)); b = np.empty((count, m)); c = np.empty((count, n))
    s = np.zeros((count, m), dtype=np.int8); f = np.zeros(count, dtype=np.int32)
    rc = lib.blpg_gen_config(config, seed, shape_idx, first, count, bc, m, n, _dp(A), _dp(b),
                             _dp(c), s.ctypes.data_as(C.POINTER(C.c_int8)),
                             f.ctypes.data_as(C.POINTER(C.c_int32)))
    if rc != 0:
        raise BlpgError(rc, "blpg_gen_config failed")
    use_sense = config == 3
    return [LP(A[k], b[k], c[k], s[k] if use_sense else None,
               family=shape_idx if config == 4 else -1, flags=int(f[k])) for k in range(count)]


@dataclass
class Results:
    status: np.ndarray
    pivots: np.ndarray
    phase1_pivots: np.ndarray
    fallback: np.ndarray
    objective: np.ndarray
    kkt_primal: np.ndarray
    kkt_dual: np.ndarray
    x: list = field(default_factory=list)
    basis: list = field(default_factory=list)

    def __len__(self):
        return len(self.status)


def _alloc_results(lps: Sequence[LP]):
    k = len(lps)
    res = (_Result * k)()
    xs = [np.zeros(lp.n) for lp in lps]
    bs = [np.zeros(lp.m, dtype=np.int32) for lp in lps]
    for i in range(k):
        res[i].x = _dp(xs[i])
        res[i].basis = bs[i].ctypes.data_as(C.POINTER(C.c_int32))
    return res, xs, bs


def _collect(res, xs, bs) -> Results:
    k = len(xs)
    # view the result records through a pointer-free mirror of blpg_result
    dt = np.dtype([("status", np.int32), ("pivots", np.int32), ("phase1_pivots", np.int32),
                   ("fallback_used", np.int32), ("objective", np.float64),
                   ("kkt_primal", np.float64), ("kkt_dual", np.float64), ("x", np.uintp),
                   ("basis", np.uintp)])
    assert dt.itemsize == C.sizeof(_Result)
    arr = np.frombuffer((C.c_char * (C.sizeof(_Result) * k)).from_address(C.addressof(res)),
                        dtype=dt, count=k) if k else np.zeros(0, dtype=dt)
    return Results(arr["status"].copy(), arr["pivots"].copy(), arr["phase1_pivots"].copy(),
                   arr["fallback_used"].copy(), arr["objective"].copy(),
                   arr["kkt_primal"].copy(), arr["kkt_dual"].copy(), xs, bs)


class Batch:
    """A device-resident batch (inputs uploaded once; solve() may be called many times)."""

    def __init__(self, solver: "Solver", lps: Sequence[LP]):
        self.solver = solver
        self.lps = list(lps)
        arr = (_LP * len(self.lps))(*[_lp_struct(lp) for lp in self.lps])
        h = C.c_void_p()
        solver._check(solver.lib.blpg_batch_upload(solver.ctx, arr, len(self.lps), C.byref(h)))
        self.h = h

    def solve(self) -> None:
        self.solver._check(self.solver.lib.blpg_batch_solve(self.solver.ctx, self.h))

    def sync(self) -> None:
        self.solver._check(self.solver.lib.blpg_batch_sync(self.solver.ctx))

    def download(self) -> Results:
        res, xs, bs = _alloc_results(self.lps)
        self.solver._check(self.solver.lib.blpg_batch_download(self.solver.ctx, self.h, res))
        return _collect(res, xs, bs)

    def info(self) -> dict:
        inf = _Info()
        self.solver._check(self.solver.lib.blpg_batch_info_get(self.solver.ctx, self.h,
                                                               C.byref(inf)))
        nb = inf.nbins
        return {"bins": [{"kernel": inf.kernel[i], "rcap": inf.rcap[i], "ccap": inf.ccap[i],
                          "count": inf.count[i], "kernel_ms": inf.kernel_ms[i]}
                         for i in range(nb)],
                "verify_ms": inf.verify_ms, "fallback_ms": inf.fallback_ms,
                "fallback_count": inf.fallback_count, "total_pivots": inf.total_pivots,
                "total_ms": inf.total_ms}

    def free(self) -> None:
        if self.h:
            self.solver.lib.blpg_batch_free(self.solver.ctx, self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Solver:
    """One blpg context (device, stream, workspaces)."""

    def __init__(self, precision: str = "fp64", kernel: str = "auto", verify: bool = True,
                 max_iter: int = 0, bland_after: int = 0, eps_d: float = 0.0,
                 eps_p: float = 0.0, tol_feas: float = 0.0, kkt_tol: float = 0.0,
                 hbm_budget: int = 0, device: int = 0, use_graph: bool = True):
        self.lib = load_library()
        o = _Options()
        self.lib.blpg_default_options(C.byref(o))
        o.precision = {"fp64": FP64, "fp32": FP32}[precision]
        o.kernel = _KERNELS[kernel]
        o.verify = int(verify)
        o.max_iter, o.bland_after = max_iter, bland_after
        o.eps_d, o.eps_p, o.tol_feas, o.kkt_tol = eps_d, eps_p, tol_feas, kkt_tol
        o.hbm_budget, o.device, o.use_graph = hbm_budget, device, int(use_graph)
        self.opts = o
        h = C.c_void_p()
        rc = self.lib.blpg_create(C.byref(h), C.byref(o))
        if rc != 0:
            raise BlpgError(rc, "blpg_create failed (no HIP device?)")
        self.ctx = h

    def _check(self, rc: int) -> None:
        if rc != 0:
            raise BlpgError(rc, self.lib.blpg_last_error(self.ctx).decode())

    def set(self, **kw) -> None:
        o = self.opts
        for k, v in kw.items():
            if k == "precision":
                v = {"fp64": FP64, "fp32": FP32}[v] if isinstance(v, str) else v
            if k == "kernel" and isinstance(v, str):
                v = _KERNELS[v]
            setattr(o, k, v)
        self._check(self.lib.blpg_set_options(self.ctx, C.byref(o)))

    def solve(self, lps: Sequence[LP]) -> Results:
        lps = list(lps)
        arr = (_LP * len(lps))(*[_lp_struct(lp) for lp in lps])
        res, xs, bs = _alloc_results(lps)
        self._check(self.lib.blpg_solve_batch(self.ctx, arr, len(lps), res))
        return _collect(res, xs, bs)

    def solve_shared_A(self, A, b, Cmat, sense=None) -> Results:
        A = np.ascontiguousarray(A, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        Cmat = np.ascontiguousarray(Cmat, dtype=np.float64)
        m, n = A.shape
        k = Cmat.shape[0]
        s = None if sense is None else np.ascontiguousarray(sense, dtype=np.int8)
        fake = [LP(A, b, Cmat[i], s) for i in range(k)]
        res, xs, bs = _alloc_results(fake)
        self._check(self.lib.blpg_solve_shared_A(
            self.ctx, m, n, _dp(A), _dp(b),
            None if s is None else s.ctypes.data_as(C.POINTER(C.c_int8)), _dp(Cmat), k, res))
        return _collect(res, xs, bs)

    def upload(self, lps: Sequence[LP]) -> Batch:
        return Batch(self, lps)

    def close(self) -> None:
        if getattr(self, "ctx", None):
            self.lib.blpg_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def default_max_iter(m: int, n: int) -> int:
    """The device's default pivot cap (blpg_options.max_iter == 0)."""
    return 50 * (m + n) + 1000


def default_tolerances(precision: str) -> dict:
    """Per-precision defaults applied by the library when an option is 0."""
    if precision == "fp32":
        return {"eps_d": 1e-6, "eps_p": 1e-6, "tol_feas": 1e-4}
    return {"eps_d": 1e-9, "eps_p": 1e-9, "tol_feas": 1e-7}
