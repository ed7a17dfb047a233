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
KERNEL_AUTO, KERNEL_STREAM, KERNEL_LDS, KERNEL_TILE, KERNEL_WAVE = 0, 1, 2, 3, 4
_KERNELS = {"auto": 0, "stream": 1, "lds": 2, "tile": 3, "wave": 4}
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
                ("sibling_warm", C.c_int32)]


class _Info(C.Structure):
    _fields_ = [("nbins", C.c_int32), ("kernel", C.c_int32 * 16), ("rcap", C.c_int32 * 16),
                ("ccap", C.c_int32 * 16), ("count", C.c_int64 * 16),
                ("kernel_ms", C.c_double * 16), ("verify_ms", C.c_double),
                ("fallback_ms", C.c_double), ("fallback_count", C.c_int64),
                ("total_pivots", C.c_int64), ("total_ms", C.c_double),
                ("warmstart_ms", C.c_double), ("param", C.c_int32 * 16)]


class _BatchResults(C.Structure):
    _fields_ = [("count", C.c_int64), ("status", C.POINTER(C.c_int32)),
                ("pivots", C.POINTER(C.c_int32)), ("phase1_pivots", C.POINTER(C.c_int32)),
                ("fallback_used", C.POINTER(C.c_int32)), ("objective", C.POINTER(C.c_double)),
                ("kkt_primal", C.POINTER(C.c_double)), ("kkt_dual", C.POINTER(C.c_double)),
                ("x", C.POINTER(C.c_double)), ("basis", C.POINTER(C.c_int32)),
                ("x_offset", C.POINTER(C.c_int64)), ("basis_offset", C.POINTER(C.c_int64))]


class _ServerConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_int64), ("max_batch", C.c_int32), ("max_wait_us", C.c_int32),
                ("precision", C.c_int32), ("arenas", C.c_int32)]


class _ServerResult(C.Structure):
    _fields_ = [("ticket", C.c_int64), ("status", C.c_int32), ("pivots", C.c_int32),
                ("objective", C.c_double), ("latency_us", C.c_double)]


class _ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("rejected", C.c_int64),
                ("batches", C.c_int64), ("mean_batch", C.c_double), ("p50_us", C.c_double),
                ("p99_us", C.c_double), ("max_us", C.c_double), ("resident_bytes", C.c_int64)]


class _WsConfig(C.Structure):
    _fields_ = [("hidden", C.c_int32), ("steps", C.c_int32), ("lr", C.c_float),
                ("weight_decay", C.c_float), ("seed", C.c_uint64)]


# every symbol include/blpg/blpg.h declares (checked by the CPU test suite)
ABI_SYMBOLS = [
    "blpg_default_options", "blpg_create", "blpg_destroy", "blpg_last_error",
    "blpg_set_options", "blpg_solve_batch", "blpg_solve_shared_A", "blpg_batch_upload",
    "blpg_batch_solve", "blpg_batch_sync", "blpg_batch_download", "blpg_batch_free",
    "blpg_batch_repack",
    "blpg_batch_fetch", "blpg_autotune", "blpg_tune_set", "blpg_tune_export", "blpg_tune_import",
    "blpg_batch_info_get", "blpg_gen_config", "blpg_version", "blpg_ws_create",
    "blpg_ws_destroy", "blpg_ws_train", "blpg_ws_predict", "blpg_batch_set_warmstart",
    "blpg_ws_params", "blpg_ws_norm", "blpg_ws_gradient", "blpg_server_create",
    "blpg_server_destroy", "blpg_server_submit", "blpg_server_poll", "blpg_server_stats_get",
    "blpg_server_replay", "blpg_server_set_warmstart", "blpg_solve_shared_A_noalloc",
    "blpg_noalloc_max_count",
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
    lib.blpg_batch_repack.argtypes = [C.c_void_p, C.c_void_p, P(_LP), C.c_int64]
    lib.blpg_batch_fetch.argtypes = [C.c_void_p, C.c_void_p, P(_BatchResults)]
    lib.blpg_autotune.argtypes = [C.c_void_p, P(_LP), C.c_int64, C.c_int32]
    lib.blpg_tune_set.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                  C.c_int32]
    lib.blpg_tune_export.argtypes = [C.c_void_p, C.c_char_p, C.c_int64, P(C.c_int64)]
    lib.blpg_tune_import.argtypes = [C.c_void_p, C.c_char_p]
    lib.blpg_batch_info_get.argtypes = [C.c_void_p, C.c_void_p, P(_Info)]
    lib.blpg_gen_config.argtypes = [C.c_int32, C.c_uint64, C.c_int32, C.c_int64, C.c_int64,
                                    C.c_int64, C.c_int32, C.c_int32, P(C.c_double),
                                    P(C.c_double), P(C.c_double), P(C.c_int8), P(C.c_int32)]
    lib.blpg_version.restype = C.c_char_p
    V = C.c_void_p  # device pointers and the stream
    lib.blpg_solve_shared_A_noalloc.argtypes = [C.c_int32, C.c_int32, V, V, V, V, C.c_int64,
                                                V, V, V, V, V, V]
    lib.blpg_noalloc_max_count.argtypes = [C.c_int32, C.c_int32]
    lib.blpg_noalloc_max_count.restype = C.c_int64
    F, I32 = P(C.c_float), P(C.c_int32)
    lib.blpg_server_create.argtypes = [C.c_void_p, P(_ServerConfig), P(C.c_void_p)]
    lib.blpg_server_destroy.argtypes = [C.c_void_p]
    lib.blpg_server_submit.argtypes = [C.c_void_p, P(_LP), P(C.c_double), I32, P(C.c_int64)]
    lib.blpg_server_poll.argtypes = [C.c_void_p, P(_ServerResult), C.c_int32, C.c_int32, I32]
    lib.blpg_server_stats_get.argtypes = [C.c_void_p, P(_ServerStats)]
    lib.blpg_server_set_warmstart.argtypes = [C.c_void_p, C.c_void_p]
    lib.blpg_server_replay.argtypes = [C.c_void_p, P(_LP), P(C.c_double), C.c_int64,
                                       P(_ServerResult)]
    lib.blpg_ws_create.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, P(_WsConfig),
                                   P(C.c_void_p)]
    lib.blpg_ws_destroy.argtypes = [C.c_void_p, C.c_void_p]
    lib.blpg_ws_train.argtypes = [C.c_void_p, C.c_void_p, P(_LP), I32, C.c_int64, F]
    lib.blpg_ws_predict.argtypes = [C.c_void_p, C.c_void_p, P(_LP), C.c_int64, F, I32]
    lib.blpg_batch_set_warmstart.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    lib.blpg_ws_params.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, F, F, P(C.c_int64)]
    lib.blpg_ws_norm.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, F, F]
    lib.blpg_ws_gradient.argtypes = [C.c_void_p, C.c_void_p, P(_LP), I32, C.c_int64, F, F]
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
    if config == 4 and count > 0 and np.array_equal(A, np.broadcast_to(A[0], A.shape)):
        # a parametric family shares one A: every LP refers to the same array, so a batch
        # upload stores it once (the pool deduplicates by pointer, as for shared-A batches)
        A0 = A[0].copy()
        return [LP(A0, b[k], c[k], None, family=shape_idx, flags=int(f[k]))
                for k in range(count)]
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

    def repack(self, lps: Sequence[LP]) -> None:
        """New LPs of the same shapes into the batch's pinned staging (host work); the next
        solve() copies them to the device before its kernels."""
        lps = list(lps)
        arr = (_LP * len(lps))(*[_lp_struct(lp) for lp in lps])
        self.solver._check(self.solver.lib.blpg_batch_repack(self.solver.ctx, self.h, arr,
                                                             len(lps)))
        self.lps = lps

    def sync(self) -> None:
        self.solver._check(self.solver.lib.blpg_batch_sync(self.solver.ctx))

    def download(self) -> Results:
        res, xs, bs = _alloc_results(self.lps)
        self.solver._check(self.solver.lib.blpg_batch_download(self.solver.ctx, self.h, res))
        return _collect(res, xs, bs)

    def fetch_raw(self) -> _BatchResults:
        """blpg_batch_fetch: results in the batch's pinned host buffers (no per-LP copies)."""
        r = _BatchResults()
        self.solver._check(self.solver.lib.blpg_batch_fetch(self.solver.ctx, self.h, C.byref(r)))
        return r

    def fetch(self) -> Results:
        r = self.fetch_raw()
        k = r.count
        if k == 0:
            z = np.zeros(0, np.int32)
            return Results(z, z, z, z, np.zeros(0), np.zeros(0), np.zeros(0), [], [])
        arr = lambda p, n: np.ctypeslib.as_array(p, shape=(n,)).copy()
        nx = sum(lp.n for lp in self.lps)
        nb = sum(lp.m for lp in self.lps)
        xo = arr(r.x_offset, k)
        bo = arr(r.basis_offset, k)
        xpool, bpool = arr(r.x, nx), arr(r.basis, nb)
        xs = [xpool[xo[i]:xo[i] + lp.n] for i, lp in enumerate(self.lps)]
        bs = [bpool[bo[i]:bo[i] + lp.m] for i, lp in enumerate(self.lps)]
        return Results(arr(r.status, k), arr(r.pivots, k), arr(r.phase1_pivots, k),
                       arr(r.fallback_used, k), arr(r.objective, k), arr(r.kkt_primal, k),
                       arr(r.kkt_dual, k), xs, bs)

    def info(self) -> dict:
        inf = _Info()
        self.solver._check(self.solver.lib.blpg_batch_info_get(self.solver.ctx, self.h,
                                                               C.byref(inf)))
        nb = inf.nbins
        return {"bins": [{"kernel": inf.kernel[i], "rcap": inf.rcap[i], "ccap": inf.ccap[i],
                          "count": inf.count[i], "kernel_ms": inf.kernel_ms[i],
                          "param": inf.param[i]}
                         for i in range(nb)],
                "verify_ms": inf.verify_ms, "fallback_ms": inf.fallback_ms,
                "fallback_count": inf.fallback_count, "total_pivots": inf.total_pivots,
                "total_ms": inf.total_ms, "warmstart_ms": inf.warmstart_ms}

    def set_warmstart(self, model: Optional["WarmStart"]) -> None:
        """Crash the batch's family LPs from `model`'s prediction on every later solve()."""
        self.model = model
        self.solver._check(self.solver.lib.blpg_batch_set_warmstart(
            self.solver.ctx, self.h, model.h if model is not None else None))

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
                 hbm_budget: int = 0, device: int = 0, use_graph: bool = True,
                 sibling_warm: bool = False):
        self.lib = load_library()
        o = _Options()
        self.lib.blpg_default_options(C.byref(o))
        o.precision = {"fp64": FP64, "fp32": FP32}[precision]
        o.kernel = _KERNELS[kernel]
        o.verify = int(verify)
        o.max_iter, o.bland_after = max_iter, bland_after
        o.eps_d, o.eps_p, o.tol_feas, o.kkt_tol = eps_d, eps_p, tol_feas, kkt_tol
        o.hbm_budget, o.device, o.use_graph = hbm_budget, device, int(use_graph)
        o.sibling_warm = int(sibling_warm)
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

    def autotune(self, lps: Sequence[LP], reps: int = 3) -> list[tuple]:
        """Measure every kernel instance per LP shape and route to the fastest, and the
        shape's check plan (kernel 0 rows: param 1 skips the first check tier); returns the
        tuning table as (precision, m, n, kernel, param, ms) tuples."""
        lps = list(lps)
        arr = (_LP * len(lps))(*[_lp_struct(lp) for lp in lps])
        self._check(self.lib.blpg_autotune(self.ctx, arr, len(lps), reps))
        return self.tune_table()

    def tune_set(self, m: int, n: int, kernel: str, param: int, precision: Optional[str] = None):
        p = self.opts.precision if precision is None else {"fp64": FP64, "fp32": FP32}[precision]
        self._check(self.lib.blpg_tune_set(self.ctx, p, m, n, _KERNELS[kernel], param))

    def tune_export(self) -> str:
        n = C.c_int64()
        self._check(self.lib.blpg_tune_export(self.ctx, None, 0, C.byref(n)))
        buf = C.create_string_buffer(n.value + 1)
        self._check(self.lib.blpg_tune_export(self.ctx, buf, n.value + 1, C.byref(n)))
        return buf.value.decode()

    def tune_import(self, text: str) -> None:
        self._check(self.lib.blpg_tune_import(self.ctx, text.encode()))

    def tune_table(self) -> list[tuple]:
        rows = []
        for line in self.tune_export().splitlines():
            p, m, n, k, q, ms = line.split()
            rows.append((int(p), int(m), int(n), int(k), int(q), float(ms)))
        return rows

    def close(self) -> None:
        if getattr(self, "ctx", None):
            self.lib.blpg_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _f32p(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_float))


def _i32p(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_int32))


class WarmStart:
    """Learned warm start (blpg_ws_*): one network per parametric family of shape m x n,
    trained on the GPU from optimal bases, evaluated on the GPU at solve time."""

    def __init__(self, solver: Solver, families: int, m: int, n: int, hidden: int = 256,
                 steps: int = 300, lr: float = 1e-3, weight_decay: float = 0.0, seed: int = 0):
        self.solver, self.families, self.m, self.n = solver, families, m, n
        cfg = _WsConfig(hidden, steps, lr, weight_decay, seed)
        self.steps = steps
        h = C.c_void_p()
        solver._check(solver.lib.blpg_ws_create(solver.ctx, families, m, n, C.byref(cfg),
                                                C.byref(h)))
        self.h = h
        cnt = C.c_int64()
        solver._check(solver.lib.blpg_ws_params(solver.ctx, h, 0, None, None, C.byref(cnt)))
        self.n_params = cnt.value

    def _lps(self, lps):
        lps = list(lps)
        return lps, (_LP * len(lps))(*[_lp_struct(lp) for lp in lps])

    def train(self, lps: Sequence[LP], bases: Sequence[np.ndarray]) -> np.ndarray:
        lps, arr = self._lps(lps)
        B = np.ascontiguousarray(np.stack(bases), dtype=np.int32)
        loss = np.zeros(self.steps, dtype=np.float32)
        self.solver._check(self.solver.lib.blpg_ws_train(self.solver.ctx, self.h, arr, _i32p(B),
                                                         len(lps), _f32p(loss)))
        return loss

    def predict(self, lps: Sequence[LP]):
        lps, arr = self._lps(lps)
        logits = np.zeros((len(lps), self.n + self.m), dtype=np.float32)
        warm = np.zeros((len(lps), self.m), dtype=np.int32)
        self.solver._check(self.solver.lib.blpg_ws_predict(self.solver.ctx, self.h, arr, len(lps),
                                                           _f32p(logits), _i32p(warm)))
        return logits, warm

    def params(self, family: int, set: Optional[np.ndarray] = None) -> np.ndarray:
        out = np.zeros(self.n_params, dtype=np.float32)
        src = None if set is None else np.ascontiguousarray(set, dtype=np.float32)
        self.solver._check(self.solver.lib.blpg_ws_params(self.solver.ctx, self.h, family,
                                                          _f32p(out), _f32p(src), None))
        return out

    def norm(self, family: int):
        d = self.m + self.n
        mu = np.zeros(d, dtype=np.float32); isd = np.zeros(d, dtype=np.float32)
        self.solver._check(self.solver.lib.blpg_ws_norm(self.solver.ctx, self.h, family,
                                                        _f32p(mu), _f32p(isd)))
        return mu, isd

    def gradient(self, lps: Sequence[LP], bases: Sequence[np.ndarray]):
        lps, arr = self._lps(lps)
        B = np.ascontiguousarray(np.stack(bases), dtype=np.int32)
        g = np.zeros((self.families, self.n_params), dtype=np.float32)
        loss = np.zeros(self.families, dtype=np.float32)
        self.solver._check(self.solver.lib.blpg_ws_gradient(self.solver.ctx, self.h, arr,
                                                            _i32p(B), len(lps), _f32p(g),
                                                            _f32p(loss)))
        return g, loss

    def layout(self):
        """(name, shape) of the parameter blocks in params() / gradient() order."""
        H, d = self.params_hidden(), self.m + self.n
        return [("W1", (H, d)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)),
                ("W3", (d, H)), ("b3", (d,))]

    def params_hidden(self) -> int:
        d = self.m + self.n
        # n_params = H*d + H + H*H + H + d*H + d
        a, b, c = 1, 2 * d + 2, d - self.n_params
        return int(round((-b + (b * b - 4 * a * c) ** 0.5) / 2))

    def close(self) -> None:
        if getattr(self, "h", None):
            self.solver.lib.blpg_ws_destroy(self.solver.ctx, self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Serving loop (blpg_server_*): single LPs in, batches formed on the device side of a
    queue, replayed through captured hipGraphs, results out with their latency."""

    def __init__(self, solver: Solver, hbm_budget: int = 0, max_batch: int = 0,
                 max_wait_us: int = 0, precision: str = "fp64", arenas: int = 0):
        self.solver = solver
        cfg = _ServerConfig(hbm_budget, max_batch, max_wait_us,
                            {"fp64": FP64, "fp32": FP32}[precision], arenas)
        h = C.c_void_p()
        solver._check(solver.lib.blpg_server_create(solver.ctx, C.byref(cfg), C.byref(h)))
        self.h = h
        self._keep = {}

    def submit(self, lp: LP, want_x: bool = False) -> int:
        """Returns the ticket; raises BlpgError(E_BUDGET) when admission is refused."""
        s = _lp_struct(lp)
        x = np.zeros(lp.n) if want_x else None
        bs = np.zeros(lp.m, dtype=np.int32) if want_x else None
        t = C.c_int64()
        rc = self.solver.lib.blpg_server_submit(self.h, C.byref(s), _dp(x), _i32p(bs), C.byref(t))
        if rc != 0:
            raise BlpgError(rc, "blpg_server_submit refused the LP")
        self._keep[t.value] = (lp, x, bs)
        return t.value

    def poll(self, max_results: int = 1024, wait_us: int = 0) -> list[dict]:
        buf = (_ServerResult * max_results)()
        n = C.c_int32()
        self.solver._check(self.solver.lib.blpg_server_poll(self.h, buf, max_results, wait_us,
                                                            C.byref(n)))
        out = []
        for i in range(n.value):
            r = buf[i]
            lp, x, bs = self._keep.pop(r.ticket, (None, None, None))
            out.append({"ticket": r.ticket, "status": r.status, "pivots": r.pivots,
                        "objective": r.objective, "latency_us": r.latency_us, "x": x,
                        "basis": bs})
        return out

    def set_warmstart(self, model: Optional["WarmStart"]) -> None:
        """Crash requests tagged with one of `model`'s families (lp.family) from its
        prediction; call before submitting tagged requests."""
        self.solver._check(self.solver.lib.blpg_server_set_warmstart(
            self.h, model.h if model is not None else None))
        self._model = model  # kept alive while attached

    def stats(self) -> dict:
        st = _ServerStats()
        self.solver._check(self.solver.lib.blpg_server_stats_get(self.h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in _ServerStats._fields_}

    def replay(self, lps: Sequence[LP], arrival_us: np.ndarray):
        """Submit lps[k] at arrival_us[k] (from now) and wait for all; per-LP results."""
        lps = list(lps)
        arr = (_LP * len(lps))(*[_lp_struct(lp) for lp in lps])
        t = np.ascontiguousarray(arrival_us, dtype=np.float64)
        out = (_ServerResult * len(lps))()
        rc = self.solver.lib.blpg_server_replay(self.h, arr, _dp(t), len(lps), out)
        if rc != 0:  # the replay's own message (a stall report) is kept per calling thread
            raise BlpgError(rc, self.solver.lib.blpg_last_error(None).decode())
        dt = np.dtype([("ticket", np.int64), ("status", np.int32), ("pivots", np.int32),
                       ("objective", np.float64), ("latency_us", np.float64)])
        assert dt.itemsize == C.sizeof(_ServerResult)
        return np.frombuffer((C.c_char * (C.sizeof(_ServerResult) * len(lps))).from_address(
            C.addressof(out)), dtype=dt, count=len(lps)).copy()

    def close(self) -> None:
        if getattr(self, "h", None):
            self.solver.lib.blpg_server_destroy(self.h)
            self.h = None

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
