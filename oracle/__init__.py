"""Python loader for the CPU oracle (liboracle.so).  TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this module,
and only as the checker.  The product (simplexprojects_amd) never imports it.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path
from typing import Optional

import numpy as np

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "liboracle.so"

OPTIMAL, INFEASIBLE, UNBOUNDED, ITER_LIMIT, NUMERIC = 5, 4, 6, 100, 101


class _LP(C.Structure):
    _fields_ = [("m", C.c_int32), ("n", C.c_int32), ("A", C.POINTER(C.c_double)),
                ("b", C.POINTER(C.c_double)), ("c", C.POINTER(C.c_double)),
                ("sense", C.POINTER(C.c_int8))]


class _Opts(C.Structure):
    _fields_ = [("eps_d", C.c_double), ("eps_p", C.c_double), ("tol_feas", C.c_double),
                ("max_iter", C.c_int32), ("bland_after", C.c_int32)]


class _Res(C.Structure):
    _fields_ = [("status", C.c_int32), ("pivots", C.c_int32), ("phase1_pivots", C.c_int32),
                ("unb_var", C.c_int32), ("objective", C.c_double), ("w_final", C.c_double),
                ("x", C.POINTER(C.c_double)), ("basis", C.POINTER(C.c_int32))]


_lib: Optional[C.CDLL] = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise FileNotFoundError(f"{LIB_PATH} not built (make -C oracle)")
        L = C.CDLL(str(LIB_PATH))
        P = C.POINTER
        for nm in ("orc_solve_d", "orc_solve_f"):
            getattr(L, nm).argtypes = [P(_LP), P(_Opts), P(_Res)]
        for nm in ("orc_solve_warm_d", "orc_solve_warm_f"):
            getattr(L, nm).argtypes = [P(_LP), P(_Opts), P(C.c_int32), P(_Res)]
        L.orc_solve_tab_d.argtypes = [P(_LP), P(_Opts), P(_Res), P(C.c_double), P(C.c_int32)]
        L.orc_solve_tab_f.argtypes = [P(_LP), P(_Opts), P(_Res), P(C.c_float), P(C.c_int32)]
        L.orc_ref_rule_solve.argtypes = [C.c_int32, C.c_int32, P(C.c_double), P(C.c_double),
                                         P(C.c_double), C.c_int32, P(C.c_double), P(C.c_int32)]
        L.orc_kkt_check.argtypes = [P(_LP), P(C.c_int32), P(C.c_double), P(C.c_double),
                                    P(C.c_double), P(C.c_double), P(C.c_double)]
        L.orc_solve_batch_d.argtypes = [C.c_int64, P(C.c_int32), P(C.c_int32), P(C.c_int64),
                                        P(C.c_int64), P(C.c_int64), P(C.c_int64),
                                        P(C.c_double), P(C.c_int8), P(_Opts), P(C.c_int32),
                                        P(C.c_int32), P(C.c_double), C.c_int]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _mk(A, b, c, sense):
    A = np.ascontiguousarray(A, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    c = np.ascontiguousarray(c, dtype=np.float64)
    s = None if sense is None else np.ascontiguousarray(sense, dtype=np.int8)
    L = _LP()
    L.m, L.n = A.shape
    L.A, L.b, L.c = _dp(A), _dp(b), _dp(c)
    L.sense = None if s is None else s.ctypes.data_as(C.POINTER(C.c_int8))
    return L, (A, b, c, s)


def defaults(precision: str, m: int, n: int) -> dict:
    """The library's option defaults (blpg_options fields left 0)."""
    if precision == "fp32":
        d = {"eps_d": 1e-6, "eps_p": 1e-6, "tol_feas": 1e-4}
    else:
        d = {"eps_d": 1e-9, "eps_p": 1e-9, "tol_feas": 1e-7}
    d["max_iter"] = 50 * (m + n) + 1000
    d["bland_after"] = 50
    return d


def solve(A, b, c, sense=None, precision: str = "fp64", tab: bool = False, warm=None,
          **kw) -> dict:
    """Replay the build's pivot rule on the CPU.  Keyword overrides: eps_d, eps_p,
    tol_feas, max_iter, bland_after.  `warm`: target basis (m variable ids) to crash in."""
    L, keep = _mk(A, b, c, sense)
    m, n = keep[0].shape
    d = defaults(precision, m, n)
    d.update({k: v for k, v in kw.items() if v is not None})
    o = _Opts(d["eps_d"], d["eps_p"], d["tol_feas"], d["max_iter"], d["bland_after"])
    x = np.zeros(n)
    basis = np.zeros(m, dtype=np.int32)
    r = _Res()
    r.x = _dp(x)
    r.basis = basis.ctypes.data_as(C.POINTER(C.c_int32))
    out = {}
    if warm is not None:
        w = np.ascontiguousarray(warm, dtype=np.int32)
        fn = lib().orc_solve_warm_f if precision == "fp32" else lib().orc_solve_warm_d
        rc = fn(C.byref(L), C.byref(o), w.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(r))
    elif tab:
        dt = np.float32 if precision == "fp32" else np.float64
        T = np.zeros((m + 2, n + 2), dtype=dt)
        fn = lib().orc_solve_tab_f if precision == "fp32" else lib().orc_solve_tab_d
        ptr = T.ctypes.data_as(C.POINTER(C.c_float if precision == "fp32" else C.c_double))
        nb = np.zeros(n + 2, dtype=np.int32)
        rc = fn(C.byref(L), C.byref(o), C.byref(r), ptr, nb.ctypes.data_as(C.POINTER(C.c_int32)))
        out["tab"] = T
        out["nb"] = nb
    else:
        fn = lib().orc_solve_f if precision == "fp32" else lib().orc_solve_d
        rc = fn(C.byref(L), C.byref(o), C.byref(r))
    if rc != 0:
        raise ValueError("oracle rejected the LP")
    out.update(status=r.status, pivots=r.pivots, phase1_pivots=r.phase1_pivots,
               unb_var=r.unb_var, objective=r.objective, w_final=r.w_final, x=x, basis=basis)
    return out


def ref_rule_solve(A, b, c, max_iter: int = 100000):
    """The reference kernel's own rule (simplex.cu:35-385), phase II only."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    c = np.ascontiguousarray(c, dtype=np.float64)
    obj = C.c_double()
    piv = C.c_int32()
    rc = lib().orc_ref_rule_solve(A.shape[0], A.shape[1], _dp(A), _dp(b), _dp(c), max_iter,
                                  C.byref(obj), C.byref(piv))
    if rc != 0:
        raise ValueError("reference rule needs b >= 0")
    return obj.value, piv.value


def kkt_check(A, b, c, basis, sense=None) -> Optional[dict]:
    L, keep = _mk(A, b, c, sense)
    m, n = keep[0].shape
    basis = np.ascontiguousarray(basis, dtype=np.int32)
    x = np.zeros(n)
    y = np.zeros(m)
    obj, rp, rd = C.c_double(), C.c_double(), C.c_double()
    rc = lib().orc_kkt_check(C.byref(L), basis.ctypes.data_as(C.POINTER(C.c_int32)), _dp(x),
                             _dp(y), C.byref(obj), C.byref(rp), C.byref(rd))
    if rc != 0:
        return None
    return {"x": x, "y": y, "objective": obj.value, "res_primal": rp.value,
            "res_dual": rd.value}


def solve_batch(lps, threads: int = 1, precision: str = "fp64") -> dict:
    """FP64 replay of a whole batch over `threads` OpenMP threads (CPU baseline)."""
    return solve_packed(pack_batch(lps), threads)


def pack_batch(lps) -> dict:
    """The batch in the flat layout solve_packed takes (done once, outside any timing)."""
    k = len(lps)
    pool_parts, spool_parts = [], []
    a_off = np.zeros(k, np.int64); b_off = np.zeros(k, np.int64)
    c_off = np.zeros(k, np.int64); s_off = np.full(k, -1, np.int64)
    m = np.zeros(k, np.int32); n = np.zeros(k, np.int32)
    pos = 0; spos = 0
    for i, lp in enumerate(lps):
        m[i], n[i] = lp.A.shape
        for arr, off in ((lp.A, a_off), (lp.b, b_off), (lp.c, c_off)):
            off[i] = pos
            flat = np.ascontiguousarray(arr, dtype=np.float64).reshape(-1)
            pool_parts.append(flat)
            pos += flat.size
        if lp.sense is not None:
            s_off[i] = spos
            spool_parts.append(np.ascontiguousarray(lp.sense, dtype=np.int8))
            spos += lp.sense.size
    pool = np.concatenate(pool_parts)
    spool = np.concatenate(spool_parts) if spool_parts else np.zeros(1, np.int8)
    return {"k": k, "m": m, "n": n, "a_off": a_off, "b_off": b_off, "c_off": c_off,
            "s_off": s_off, "pool": pool, "spool": spool}


def solve_packed(pk: dict, threads: int = 1, precision: str = "fp64") -> dict:
    """FP64 replay of a packed batch over `threads` OpenMP threads."""
    k, m, n = pk["k"], pk["m"], pk["n"]
    a_off, b_off, c_off, s_off = pk["a_off"], pk["b_off"], pk["c_off"], pk["s_off"]
    pool, spool = pk["pool"], pk["spool"]
    mm, nn = int(m.max()), int(n.max())
    d = defaults(precision, mm, nn)
    d["max_iter"] = 0  # per-LP default below
    status = np.zeros(k, np.int32); piv = np.zeros(k, np.int32); obj = np.zeros(k)
    # the batch driver takes one option set; use the largest LP's cap
    o = _Opts(d["eps_d"], d["eps_p"], d["tol_feas"], 50 * (mm + nn) + 1000, d["bland_after"])
    P = lambda a, t: a.ctypes.data_as(C.POINTER(t))
    rc = lib().orc_solve_batch_d(k, P(m, C.c_int32), P(n, C.c_int32), P(a_off, C.c_int64),
                                 P(b_off, C.c_int64), P(c_off, C.c_int64), P(s_off, C.c_int64),
                                 _dp(pool), P(spool, C.c_int8), C.byref(o),
                                 P(status, C.c_int32), P(piv, C.c_int32), _dp(obj), threads)
    if rc != 0:
        raise ValueError("oracle batch failed")
    return {"status": status, "pivots": piv, "objective": obj}
