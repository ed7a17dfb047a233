"""Shared test helpers: fixtures loading, independent checkers (HiGHS, vertex enumeration,
exact-rational simplex).  Test infrastructure only."""
from __future__ import annotations

import hashlib
import itertools
import json
from fractions import Fraction
from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden"

# the 7 hand LPs of BLPG-Solver/src/parseBenchmark.cpp:128-319 (all maximisation) with the
# optimum established here (HiGHS, vertex enumeration); three comments there are wrong:
# parseSmallLP is optimal at 2 (":135 unbounded"), parseSmallLP5 is 8 (":230 39.444"),
# parse5Dim with c = e2 is 0.01 (":274 -0.99").
HAND_LPS = {
    "parseSmallLP": ([[2, -1], [1, -5]], [2, -4], [2, -1], 2.0),
    "parseSmallLP2": ([[1, 1], [-1, 1], [5, 4]], [10, -3, 35], [5, 6], 355.0 / 9.0),
    "parseSmallLP3": ([[1, -1], [-1, -1], [-1, 4]], [8, -3, 2], [1, 3], 64.0 / 3.0),
    "parseSmallLP4": ([[1, 1, -1], [-3, 1, 0], [1, -2, -2]], [7, -8, 0], [2, 0, -6], 28.0 / 3.0),
    "parseSmallLP5": ([[4, -1], [2, 1], [-5, 2]], [8, 10, 2], [1, 1], 8.0),
}


def parse5dim():
    A = np.zeros((10, 5))
    for i in range(5):
        A[2 * i, i] = 1
        A[2 * i + 1, i] = -1
    b = np.array([1.01, -0.99, 0.01, 0.01, 0.01, 0.01, 0.01, 0.01, 0.01, 0.01])
    c = np.array([0, 1.0, 0, 0, 0])
    return A, b, c, 0.01


def parse28dim_helicopter():
    A = np.zeros((56, 28))
    b = np.zeros(56)
    for i in range(28):
        A[2 * i, i] = 1
        A[2 * i + 1, i] = -1
        if i < 8:
            b[2 * i] = 0.1
    c = np.zeros(28)
    c[0] = 1
    return A, b, c, 0.1


def load_netlib():
    meta = json.loads((GOLDEN / "netlib19.json").read_text())
    z = np.load(GOLDEN / "netlib19.npz")
    out = {}
    for name, mt in meta.items():
        k = mt["key"]
        m, n = (int(v) for v in z[k + "_shape"])
        A = np.zeros((m, n))
        A[z[k + "_rows"], z[k + "_cols"]] = z[k + "_vals"]
        # the reference solves min c^T x as max (-c)^T x (GPU_Tester.cu:119-127)
        out[name] = (A, z[k + "_b"].copy(), -z[k + "_c"], mt)
    return out


def load_highs(cfg: int) -> dict:
    return json.loads((GOLDEN / f"highs_config{cfg}.json").read_text())


def data_hash(lps) -> str:
    h = hashlib.sha256()
    for lp in lps:
        for a in (lp.A, lp.b, lp.c):
            h.update(np.ascontiguousarray(a, dtype=np.float64).tobytes())
        if lp.sense is not None:
            h.update(np.ascontiguousarray(lp.sense, dtype=np.int8).tobytes())
    return h.hexdigest()


def rel_err(a, b) -> float:
    return abs(a - b) / max(1.0, abs(b))


def highs(A, b, c, sense=None):
    """Independent checker: scipy's HiGHS -> (status 5/4/6, objective or None)."""
    from scipy.optimize import linprog
    m = A.shape[0]
    s = np.zeros(m, np.int8) if sense is None else np.asarray(sense)
    ub, eq = s != 1, s == 1
    Aub = np.where((s[ub] == 2)[:, None], -A[ub], A[ub])
    bub = np.where(s[ub] == 2, -b[ub], b[ub])
    kw = dict(A_ub=Aub if ub.any() else None, b_ub=bub if ub.any() else None,
              A_eq=A[eq] if eq.any() else None, b_eq=b[eq] if eq.any() else None,
              bounds=(0, None))
    for method in ("highs", "highs-ds", "highs-ipm"):
        r = linprog(-c, method=method, **kw)
        if r.status == 0:
            return 5, float(-r.fun)
        if r.status == 3:
            return 6, None
        if r.status == 2:
            # HiGHS' presolve can say "infeasible" for "infeasible or unbounded":
            # a feasibility-only solve decides
            f = linprog(np.zeros_like(c), method="highs", **kw)
            return (6 if f.status == 0 else 4), None
    raise RuntimeError("HiGHS undecided")


def vertex_enum(A, b, c):
    """Brute force over all bases of [A | I] (x >= 0, A x <= b): the best feasible basic
    solution's objective, or None if no basic solution is feasible."""
    m, n = A.shape
    M = np.hstack([A, np.eye(m)])
    cc = np.concatenate([c, np.zeros(m)])
    best = None
    for cols in itertools.combinations(range(n + m), m):
        B = M[:, cols]
        if abs(np.linalg.det(B)) < 1e-12:
            continue
        xb = np.linalg.solve(B, b)
        if (xb < -1e-9).any():
            continue
        z = float(cc[list(cols)] @ xb)
        best = z if best is None else max(best, z)
    return best


def exact_simplex(A, b, c, max_iter=500):
    """The build's pivot rule (phase II, b >= 0) in exact rational arithmetic:
    Dantzig pricing with lowest-index ties; ratio key theta (or -a for degenerate rows),
    lowest row on ties.  Returns (status, objective Fraction, pivots)."""
    m, n = len(b), len(c)
    T = [[Fraction(0)] * (n + 2) for _ in range(m + 2)]
    for j in range(n):
        T[0][1 + j] = -Fraction(c[j])
    for i in range(m):
        T[2 + i][0] = Fraction(b[i])
        for j in range(n):
            T[2 + i][1 + j] = Fraction(A[i][j])
    live = [False] + [True] * n + [False]
    piv = 0
    while piv < max_iter:
        cand = [(T[0][j], j) for j in range(1, n + 2) if live[j] and T[0][j] < 0]
        if not cand:
            return 5, T[0][0], piv
        s = min(cand)[1]
        rows = []
        for i in range(2, m + 2):
            a = T[i][s]
            if a > 0:
                th = max(T[i][0], Fraction(0)) / a
                rows.append((th if th > 0 else -a, i))
        if not rows:
            return 6, None, piv
        r = min(rows)[1]
        p = T[r][s]
        prow = [T[r][j] / p for j in range(n + 2)]
        prow[s] = 1 / p
        for i in range(m + 2):
            if i == r:
                continue
            ci = T[i][s]
            for j in range(n + 2):
                old = Fraction(0) if j == s else T[i][j]
                T[i][j] = old - ci * prow[j]
        T[r] = prow
        piv += 1
    return 100, None, piv


# Hall & McKinnon's 2x4 LP that cycles under Dantzig's rule whatever the ratio-test
# tie-break; unbounded (HiGHS).  With the row sum(x) <= 1 added it is optimal at 0.875.
HALL_MCKINNON = (np.array([[0.4, 0.2, -1.4, -0.2], [-7.8, -1.4, 7.8, 0.4]]), np.zeros(2),
                 np.array([2.3, 2.15, -13.55, -0.4]))
# Beale's example, max form; optimum 5/4.
BEALE = (np.array([[0.25, -8, -1, 9], [0.5, -12, -0.5, 3], [0, 0, 1, 0]]), np.array([0, 0, 1.0]),
         np.array([0.75, -20, 0.5, -6]))


def two_phase_feasible(m, n, count, seed):
    """Planted-feasible two-phase LPs of the two-phase batch's shapes: A ~ U[-1,1) with 50%
    of entries zeroed, round(0.3 m) '=' rows and the other rows '<=' / '>=' half and half,
    built around a planted point x* ~ U[0,2) (b = A x*, plus U[0.1,1) of slack on
    inequality rows), so b has mixed signs and phase I is needed; the last row
    sum(x) <= sum(x*) + 5 keeps the LP bounded; c ~ N(0,1).  Every LP is feasible and
    bounded.  Returns (A, b, c, sense) tuples (numpy PCG64, identical on every host)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        A = rng.uniform(-1.0, 1.0, (m, n))
        A[rng.random((m, n)) < 0.5] = 0.0
        x = rng.uniform(0.0, 2.0, n)
        s = np.zeros(m, np.int8)
        rows = rng.permutation(m - 1)
        neq = int(round(0.3 * m))
        s[rows[:neq]] = 1
        s[rows[neq:neq + (m - 1 - neq) // 2]] = 2
        u = rng.uniform(0.1, 1.0, m)
        b = A @ x + np.where(s == 0, u, np.where(s == 2, -u, 0.0))
        A[-1] = 1.0
        b[-1] = x.sum() + 5.0
        out.append((A, b, rng.normal(size=n), s))
    return out


def load_two_phase():
    """HiGHS fixture of the planted-feasible two-phase LPs, with the LPs regenerated and
    checked against the fixture's data hash: [(group, [(A, b, c, sense), ...]), ...]."""
    import hashlib
    import json
    g = json.loads((GOLDEN / "highs_two_phase.json").read_text())
    out = []
    for grp in g["groups"]:
        lps = two_phase_feasible(grp["m"], grp["n"], grp["count"], grp["seed"])
        h = hashlib.sha256()
        for A, b, c, s in lps:
            for a in (A, b, c):
                h.update(np.ascontiguousarray(a, dtype=np.float64).tobytes())
            h.update(s.tobytes())
        assert h.hexdigest() == grp["sha256"], "two-phase generator changed"
        out.append((grp, lps))
    return out
