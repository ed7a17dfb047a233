"""CPU restatement of the FP64 check by iterative refinement (TEST INFRASTRUCTURE).

Given an LP and the solve's final condensed tableau T ((m+2) x (n+2), solve precision) with
the variable id of every column, the check rebuilds the basic solution and the duals in
FP64 without factorising B: the column of a nonbasic slack s_p is B^-1 e_p, a basic
slack's column of B^-1 is the unit vector of its row.  The device kernel does the same
sweeps (kernel_verify.hip); this module is its checker, written independently in numpy.
"""
from __future__ import annotations

import numpy as np

MAX_SWEEPS = 12


def _binv(T, bs, colp, n, m):
    """B~^-1 as an m x m matrix read from the tableau (rows: basis rows, cols: constraints)."""
    Bi = np.zeros((m, m))
    for p in range(m):
        if colp[p] >= 0:
            Bi[:, p] = T[2:, colp[p]].astype(np.float64)
        else:
            q = int(np.nonzero(bs == n + p)[0][0])
            Bi[q, p] = 1.0
    return Bi


def _refine(v, resid, correct, scale0, amax):
    """Sweeps until the rounding floor, MAX_SWEEPS, or a sweep that shrinks the residual
    by less than 10%.  Returns (v, max |residual|, max |v|)."""
    prev = np.inf
    sweep = 0
    while True:
        r = resid(v)
        res, vm = np.abs(r).max(initial=0.0), np.abs(v).max(initial=0.0)
        if sweep == MAX_SWEEPS or res <= 1e-14 * (scale0 + amax * vm) or res > 0.9 * prev:
            return v, res, vm
        prev = res
        v = v + correct(r)
        sweep += 1


def farkas_test(Ap, bp, eq, y, amax, tol):
    """FP64 Farkas test of a candidate y for {x >= 0, A'x <= b' ('=' on eq rows)}, with
    A', b' the sign-applied rows: y is clamped to >= 0 on inequality rows, then the LP is
    certified infeasible when (y^T A')_j >= -tol |A|_max |y|_inf for every column j and
    y^T b' <= -sqrt(tol) (1 + |b|_inf) |y|_inf (0 <= y^T A' x <= y^T b' < 0 for any
    feasible x).  Returns (ok, violation, margin), both relative to |y|_inf."""
    y = np.asarray(y, np.float64)
    ym = np.abs(y).max(initial=0.0)
    yneg = max(0.0, float((-y[~eq]).max(initial=0.0)))
    yc = np.where(eq, y, np.maximum(y, 0.0))
    vg = max(0.0, float((-(Ap.T @ yc)).max(initial=0.0)))
    h = float(bp @ yc)
    bm = np.abs(bp).max(initial=0.0)
    tiny = np.finfo(np.float64).tiny
    ok = bool(ym > 0 and vg <= tol * amax * ym and yneg <= tol * ym and
              -h >= np.sqrt(tol) * (1 + bm) * ym)
    return ok, max(vg / max(amax * ym, tiny), yneg / max(ym, tiny)), -h / ((1 + bm) * max(ym, tiny))


def _infeasible_check(Ap, bp, eq, bs, T, nb, amax, tol):
    """Phase-I duals of an FP32 INFEASIBLE result (x0, id n+m, basic): row 1 of the
    tableau at the nonbasic slack columns, refined against B^T y = c_B of the auxiliary
    problem (0 but at x0, whose row is left out), then the Farkas test."""
    m, n = Ap.shape
    fail = dict(ok=False, objective=0.0, kkt_p=-1.0, kkt_d=-1.0, x=None)
    if np.any(bs < 0) or np.any(bs > n + m) or int((bs == n + m).sum()) != 1:
        return fail
    colp = -np.ones(m, np.int64)
    for j in range(1, n + 2):
        v = int(nb[j])
        if n <= v < n + m:
            colp[v - n] = j
    if (colp >= 0).sum() != (bs < n).sum() + 1:
        return fail
    Bi = _binv(T, bs, colp, n, m)

    def resid(v):
        r = np.zeros(m)
        for i, bv in enumerate(bs):
            if bv < n:
                r[i] = -float(Ap[:, bv] @ v)
            elif bv < n + m:
                r[i] = -v[bv - n]
        return r  # (x0's row: 0)
    y0 = np.array([float(T[1, colp[p]]) if colp[p] >= 0 else 0.0 for p in range(m)])
    y, _, _ = _refine(y0, resid, lambda r: Bi.T @ r, 0.0, amax)
    ok, viol, margin = farkas_test(Ap, bp, eq, y, amax, tol)
    return dict(ok=ok, objective=0.0, kkt_p=viol, kkt_d=margin, x=None, y=y)


def refine_check(A, b, c, sense, status, basis, T, nb, unb_var, amax, tol):
    """Returns dict(ok, objective, kkt_p, kkt_d, x) for an OPTIMAL or UNBOUNDED result
    (ok is None when the basis is not a proper slack/structural basis), and for an FP32
    INFEASIBLE one (phase I ended with x0 basic; ok False without x0) dict(ok, kkt_p =
    relative Farkas violation, kkt_d = relative margin of y^T b' < 0, y)."""
    A = np.asarray(A, np.float64)
    m, n = A.shape
    sg = np.ones(m) if sense is None else np.where(np.asarray(sense) == 2, -1.0, 1.0)
    eq = np.zeros(m, bool) if sense is None else np.asarray(sense) == 1
    Ap = A * sg[:, None]
    bp = np.asarray(b, np.float64) * sg
    c = np.asarray(c, np.float64)
    bs = np.asarray(basis, np.int64)
    if status == "infeasible":
        return _infeasible_check(Ap, bp, eq, bs, T, nb, amax, tol)
    if np.any(bs < 0) or np.any(bs >= n + m):
        return dict(ok=None)
    colp = -np.ones(m, np.int64)
    for j in range(1, n + 2):
        v = int(nb[j])
        if n <= v < n + m:
            colp[v - n] = j
    if (bs < n).sum() != (colp >= 0).sum():
        return dict(ok=None)
    # basis matrix of [A' I] in FP64
    Bm = np.zeros((m, m))
    for i, v in enumerate(bs):
        if v < n:
            Bm[:, i] = Ap[:, v]
        else:
            Bm[v - n, i] = 1.0
    Bi = _binv(T, bs, colp, n, m)
    bm = np.abs(b).max(initial=0.0)
    cm = np.abs(c).max(initial=0.0)
    sq = np.sqrt(tol)

    x, rres, xm = _refine(T[2:, 0].astype(np.float64), lambda v: bp - Bm @ v,
                          lambda r: Bi @ r, 1 + bm, amax)
    pv = rres
    for i, v in enumerate(bs):
        if v < n:
            pv = max(pv, -x[i])
        else:
            pv = max(pv, abs(x[i]) if eq[v - n] else -x[i])
    kkt_p = pv / (1 + bm + amax * xm)
    pconv = rres <= sq * (1 + bm)

    cB = np.array([c[v] if v < n else 0.0 for v in bs])
    y0 = np.array([float(T[0, colp[p]]) if colp[p] >= 0 else 0.0 for p in range(m)])
    y, rr, ym = _refine(y0, lambda v: cB - Bm.T @ v, lambda r: Bi.T @ r, 1 + cm, amax)
    basic = np.zeros(n + m, bool)
    basic[bs] = True
    red = c - Ap.T @ y
    dv = max(rr, red[~basic[:n]].max(initial=0.0),
             max((-y[p] for p in range(m) if colp[p] >= 0 and not eq[p]), default=0.0))
    kkt_d = dv / (1 + cm + amax * ym)
    dconv = rr <= sq * (1 + cm)
    xs = np.zeros(n)
    for i, v in enumerate(bs):
        if v < n:
            xs[v] = x[i]
    obj = float(c @ xs)
    if status == "optimal":
        return dict(ok=bool(kkt_p <= tol and kkt_d <= tol and pconv and dconv),
                    objective=obj, kkt_p=kkt_p, kkt_d=kkt_d, x=xs)
    # Unbounded: a feasible point and an improving ray.  The point is the origin when it is
    # feasible (b' >= 0 on inequality rows, b = 0 on equality rows), else the final basis,
    # which must then pass the primal test.  The ray d = B^-1 a'_e (refined) must keep every
    # basic variable from decreasing (d <= 0 in tableau terms, |d| on basic '=' slacks) and
    # gain c_e - c_B^T d > 0.
    origin_ok = bool(np.all(np.where(eq, bp == 0, bp >= 0)))
    v = int(unb_var)
    jv = [j for j in range(1, n + 2) if int(nb[j]) == v]
    if not (0 <= v < n + m) or basic[v] or not jv or \
            (not origin_ok and (kkt_p > tol or not pconv)):
        return dict(ok=False, objective=obj, kkt_p=kkt_p, kkt_d=-1.0, x=xs)
    av = Ap[:, v] if v < n else np.eye(m)[v - n]
    d, rres2, dmax = _refine(T[2:, jv[0]].astype(np.float64), lambda u: av - Bm @ u,
                             lambda r: Bi @ r, 1.0, amax)
    viol = max((abs(d[i]) if (bs[i] >= n and eq[bs[i] - n]) else d[i]) for i in range(m))
    gain = (c[v] if v < n else 0.0) - sum(float(c[bs[i]] * d[i]) for i in range(m) if bs[i] < n)
    good = viol <= tol * (1 + dmax) and rres2 <= tol * (1 + amax * dmax) and \
        gain > tol * (1 + cm * (1 + dmax))
    return dict(ok=bool(good), objective=obj, kkt_p=kkt_p, kkt_d=-1.0, x=xs)
