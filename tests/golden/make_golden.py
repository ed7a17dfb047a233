#!/usr/bin/env python3
"""Writes the golden fixtures under tests/golden/ (run in the build container; the GPU
box only reads the committed outputs).

  netlib19.npz        the 19 dense netlib conversions the reference ships
                      (CPLEXvsGLPK/benchmarksGLPKandCPLEX/benchmarks/<NAME>_{A,B,C}.txt,
                      format read as in BLPG-Solver/src/parseBenchmark.cpp:59-126:
                      "rows cols" header then dense A rows; one value per line for b, c),
                      stored sparse (row, col, value) + b + c per LP.
  netlib19.json       per LP: shape, the netlib optimal value quoted in the reference
                      (CPLEXvsGLPK/SourceCode/src/parseBenchmark.cpp:18-38; minimisation),
                      and the HiGHS optimum of the shipped dense file (status, objective).
  highs_config{1,2,3,4}.json
                      HiGHS status / objective for seeded subsets of the synthetic
                      configurations, with a SHA-256 of the generated data so that the
                      generator is pinned too.
  highs_two_phase.json
                      HiGHS optima of planted-feasible two-phase LPs at 174x142 and
                      205x203 (tests/common.py two_phase_feasible: mixed senses and
                      mixed-sign b around a planted point), with a SHA-256 of the data.

HiGHS is scipy.optimize.linprog(method="highs") (scipy 1.15.3 in this image).  When it
reports "unknown" the model is re-solved with highs-ds, highs-ipm and as a pure
feasibility problem, and the first definite answer is recorded.
"""
from __future__ import annotations

import hashlib
import json
import sys
from pathlib import Path

import numpy as np
from scipy.optimize import linprog

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))

REF_NETLIB = Path("/root/reference/CPLEXvsGLPK/benchmarksGLPKandCPLEX/benchmarks")
REF_BLPG = Path("/root/reference/benchmarks")

# netlib optimal values (minimisation), CPLEXvsGLPK/SourceCode/src/parseBenchmark.cpp:18-38
NETLIB_OPT = {
    "ADLITTLE": 2.2549496316e05, "AFIRO": -4.6475314286e02, "BEACONFD": 3.3592485807e04,
    "BLEND": -3.0812149846e01, "BOEING2": -3.1501872802e02, "BRANDY": 1.5185098965e03,
    "E226": -1.8751929066e01, "ISRAEL": -8.9664482186e05, "KB2": -1.7499001299e03,
    "RECIPELP": -2.6661600000e02, "SC105": -5.2202061212e01, "SC205": -5.2202061212e01,
    "SC50A": -6.4575077059e01, "SC50B": -7.0000000000e01, "SCAGR7": -2.3313892548e06,
    "SHARE1B": -7.6589318579e04, "SHARE2B": -4.1573224074e02, "STOCFOR1": -4.1131976219e04,
    "VTP-BASE": 1.2983146246e05,
}
BLPG8 = ["ADLITTLE", "AFIRO", "BLEND", "ISRAEL", "SC105", "SC205", "SC50A", "SC50B"]


def read_dense(prefix: Path):
    lines = (prefix.parent / (prefix.name + "_A.txt")).read_text().splitlines()
    m, n = (int(t) for t in lines[0].split()[:2])
    A = np.zeros((m, n))
    for i, line in enumerate(lines[1:1 + m]):
        vals = [float(t) for t in line.split()]
        A[i, :len(vals)] = vals
    b = np.array([float(l.split()[0]) for l in
                  (prefix.parent / (prefix.name + "_B.txt")).read_text().splitlines() if l.strip()])
    c = np.array([float(l.split()[0]) for l in
                  (prefix.parent / (prefix.name + "_C.txt")).read_text().splitlines() if l.strip()])
    assert b.size == m and c.size == n, prefix
    return A, b, c


def highs(A, b, c, sense=None, maximize=True):
    """-> (status 5/4/6, objective or None) for max/min c^T x, A x (<=,=,>=) b, x >= 0."""
    m = A.shape[0]
    s = np.zeros(m, np.int8) if sense is None else np.asarray(sense)
    ub, eq = s != 1, s == 1
    Aub = np.where((s[ub] == 2)[:, None], -A[ub], A[ub])
    bub = np.where(s[ub] == 2, -b[ub], b[ub])
    kw = dict(A_ub=Aub if ub.any() else None, b_ub=bub if ub.any() else None,
              A_eq=A[eq] if eq.any() else None, b_eq=b[eq] if eq.any() else None,
              bounds=(0, None))
    obj = -c if maximize else c
    for method in ("highs", "highs-ds", "highs-ipm"):
        r = linprog(obj, method=method, **kw)
        if r.status == 0:
            return 5, float(-r.fun if maximize else r.fun)
        if r.status == 3:
            return 6, None
        if r.status == 2:
            # HiGHS' presolve can say "infeasible" for "infeasible or unbounded":
            # a feasibility-only solve decides
            f = linprog(np.zeros_like(c), method="highs", **kw)
            return (6 if f.status == 0 else 4), None
    r = linprog(np.zeros_like(c), method="highs", **kw)
    if r.status == 2:
        return 4, None
    raise RuntimeError("HiGHS could not decide")


def data_hash(lps) -> str:
    h = hashlib.sha256()
    for lp in lps:
        for a in (lp.A, lp.b, lp.c):
            h.update(np.ascontiguousarray(a, dtype=np.float64).tobytes())
        if lp.sense is not None:
            h.update(np.ascontiguousarray(lp.sense, dtype=np.int8).tobytes())
    return h.hexdigest()


def make_netlib():
    arrays, meta = {}, {}
    for name in sorted(NETLIB_OPT):
        A, b, c = read_dense(REF_NETLIB / name)
        if name in BLPG8:  # the BLPG-Solver copy must be the same data
            A2, b2, c2 = read_dense(REF_BLPG / name)
            assert np.array_equal(A, A2) and np.array_equal(b, b2) and np.array_equal(c, c2), name
        r, cc = np.nonzero(A)
        key = name.replace("-", "_")
        arrays[f"{key}_shape"] = np.array(A.shape, np.int32)
        arrays[f"{key}_rows"] = r.astype(np.int32)
        arrays[f"{key}_cols"] = cc.astype(np.int32)
        arrays[f"{key}_vals"] = A[r, cc]
        arrays[f"{key}_b"] = b
        arrays[f"{key}_c"] = c
        st, obj = highs(A, b, c, maximize=False)
        meta[name] = {"key": key, "m": int(A.shape[0]), "n": int(A.shape[1]),
                      "negative_b": int((b < 0).sum()), "netlib_opt_min": NETLIB_OPT[name],
                      "highs_status": st, "highs_opt_min": obj, "blpg8": name in BLPG8}
        print(name, A.shape, st, obj, NETLIB_OPT[name], flush=True)
    np.savez_compressed(HERE / "netlib19.npz", **arrays)
    (HERE / "netlib19.json").write_text(json.dumps(meta, indent=1, sort_keys=True))


CONFIGS = {
    1: {"seed": 0, "batch": 1024, "shapes": [(27, 32)], "count": 64},
    2: {"seed": 2, "batch": 1024, "shapes": [(50, 48), (56, 97), (74, 83), (105, 103)], "count": 64},
    3: {"seed": 3, "batch": 8192, "shapes": [(174, 142), (205, 203)], "count": 24},
    4: {"seed": 4, "batch": 1024, "shapes": [(105, 103)], "count": 32, "families": [0, 1]},
}


def make_configs():
    import simplexprojects_amd as sa
    for cfg, spec in CONFIGS.items():
        entries = []
        groups = ([(f, spec["shapes"][0]) for f in spec["families"]] if cfg == 4
                  else list(enumerate(spec["shapes"])))
        for si, (m, n) in groups:
            lps = sa.gen_config(cfg, spec["seed"], m, n, spec["count"], shape_idx=si,
                                batch_count=spec["batch"])
            st, ob = [], []
            for lp in lps:
                s, o = highs(lp.A, lp.b, lp.c, lp.sense)
                st.append(s)
                ob.append(o)
            entries.append({"shape_idx": si, "m": m, "n": n, "first": 0,
                            "count": spec["count"], "batch_count": spec["batch"],
                            "sha256": data_hash(lps), "status": st, "objective": ob,
                            "flags": [int(lp.flags) for lp in lps]})
            print(cfg, m, n, {k: st.count(k) for k in set(st)}, flush=True)
        (HERE / f"highs_config{cfg}.json").write_text(
            json.dumps({"config": cfg, "seed": spec["seed"], "groups": entries}, indent=0))


TWO_PHASE = {"seed": 17, "shapes": [(174, 142), (205, 203)], "count": 12}


def make_two_phase():
    from tests.common import two_phase_feasible
    groups = []
    for si, (m, n) in enumerate(TWO_PHASE["shapes"]):
        lps = two_phase_feasible(m, n, TWO_PHASE["count"], TWO_PHASE["seed"] + si)
        h = hashlib.sha256()
        st, ob = [], []
        for A, b, c, s in lps:
            for a in (A, b, c):
                h.update(np.ascontiguousarray(a, dtype=np.float64).tobytes())
            h.update(s.tobytes())
            r = highs(A, b, c, s)
            st.append(r[0])
            ob.append(r[1])
        groups.append({"m": m, "n": n, "seed": TWO_PHASE["seed"] + si, "count": len(lps),
                       "sha256": h.hexdigest(), "status": st, "objective": ob})
        print("two-phase", m, n, {k: st.count(k) for k in set(st)}, flush=True)
    (HERE / "highs_two_phase.json").write_text(json.dumps({"groups": groups}, indent=0))


if __name__ == "__main__":
    # `make_golden.py two_phase` writes only highs_two_phase.json (no reference files read)
    if sys.argv[1:] == ["two_phase"]:
        make_two_phase()
    else:
        make_netlib()
        make_configs()
        make_two_phase()
