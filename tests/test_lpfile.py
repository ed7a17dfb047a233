"""CPU tests of the CPLEX-LP writer and the glpsol report parser (the GLPK baseline runner
itself needs glpsol on PATH, which this environment does not have)."""
import re

import numpy as np

import oracle
import simplexprojects_amd as sa
from simplexprojects_amd.lpfile import parse_glpsol_report, write_cplex_lp
from tests.common import rel_err


def read_back(path, m, n):
    """An independent reader for the subset of CPLEX-LP the writer produces."""
    text = open(path).read()
    def terms(expr):
        v = np.zeros(n)
        for sign, coef, j in re.findall(r"([+-]?)\s*([0-9.eE+-]+)\s+x(\d+)", expr):
            v[int(j) - 1] += (-1.0 if sign == "-" else 1.0) * float(coef)
        return v
    c = terms(re.search(r"obj:(.*)", text).group(1))
    A = np.zeros((m, n)); b = np.zeros(m); s = np.zeros(m, np.int8)
    for i, (expr, op, rhs) in enumerate(re.findall(r"^ r\d+: (.*) (<=|>=|=) (\S+)$", text, re.M)):
        A[i] = terms(expr); b[i] = float(rhs); s[i] = {"<=": 0, "=": 1, ">=": 2}[op]
    return A, b, c, s


def test_round_trip_and_same_optimum(tmp_path):
    lps = sa.gen_config(2, 2, 50, 48, 6, shape_idx=0, batch_count=1024)
    rng = np.random.default_rng(3)
    A = rng.uniform(-1, 3, (12, 9)); x0 = rng.uniform(0, 1, 9)
    s = np.array([0, 1, 2] * 4, np.int8)
    b = A @ x0 + np.where(s == 0, 0.5, np.where(s == 2, -0.5, 0.0))
    A[-1] = 1.0; b[-1] = x0.sum() + 3; s[-1] = 0
    lps.append(sa.LP(A, b, rng.uniform(-1, 1, 9), s))
    for k, lp in enumerate(lps):
        p = str(tmp_path / f"lp{k}.lp")
        write_cplex_lp(lp, p)
        A2, b2, c2, s2 = read_back(p, lp.m, lp.n)
        assert np.array_equal(A2, lp.A) and np.array_equal(b2, lp.b) and np.array_equal(c2, lp.c)
        assert np.array_equal(s2, lp.sense if lp.sense is not None else np.zeros(lp.m, np.int8))
        o1 = oracle.solve(lp.A, lp.b, lp.c, lp.sense)
        o2 = oracle.solve(A2, b2, c2, s2)
        assert o1["status"] == o2["status"]
        if o1["status"] == oracle.OPTIMAL:
            assert rel_err(o1["objective"], o2["objective"]) <= 1e-12


def test_glpsol_report_parser():
    rpt = ("Problem:    \nRows:       3\nColumns:    2\nNon-zeros:  6\n"
           "Status:     OPTIMAL\nObjective:  obj = 21.33333333 (MAXimum)\n")
    assert parse_glpsol_report(rpt) == (5, 21.33333333)
    assert parse_glpsol_report(rpt.replace("OPTIMAL", "UNBOUNDED"))[0] == 6
    assert parse_glpsol_report(rpt.replace("OPTIMAL", "INFEASIBLE (FINAL)"))[0] == 4
    assert parse_glpsol_report("garbage")[0] == 1
