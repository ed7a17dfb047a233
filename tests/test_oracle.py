"""CPU tests that pin the oracle (oracle/) before it is trusted as the GPU checker.

Pins, each chosen so that a plausible mistake in the oracle fails one of them:
  - netlib known answers quoted in the reference (BLPG-Solver/src/parseBenchmark.cpp:17-25)
    and HiGHS on all 19 dense conversions it ships;
  - the reference's hand LPs (parseBenchmark.cpp:128-319);
  - committed HiGHS golden fixtures for the synthetic configurations (with a data hash);
  - brute-force vertex enumeration on tiny LPs;
  - an exact-rational replay of the same pivot rule (pivot sequence and exact optimum);
  - the reference kernel's own rule (restated in oracle/ref_rule.c) on the configuration
    where it coincides with the build's rule, and on the reference's recorded outcomes;
  - anti-cycling (Hall-McKinnon) and the iteration cap; the KKT checker at known optima
    and on planted errors.
"""
import numpy as np
import pytest

import oracle
import simplexprojects_amd as sa
from tests.common import (BEALE, HALL_MCKINNON, HAND_LPS, exact_simplex, highs, load_highs,
                          load_netlib, data_hash, parse5dim, parse28dim_helicopter, rel_err,
                          vertex_enum)

NETLIB = load_netlib()


@pytest.mark.parametrize("name", sorted(NETLIB))
def test_netlib19_matches_highs(name):
    A, b, c, mt = NETLIB[name]
    r = oracle.solve(A, b, c)
    assert r["status"] == mt["highs_status"]
    if mt["highs_opt_min"] is not None:
        assert rel_err(-r["objective"], mt["highs_opt_min"]) <= 1e-9


@pytest.mark.parametrize("name", ["AFIRO", "BLEND", "ISRAEL", "SC105", "SC205", "SC50A", "SC50B"])
def test_netlib_quoted_optimum(name):
    # parseBenchmark.cpp:17-25 quotes these optima (minimisation, 11 digits)
    A, b, c, mt = NETLIB[name]
    r = oracle.solve(A, b, c)
    assert r["status"] == oracle.OPTIMAL
    assert rel_err(-r["objective"], mt["netlib_opt_min"]) <= 1e-10


def test_adlittle_conversion_optimum():
    # the shipped dense ADLITTLE differs from netlib by conversion rounding (2.7e-6):
    # its own optimum is 2.2549435412e+05 (HiGHS), the netlib value 2.2549496316e+05
    A, b, c, mt = NETLIB["ADLITTLE"]
    r = oracle.solve(A, b, c)
    assert abs(-r["objective"] - 2.2549435412e05) < 1e-3
    assert rel_err(-r["objective"], mt["netlib_opt_min"]) > 1e-6


@pytest.mark.parametrize("name", sorted(HAND_LPS))
def test_hand_lps(name):
    A, b, c, want = HAND_LPS[name]
    A, b, c = np.array(A, float), np.array(b, float), np.array(c, float)
    r = oracle.solve(A, b, c)
    assert r["status"] == oracle.OPTIMAL
    assert abs(r["objective"] - want) < 1e-12
    st, obj = highs(A, b, c)
    assert st == 5 and abs(obj - want) < 1e-9


@pytest.mark.parametrize("maker", [parse5dim, parse28dim_helicopter])
def test_reachability_lps(maker):
    A, b, c, want = maker()
    r = oracle.solve(A, b, c)
    assert r["status"] == oracle.OPTIMAL and abs(r["objective"] - want) < 1e-12


@pytest.mark.parametrize("cfg", [1, 2, 3, 4])
def test_highs_golden(cfg):
    g = load_highs(cfg)
    for grp in g["groups"]:
        lps = sa.gen_config(cfg, g["seed"], grp["m"], grp["n"], grp["count"],
                            shape_idx=grp["shape_idx"], first=grp["first"],
                            batch_count=grp["batch_count"])
        assert data_hash(lps) == grp["sha256"], "generator changed"
        for k, lp in enumerate(lps):
            r = oracle.solve(lp.A, lp.b, lp.c, lp.sense)
            assert r["status"] == grp["status"][k], (cfg, grp["m"], k)
            if grp["objective"][k] is not None:
                assert rel_err(r["objective"], grp["objective"][k]) <= 1e-9


def test_vertex_enumeration_tiny():
    rng = np.random.default_rng(7)
    checked = 0
    for _ in range(300):
        m, n = rng.integers(1, 5), rng.integers(1, 5)
        A = rng.integers(-5, 6, size=(m, n)).astype(float)
        b = rng.integers(0, 8, size=m).astype(float)
        c = rng.integers(-4, 6, size=n).astype(float)
        r = oracle.solve(A, b, c)
        st, obj = highs(A, b, c)
        assert r["status"] == st
        if st == 5:
            best = vertex_enum(A, b, c)
            assert best is not None and abs(r["objective"] - best) < 1e-9
            checked += 1
    assert checked > 100


def test_exact_rational_replay():
    rng = np.random.default_rng(11)
    same_seq = 0
    for t in range(60):
        m, n = rng.integers(2, 7), rng.integers(2, 7)
        A = rng.integers(-20, 21, size=(m, n))
        b = rng.integers(0, 30, size=m)  # zeros make degenerate pivots
        c = rng.integers(-10, 15, size=n)
        st, obj, piv = exact_simplex(A.tolist(), b.tolist(), c.tolist())
        r = oracle.solve(A.astype(float), b.astype(float), c.astype(float))
        assert r["status"] == st, t
        if st == 5:
            assert abs(r["objective"] - float(obj)) <= 1e-9 * max(1.0, abs(float(obj)))
        same_seq += r["pivots"] == piv
    assert same_seq >= 57  # identical pivot sequences barring FP near-ties


def test_reference_rule_matches_on_config1():
    # With c > 0, b > 0 and no ties, the reference kernel's rule (simplex.cu:93-376) and the
    # build's rule take the same pivots.
    lps = sa.gen_config(1, 0, 27, 32, 64, batch_count=1024)
    for lp in lps:
        obj_ref, piv_ref = oracle.ref_rule_solve(lp.A, lp.b, lp.c)
        r = oracle.solve(lp.A, lp.b, lp.c)
        assert r["pivots"] == piv_ref
        assert rel_err(r["objective"], obj_ref) <= 1e-12


def test_reference_rule_recorded_outcomes():
    # CPLEXvsGLPK/SourceCode/src/parseBenchmark.cpp:18-38 records "-0 by both GLPK and GPU"
    # for KB2 and RECIPE and "No" for BLEND; the restated rule reproduces them.
    for name in ("KB2", "RECIPELP"):
        A, b, c, _ = NETLIB[name]
        obj, _ = oracle.ref_rule_solve(A, b, c, max_iter=20000)
        assert obj == 0.0
    A, b, c, mt = NETLIB["BLEND"]
    obj, _ = oracle.ref_rule_solve(A, b, c)
    assert rel_err(-obj, mt["netlib_opt_min"]) > 1e-3


def test_hall_mckinnon_cycles_without_bland():
    A, b, c = HALL_MCKINNON
    r = oracle.solve(A, b, c, bland_after=10**9, max_iter=300)
    assert r["status"] == oracle.ITER_LIMIT
    r = oracle.solve(A, b, c)  # Bland after 50 degenerate pivots
    assert r["status"] == oracle.UNBOUNDED == highs(A, b, c)[0]
    A2, b2 = np.vstack([A, np.ones(4)]), np.array([0, 0, 1.0])
    r = oracle.solve(A2, b2, c)
    assert r["status"] == oracle.OPTIMAL and abs(r["objective"] - 0.875) < 1e-12


def test_beale():
    A, b, c = BEALE
    for ba in (1, 5, 50):
        r = oracle.solve(A, b, c, bland_after=ba)
        assert r["status"] == oracle.OPTIMAL and abs(r["objective"] - 1.25) < 1e-12


def test_iteration_cap():
    lp = sa.gen_config(2, 2, 105, 103, 1, shape_idx=3, batch_count=1024)[0]
    r = oracle.solve(lp.A, lp.b, lp.c, max_iter=3)
    assert r["status"] == oracle.ITER_LIMIT and r["pivots"] == 3


def test_kkt_checker_pins():
    lps = sa.gen_config(2, 2, 50, 48, 40, shape_idx=0, batch_count=1024)
    n_opt = 0
    for lp in lps:
        r = oracle.solve(lp.A, lp.b, lp.c)
        if r["status"] != oracle.OPTIMAL:
            continue
        n_opt += 1
        k = oracle.kkt_check(lp.A, lp.b, lp.c, r["basis"])
        assert k["res_primal"] < 1e-12 and k["res_dual"] < 1e-12
        assert rel_err(k["objective"], r["objective"]) < 1e-10
        # planted mistake 1: raising a nonbasic variable's cost makes the basis dual-infeasible
        nonbasic = sorted(set(range(lp.n)) - set(r["basis"].tolist()))
        c2 = lp.c.copy()
        c2[nonbasic[0]] = float(k["y"] @ lp.A[:, nonbasic[0]]) + 1.0  # reduced cost +1
        k1 = oracle.kkt_check(lp.A, lp.b, c2, r["basis"])
        assert k1["res_dual"] > 1e-3 and k1["res_primal"] < 1e-12
        # planted mistake 2: replacing a basic structural by a nonbasic one breaks KKT
        basis = r["basis"].copy()
        i = int(np.argmax(basis < lp.n))
        nonbasic = sorted(set(range(lp.n)) - set(basis.tolist()))
        basis[i] = nonbasic[0]
        k2 = oracle.kkt_check(lp.A, lp.b, lp.c, basis)
        assert k2 is None or k2["res_primal"] > 1e-9 or k2["res_dual"] > 1e-9
    assert n_opt >= 10


def test_warm_start_from_optimal_basis():
    for cfg, (m, n), si in [(1, (27, 32), 0), (2, (105, 103), 3), (3, (174, 142), 0)]:
        lps = sa.gen_config(cfg, cfg if cfg > 1 else 0, m, n, 6, shape_idx=si,
                            batch_count=8192 if cfg == 3 else 1024)
        for lp in lps:
            r = oracle.solve(lp.A, lp.b, lp.c, lp.sense)
            w = oracle.solve(lp.A, lp.b, lp.c, lp.sense, warm=r["basis"])
            assert w["status"] == r["status"]
            if r["status"] == oracle.OPTIMAL:
                assert rel_err(w["objective"], r["objective"]) < 1e-9
                assert w["pivots"] <= r["pivots"]


def test_fp32_oracle_mostly_agrees():
    lps = sa.gen_config(2, 2, 74, 83, 64, shape_idx=2, batch_count=1024)
    agree = sum(oracle.solve(lp.A, lp.b, lp.c, precision="fp32")["status"] ==
                oracle.solve(lp.A, lp.b, lp.c)["status"] for lp in lps)
    assert agree >= 60


def test_generator_statistics_and_subsets():
    lps = sa.gen_config(2, 2, 50, 48, 1024, shape_idx=0, batch_count=1024)
    A = np.stack([lp.A for lp in lps])
    zero = float(np.mean(A == 0))
    assert 0.38 < zero < 0.42  # 40% zeroed
    assert sum(lp.flags & 1 for lp in lps) == 51  # floor(0.05 * 1024) planted unbounded
    assert all((lp.b == 0).sum() >= 5 for lp in lps)  # floor(0.1 m) zero rows
    sub = sa.gen_config(2, 2, 50, 48, 10, shape_idx=0, first=500, batch_count=1024)
    for k in range(10):
        assert np.array_equal(sub[k].A, lps[500 + k].A) and np.array_equal(sub[k].c, lps[500 + k].c)
    l3 = sa.gen_config(3, 3, 174, 142, 200, shape_idx=0, batch_count=8192)
    assert all(int((lp.sense == 1).sum()) in (52, 51) for lp in l3 if not lp.flags & 2)


def test_dual_repair_after_rhs_change():
    # the optimal basis of one instance, crashed into an instance with a changed b, stays
    # dual feasible; dual simplex pivots restore primal feasibility (HiGHS pins the answer)
    rng = np.random.default_rng(5)
    lps = sa.gen_config(4, 4, 105, 103, 12, shape_idx=1, batch_count=1024)
    dual_used = 0
    for lp in lps:
        r0 = oracle.solve(lp.A, lp.b, lp.c)
        b2 = lp.b * (1 + 0.3 * rng.standard_normal(lp.m))
        b2 = np.maximum(b2, 0.1)
        w = oracle.solve(lp.A, b2, lp.c, warm=r0["basis"])
        st, obj = highs(lp.A, b2, lp.c)
        assert w["status"] == st == oracle.OPTIMAL
        assert rel_err(w["objective"], obj) <= 1e-9
        crash = int((r0["basis"] < lp.n).sum())
        dual_used += w["phase1_pivots"] > crash
    assert dual_used >= 3


def test_two_phase_feasible_matches_highs():
    """Planted-feasible two-phase LPs (mixed senses, mixed-sign b, 174x142 and 205x203): the
    oracle's FP64 replay runs phase I and ends at HiGHS's optimum on every one."""
    from tests.common import load_two_phase
    for grp, lps in load_two_phase():
        for k, (A, b, c, s) in enumerate(lps):
            r = oracle.solve(A, b, c, s)
            assert grp["status"][k] == oracle.OPTIMAL and r["status"] == oracle.OPTIMAL, k
            assert rel_err(r["objective"], grp["objective"][k]) <= 1e-9, k
            assert r["phase1_pivots"] > 0, k
