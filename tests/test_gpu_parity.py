"""GPU parity: the HIP kernels (through the C ABI) against the CPU oracle and the fixtures.

  - bit-exact: status, pivot count, basis and tableau objective equal the oracle's replay
    of the same rule, for every kernel (register tile, LDS, HBM-streamed) and precision;
  - golden: status equal to HiGHS on every fixture LP, objective within 1e-9 relative
    (FP64) or 1e-6 after the FP64 check (FP32);
  - netlib: the 19 dense conversions the reference ships;
  - full-size properties of the benchmark batch and the two-phase batch: every LP decided,
    FP32 and FP64 statuses agree, verified objectives agree to 1e-6, KKT residuals small,
    runs are deterministic.
"""
import numpy as np
import pytest

import oracle
import simplexprojects_amd as sa
from tests.common import load_highs, load_netlib, rel_err

pytestmark = pytest.mark.gpu

SHAPES = {1: [(27, 32)], 2: [(50, 48), (56, 97), (74, 83), (105, 103)],
          3: [(174, 142), (205, 203)]}
SEEDS = {1: 0, 2: 2, 3: 3}
BATCH = {1: 1024, 2: 1024, 3: 8192}


def sample(cfg, count):
    lps = []
    for si, (m, n) in enumerate(SHAPES[cfg]):
        lps += sa.gen_config(cfg, SEEDS[cfg], m, n, count, shape_idx=si,
                             batch_count=BATCH[cfg])
    return lps


@pytest.mark.parametrize("kernel", ["wave", "tile", "lds", "stream"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("cfg", [1, 2, 3])
def test_bit_exact_vs_oracle(cfg, precision, kernel):
    lps = sample(cfg, 12 if cfg == 3 else 24)
    S = sa.Solver(precision=precision, kernel=kernel, verify=(precision == "fp32"))
    r = S.solve(lps)
    for k, lp in enumerate(lps):
        if precision == "fp32" and r.fallback[k]:
            # the FP32 result failed the FP64 check and was re-solved in FP64
            assert r.status[k] == oracle.solve(lp.A, lp.b, lp.c, lp.sense)["status"]
            continue
        o = oracle.solve(lp.A, lp.b, lp.c, lp.sense, precision=precision)
        assert r.pivots[k] == o["pivots"], (k, kernel, precision)
        assert r.phase1_pivots[k] == o["phase1_pivots"]
        assert np.array_equal(r.basis[k], o["basis"])
        if precision == "fp64":
            assert r.status[k] == o["status"]
            assert r.objective[k] == o["objective"]  # bit-identical tableau value
        elif not r.fallback[k]:
            assert r.status[k] == o["status"]
    S.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("cfg", [1, 2, 3, 4])
def test_highs_golden(cfg, precision):
    g = load_highs(cfg)
    lps, want_s, want_o = [], [], []
    for grp in g["groups"]:
        lps += sa.gen_config(cfg, g["seed"], grp["m"], grp["n"], grp["count"],
                             shape_idx=grp["shape_idx"], batch_count=grp["batch_count"])
        want_s += grp["status"]
        want_o += grp["objective"]
    S = sa.Solver(precision=precision)
    r = S.solve(lps)
    tol = 1e-9 if precision == "fp64" else 1e-6
    for k in range(len(lps)):
        assert r.status[k] == want_s[k], (cfg, k)
        if want_o[k] is not None:
            assert rel_err(r.objective[k], want_o[k]) <= tol, (cfg, k)
    S.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_netlib19(precision):
    nl = load_netlib()
    names = sorted(nl)
    lps = [sa.LP(nl[n][0], nl[n][1], nl[n][2]) for n in names]
    S = sa.Solver(precision=precision)
    r = S.solve(lps)
    for k, name in enumerate(names):
        mt = nl[name][3]
        assert r.status[k] == mt["highs_status"], name
        if mt["highs_opt_min"] is not None:
            assert rel_err(-r.objective[k], mt["highs_opt_min"]) <= 1e-9, name
    S.close()


def test_benchmark_batch_properties():
    lps = sample(2, 1024)
    S64 = sa.Solver(precision="fp64")
    S32 = sa.Solver(precision="fp32")
    r64 = S64.solve(lps)
    r32 = S32.solve(lps)
    decided = np.isin(r64.status, [sa.OPTIMAL, sa.UNBOUNDED, sa.INFEASIBLE])
    assert decided.all()
    assert np.array_equal(r64.status, r32.status)
    opt = r64.status == sa.OPTIMAL
    rel = np.abs(r32.objective[opt] - r64.objective[opt]) / np.maximum(1, np.abs(r64.objective[opt]))
    assert rel.max() <= 1e-6
    assert (r64.kkt_primal[opt] <= 1e-9).all() and (r64.kkt_dual[opt] <= 1e-9).all()
    assert r32.fallback.sum() < 0.01 * len(lps)
    # planted unbounded columns are unbounded
    flags = np.array([lp.flags for lp in lps])
    assert (r64.status[(flags & 1) == 1] == sa.UNBOUNDED).all()
    # determinism: a second run is identical
    r64b = S64.solve(lps)
    assert np.array_equal(r64.pivots, r64b.pivots) and np.array_equal(r64.objective, r64b.objective)
    S64.close()
    S32.close()


def test_two_phase_batch_properties():
    lps = sample(3, 256)
    S = sa.Solver(precision="fp64")
    r = S.solve(lps)
    assert np.isin(r.status, [sa.OPTIMAL, sa.UNBOUNDED, sa.INFEASIBLE]).all()
    flags = np.array([lp.flags for lp in lps])
    assert (r.status[(flags & 2) == 2] == sa.INFEASIBLE).all()  # planted contradictions
    assert (r.phase1_pivots > 0).all()
    r32 = sa.Solver(precision="fp32").solve(lps)
    assert np.array_equal(r.status, r32.status)
    S.close()


def test_warm_start_matches_oracle():
    lps = sample(2, 16)
    bases = [oracle.solve(lp.A, lp.b, lp.c, precision="fp32")["basis"] for lp in lps]
    wl = [sa.LP(lp.A, lp.b, lp.c, warm_basis=bs) for lp, bs in zip(lps, bases)]
    for precision in ("fp64", "fp32"):
        S = sa.Solver(precision=precision, kernel="tile", verify=False)
        r = S.solve(wl)
        for k, lp in enumerate(lps):
            o = oracle.solve(lp.A, lp.b, lp.c, precision=precision, warm=bases[k])
            assert r.pivots[k] == o["pivots"] and r.status[k] == o["status"]
            assert np.array_equal(r.basis[k], o["basis"])
        S.close()


def scrambled(basis, n, m, rng):
    """A warm list the crash must read literally: shuffled, a few entries dropped (-1),
    repeated, or swapped for other structurals / slacks / x0."""
    w = rng.permutation(np.asarray(basis, dtype=np.int32))
    for i in range(m):
        u = rng.random()
        if u < 0.08:
            w[i] = -1
        elif u < 0.16:
            w[i] = w[rng.integers(0, m)]
        elif u < 0.24:
            w[i] = rng.integers(0, n)
        elif u < 0.28:
            w[i] = rng.integers(n, n + m + 1)
    return w


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_block_crash_matches_oracle(precision):
    # the tile kernel crashes up to a block of targets at once; every pivot must be the
    # one-at-a-time crash's (same rows, same tableau bits).  Cap 30: the first block fits
    # under the cap, the rest of the crash runs one target at a time.
    from tests.test_verify_ref import mixed_sense_lps
    rng = np.random.default_rng(17)
    lps = sample(2, 6) + [sa.LP(A, b, c, s) for A, b, c, s in mixed_sense_lps(8, m=40, n=34, seed=9)]
    cases = []
    for lp in lps:
        o64 = oracle.solve(lp.A, lp.b, lp.c, lp.sense)
        bs = o64["basis"]
        m, n = lp.A.shape
        cases.append((lp, bs, None, o64))                        # the optimal basis
        cases.append((lp, scrambled(bs, n, m, rng), None, o64))  # read as given
        cases.append((lp, bs, 30, o64))                          # cap inside the crash
    # several blocks: more basic structurals than any block holds
    assert max(int(((c[1] >= 0) & (c[1] < c[0].A.shape[1])).sum()) for c in cases) > 16
    for cap in (None, 30):
        sel = [c for c in cases if c[2] == cap]
        wl = [sa.LP(lp.A, lp.b, lp.c, lp.sense, warm_basis=w) for lp, w, _, _ in sel]
        S = sa.Solver(precision=precision, kernel="tile", verify=False, max_iter=cap or 0)
        r = S.solve(wl)
        exact = 0
        for k, (lp, w, _, o64) in enumerate(sel):
            if r.fallback[k]:  # an FP32 result the FP64 check sent to an FP64 re-solve
                continue
            o = oracle.solve(lp.A, lp.b, lp.c, lp.sense, precision=precision, warm=w,
                             max_iter=cap)
            assert r.status[k] == o["status"], (k, cap)
            assert r.pivots[k] == o["pivots"] and r.phase1_pivots[k] == o["phase1_pivots"], (k, cap)
            assert np.array_equal(r.basis[k], o["basis"]), (k, cap)
            if precision == "fp64":
                assert r.objective[k] == o["objective"], (k, cap)  # tableau bits
                assert np.array_equal(r.x[k], o["x"]), (k, cap)
            elif o["status"] == oracle.OPTIMAL:
                # FP32 reports the FP64 check's objective of the same basis
                assert rel_err(r.objective[k], o64["objective"]) <= 1e-9, (k, cap)
            exact += 1
        assert exact >= len(sel) // 3, (cap, exact)
        S.close()


def test_shared_A_matches_oracle():
    # the reference's batch shape (one A, b; many c): every LP bit-identical to the oracle
    lp0 = sa.gen_config(1, 0, 27, 32, 1, batch_count=1024)[0]
    rng = np.random.default_rng(3)
    Cm = rng.uniform(0, 1, size=(64, 32))
    S = sa.Solver(verify=False)
    rs = S.solve_shared_A(lp0.A, lp0.b, Cm)
    for k in range(64):
        o = oracle.solve(lp0.A, lp0.b, Cm[k])
        assert rs.status[k] == o["status"] and rs.pivots[k] == o["pivots"], k
        assert rs.objective[k] == o["objective"], k
        assert np.array_equal(rs.basis[k], o["basis"]), k
    S.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_two_phase_feasible_golden(precision):
    """Planted-feasible two-phase LPs at 174x142 and 205x203 (phase I, then phase II to the
    optimum): status and objective equal HiGHS's (1e-9 relative FP64, 1e-6 after the FP64
    check in FP32); FP64 runs take the oracle's pivots exactly, FP32 runs the oracle's FP32
    replay's pivots unless re-solved by the fallback."""
    from tests.common import load_two_phase
    groups = load_two_phase()
    lps = [sa.LP(A, b, c, s) for _, grp_lps in groups for (A, b, c, s) in grp_lps]
    status = np.concatenate([grp["status"] for grp, _ in groups])
    obj = np.concatenate([grp["objective"] for grp, _ in groups])
    S = sa.Solver(precision=precision)
    r = S.solve(lps)
    assert np.array_equal(r.status, status)
    err = np.abs(r.objective - obj) / np.maximum(1.0, np.abs(obj))
    assert err.max() <= (1e-9 if precision == "fp64" else 1e-6)
    assert (r.phase1_pivots > 0).all()
    for k, lp in enumerate(lps):
        o = oracle.solve(lp.A, lp.b, lp.c, lp.sense, precision=precision)
        assert r.pivots[k] == o["pivots"] or (precision == "fp32" and r.fallback[k]), k
    S.close()


def test_shared_A_sibling_warm_matches_oracle():
    """One A, b and senses (a planted-feasible two-phase LP, 174x142) with 40 objectives:
    copies of one objective, as the reference's own shared-A batch (GPU_Tester.cu:119-127),
    then that objective perturbed by 5% per entry.  With sibling_warm, LP 0 is solved cold
    and LPs 1.. are crashed from its final basis, primal feasible for all of them.  Every
    LP is bit-identical to the oracle's replay with the same warm list (status, pivots,
    objective, basis), ends at the cold solve's optimum, and skips phase I: about a fifth of
    the cold pivots."""
    from tests.common import load_two_phase
    grp, lps = load_two_phase()[0]
    A, b, c0, s = lps[0]
    rng = np.random.default_rng(5)
    Cm = np.tile(c0, (40, 1))
    Cm[8:] *= 1.0 + 0.05 * rng.normal(size=(32, len(c0)))
    S = sa.Solver(verify=False, sibling_warm=True)
    r = S.solve_shared_A(A, b, Cm, s)
    o0 = oracle.solve(A, b, Cm[0], s)
    assert o0["status"] == oracle.OPTIMAL
    assert r.status[0] == o0["status"] and r.pivots[0] == o0["pivots"]
    assert np.array_equal(r.basis[0], o0["basis"])
    warm_piv, cold_piv = [], []
    for k in range(1, len(Cm)):
        o = oracle.solve(A, b, Cm[k], s, warm=o0["basis"])
        assert r.status[k] == o["status"] and r.pivots[k] == o["pivots"], k
        assert r.objective[k] == o["objective"], k
        assert np.array_equal(r.basis[k], o["basis"]), k
        oc = oracle.solve(A, b, Cm[k], s)
        assert oc["status"] == o["status"] == oracle.OPTIMAL, k
        assert rel_err(o["objective"], oc["objective"]) <= 1e-9, k
        assert r.phase1_pivots[k] < oc["phase1_pivots"], k
        warm_piv.append(o["pivots"])
        cold_piv.append(oc["pivots"])
    assert np.mean(warm_piv) < 0.35 * np.mean(cold_piv), (np.mean(warm_piv), np.mean(cold_piv))
    S.close()
