"""GPU tests of routing and measured autotuning: every kernel instance (each register-tile
variant, LDS-resident and HBM-streamed tableaux at each workgroup size) replays the oracle
bit for bit, so the autotuner's choice changes speed only; the tuning table round-trips
through its text form; invalid pins are rejected."""
import numpy as np
import pytest

import oracle
import simplexprojects_amd as sa

pytestmark = pytest.mark.gpu

INSTANCES = ([("wave", v) for v in range(5)] + [("tile", v) for v in range(12)] +
             [(k, nt) for k in ("lds", "stream") for nt in (64, 128, 256, 512)])
WAVE_FP64 = {0}  # wave variants built in FP64 (the others' columns exceed its registers)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("kernel,param", INSTANCES)
def test_every_instance_is_bit_exact(kernel, param, precision):
    if kernel == "wave" and precision == "fp64" and param not in WAVE_FP64:
        pytest.skip("wave variant not built in FP64")
    lps = sa.gen_config(2, 2, 30, 38, 12, shape_idx=0, batch_count=1024)
    S = sa.Solver(precision=precision, verify=False)
    S.tune_set(30, 38, kernel, param)
    b = S.upload(lps)
    b.solve()
    assert b.info()["bins"][0]["kernel"] == sa._KERNELS[kernel]
    r = b.fetch()
    for k, lp in enumerate(lps):
        if r.fallback[k]:  # FP32 result failed the FP64 check: re-solved in FP64
            assert r.status[k] == oracle.solve(lp.A, lp.b, lp.c)["status"]
            continue
        o = oracle.solve(lp.A, lp.b, lp.c, precision=precision)
        assert r.status[k] == o["status"] and r.pivots[k] == o["pivots"], (k, kernel, param)
        assert np.array_equal(r.basis[k], o["basis"])
        if precision == "fp64":  # (FP32 results report the FP64 re-solve's objective)
            assert r.objective[k] == o["objective"]
    b.free()
    S.close()


def test_autotune_table_and_results():
    lps = (sa.gen_config(2, 2, 50, 48, 256, shape_idx=0, batch_count=1024) +
           sa.gen_config(2, 2, 74, 83, 256, shape_idx=2, batch_count=1024))
    S = sa.Solver(precision="fp64")
    base = S.solve(lps)
    table = S.autotune(lps, reps=1)
    shapes = {(m, n) for (_, m, n, _, _, _) in table}
    assert shapes == {(50, 48), (74, 83)}
    assert all(ms > 0 for *_, ms in table)
    tuned = S.solve(lps)
    assert np.array_equal(base.pivots, tuned.pivots)
    assert np.array_equal(base.objective, tuned.objective)
    text = S.tune_export()
    S2 = sa.Solver(precision="fp64")
    S2.tune_import(text)
    assert S2.tune_export() == text
    with pytest.raises(sa.BlpgError):
        S2.tune_import("0 50 48 9 0 1.0\n")
    with pytest.raises(sa.BlpgError):
        S2.tune_set(200, 200, "tile", 0)  # variant 0 holds 30x38 at most
    S.close()
    S2.close()


@pytest.mark.parametrize("workload", ["mixed", "families"])
@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_check_plan_skip_first_tier_same_decisions(precision, workload):
    # A shape whose check plan skips the first LDS tier (as the autotuner chooses for final
    # bases that never fit it) hands its LPs straight to the next tier: every decision,
    # objective and KKT measure is bit-identical to the default plan (both tiers stage the
    # same blocks and run the same sweeps), and the table row round-trips.  "mixed": the
    # benchmark's 105x103 LPs (33-71 basic structurals: none fit the first tier, which only
    # defers them); "families": parametric-family LPs (8-18 basic structurals at the end;
    # the first tier holds up to 17, including the 16-17 that only fit since its int
    # vectors are 4-byte aligned), checked by the first tier under the default plan and by
    # the second under the other.
    if workload == "mixed":
        lps = sa.gen_config(2, 2, 105, 103, 512, shape_idx=3, batch_count=1024)
    else:
        lps = []
        for f in range(4):
            lps += sa.gen_config(4, 4, 105, 103, 128, shape_idx=f, first=512, batch_count=1024)
    S = sa.Solver(precision=precision, verify=True)
    base = S.solve(lps)
    if workload == "families":
        ks = np.array([int((np.asarray(bb) < 103).sum()) for bb in base.basis])
        assert ((ks >= 16) & (ks <= 17)).any(), np.bincount(ks)
    p = 1 if precision == "fp32" else 0
    S.tune_import(f"{p} 105 103 0 1 0.25\n")
    assert (p, 105, 103, 0, 1, 0.25) in S.tune_table()
    r = S.solve(lps)
    for f in ("status", "fallback", "pivots"):
        assert np.array_equal(getattr(base, f), getattr(r, f)), f
    for f in ("objective", "kkt_primal", "kkt_dual"):
        assert np.array_equal(getattr(base, f), getattr(r, f), equal_nan=True), f
    with pytest.raises(sa.BlpgError):
        S.tune_import(f"{p} 105 103 0 7 0.25\n")  # unknown check-plan flags
    S.close()
