"""GPU edge cases of the drop-in path: empty and tiny batches, degenerate / cycling LPs,
equality handling, the iteration cap, invalid input, sizes beyond the register tile, and
the C++ drop-in program."""
import os
import subprocess

import numpy as np
import pytest

import oracle
import simplexprojects_amd as sa
from tests.common import HALL_MCKINNON, HAND_LPS

pytestmark = pytest.mark.gpu


def test_empty_batch():
    S = sa.Solver()
    r = S.solve([])
    assert len(r) == 0
    S.close()


def test_one_by_one():
    S = sa.Solver()
    r = S.solve([sa.LP([[2.0]], [4.0], [3.0]), sa.LP([[1.0]], [1.0], [-1.0]),
                 sa.LP([[-1.0]], [1.0], [1.0])])
    assert list(r.status) == [sa.OPTIMAL, sa.OPTIMAL, sa.UNBOUNDED]
    assert r.objective[0] == 6.0 and r.objective[1] == 0.0
    S.close()


@pytest.mark.parametrize("kernel", ["tile", "stream"])
def test_hand_lps(kernel):
    S = sa.Solver(kernel=kernel)
    names = sorted(HAND_LPS)
    r = S.solve([sa.LP(HAND_LPS[n][0], HAND_LPS[n][1], HAND_LPS[n][2]) for n in names])
    for k, n in enumerate(names):
        assert r.status[k] == sa.OPTIMAL and abs(r.objective[k] - HAND_LPS[n][3]) < 1e-12
    S.close()


def test_cycling_needs_bland():
    A, b, c = HALL_MCKINNON
    S = sa.Solver(bland_after=10**9, max_iter=300)
    assert S.solve([sa.LP(A, b, c)]).status[0] == sa.ITER_LIMIT
    S.close()
    S = sa.Solver()
    r = S.solve([sa.LP(A, b, c), sa.LP(np.vstack([A, np.ones(4)]), [0, 0, 1.0], c)])
    assert r.status[0] == sa.UNBOUNDED and r.status[1] == sa.OPTIMAL
    assert abs(r.objective[1] - 0.875) < 1e-12
    S.close()


def test_equalities_redundant_and_contradictory():
    # x1 + x2 = 2 twice (redundant), max x1: 2
    A = np.array([[1.0, 1.0], [1.0, 1.0]])
    lp1 = sa.LP(A, [2.0, 2.0], [1.0, 0.0], sense=[1, 1])
    # x1 + x2 = 2 and x1 + x2 = 3: infeasible
    lp2 = sa.LP(A, [2.0, 3.0], [1.0, 0.0], sense=[1, 1])
    # all-zero equality row with b != 0: infeasible in stage 1
    lp3 = sa.LP(np.array([[0.0, 0.0], [1.0, 1.0]]), [1.0, 5.0], [1.0, 1.0], sense=[1, 0])
    # '>=' row: x1 + x2 >= 3, x1 <= 1, min x2 (max -x2): x2 = 2
    lp4 = sa.LP(np.array([[1.0, 1.0], [1.0, 0.0]]), [3.0, 1.0], [0.0, -1.0], sense=[2, 0])
    S = sa.Solver()
    r = S.solve([lp1, lp2, lp3, lp4])
    assert r.status[0] == sa.OPTIMAL and abs(r.objective[0] - 2) < 1e-12
    assert r.status[1] == sa.INFEASIBLE and r.status[2] == sa.INFEASIBLE
    assert r.status[3] == sa.OPTIMAL and abs(r.objective[3] + 2) < 1e-12
    for k, lp in enumerate([lp1, lp2, lp3, lp4]):
        assert r.status[k] == oracle.solve(lp.A, lp.b, lp.c, lp.sense)["status"]
    S.close()


def test_iteration_cap():
    lp = sa.gen_config(2, 2, 105, 103, 1, shape_idx=3, batch_count=1024)[0]
    S = sa.Solver(max_iter=3)
    r = S.solve([lp])
    assert r.status[0] == sa.ITER_LIMIT and r.pivots[0] == 3
    S.close()


def test_invalid_input_rejected():
    S = sa.Solver()
    A = np.ones((2, 2))
    A[0, 1] = np.nan
    with pytest.raises(sa.BlpgError) as e:
        S.solve([sa.LP(A, [1, 1], [1, 1])])
    assert e.value.code == sa.E_INVALID
    with pytest.raises(sa.BlpgError):
        S.solve([sa.LP(np.ones((1, 1)), [1.0], [1.0], sense=[7])])
    S.close()


def test_large_lp_uses_streamed_kernel():
    # 300 x 280 exceeds every register-tile variant and the LDS: streamed from HBM
    rng = np.random.default_rng(5)
    lps = [sa.LP(rng.uniform(0, 1, (300, 280)) * (rng.uniform(size=(300, 280)) < 0.3),
                 rng.uniform(1, 10, 300), rng.uniform(0, 1, 280)) for _ in range(4)]
    S = sa.Solver(verify=False)
    b = S.upload(lps)
    b.solve()
    b.sync()
    assert b.info()["bins"][0]["kernel"] == sa.KERNEL_STREAM
    r = b.download()
    for k, lp in enumerate(lps):
        o = oracle.solve(lp.A, lp.b, lp.c)
        assert r.pivots[k] == o["pivots"] and r.objective[k] == o["objective"]
    S.close()


def test_hbm_budget_rejects():
    S = sa.Solver(hbm_budget=1024)
    with pytest.raises(sa.BlpgError) as e:
        S.solve(sa.gen_config(1, 0, 27, 32, 4))
    assert e.value.code == sa.E_BUDGET
    S.close()


def test_cpp_drop_in_program():
    exe = os.path.join(os.path.dirname(sa.LIB_PATH), "test_shim")
    assert os.path.exists(exe), "built by simplexprojects_amd.build"
    r = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=120)
    assert r.returncode == 0 and "ALL OK" in r.stdout, r.stdout


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_fetch_matches_download(precision):
    lps = sa.gen_config(2, 2, 56, 97, 64, shape_idx=1, batch_count=1024)
    S = sa.Solver(precision=precision)
    b = S.upload(lps)
    b.solve()
    d = b.download()
    f = b.fetch()
    for name in ("status", "pivots", "phase1_pivots", "fallback", "objective", "kkt_primal",
                 "kkt_dual"):
        assert np.array_equal(getattr(d, name), getattr(f, name)), name
    for k in range(len(lps)):
        assert np.array_equal(d.x[k], f.x[k]) and np.array_equal(d.basis[k], f.basis[k])
        if d.status[k] == sa.OPTIMAL:  # verified x is the FP64 basic solution
            assert abs(float(lps[k].c @ f.x[k]) - f.objective[k]) <= 1e-9 * max(1, abs(f.objective[k]))
            assert (lps[k].A @ f.x[k] <= lps[k].b + 1e-9).all() and (f.x[k] >= 0).all()
    b.free()
    S.close()


def test_batch_beyond_budget_is_solved_in_parts():
    # out-of-core batching: a batch over the HBM budget is split and solved part by part;
    # every part's results are the oracle's, bit for bit
    lps = sa.gen_config(2, 2, 50, 48, 600, shape_idx=0, batch_count=1024)
    ref = [oracle.solve(lp.A, lp.b, lp.c) for lp in lps]
    # ~42 KB of device memory per LP of this shape: 6 MB holds about a quarter of them
    S2 = sa.Solver(precision="fp64", verify=False, hbm_budget=6 << 20)
    r = S2.solve(lps)
    for k in range(len(lps)):
        assert r.status[k] == ref[k]["status"], k
        assert r.pivots[k] == ref[k]["pivots"], k
        assert r.objective[k] == ref[k]["objective"], k
        assert np.array_equal(r.basis[k], ref[k]["basis"]), k
    with pytest.raises(sa.BlpgError):
        S2.upload(lps)  # one device batch of all of them still exceeds the budget
    S2.close()


def test_prefix_arrays_of_one_buffer_are_distinct():
    # two LPs whose b (and A) are prefixes of one buffer with different lengths: each is
    # its own array in the device pool (the pool keys arrays by pointer and length)
    rng = np.random.default_rng(11)
    big_A = rng.uniform(-1, 1, size=(12 * 8,))
    big_b = rng.uniform(1, 5, size=12)
    c = rng.uniform(0, 1, size=8)
    lps = []
    for m in (6, 12, 9):
        lp = sa.LP(np.zeros((m, 8)), np.zeros(m), c)
        lp.A = big_A[: m * 8].reshape(m, 8)  # views: same base pointer, different lengths
        lp.b = big_b[:m]
        lps.append(lp)
    assert lps[0].A.ctypes.data == lps[1].A.ctypes.data == lps[2].A.ctypes.data
    S = sa.Solver(precision="fp64", verify=False)
    r = S.solve(lps)
    for k, lp in enumerate(lps):
        o = oracle.solve(lp.A, lp.b, lp.c)
        assert r.status[k] == o["status"] and r.pivots[k] == o["pivots"], k
        assert r.objective[k] == o["objective"], k
    S.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_repack_copies_new_inputs(precision):
    # a repacked batch solves the new LPs (the copy to the device runs inside the solve,
    # bin by bin), and a later solve without repack keeps them
    a = sa.gen_config(2, 2, 50, 48, 48, shape_idx=0, batch_count=1024)
    a += sa.gen_config(2, 2, 74, 83, 48, shape_idx=2, batch_count=1024)
    b = sa.gen_config(2, 7, 50, 48, 48, shape_idx=0, batch_count=1024)
    b += sa.gen_config(2, 7, 74, 83, 48, shape_idx=2, batch_count=1024)
    S = sa.Solver(precision=precision)
    batch = S.upload(a)
    batch.solve()
    ra = batch.fetch()
    batch.repack(b)
    batch.solve()
    rb = batch.fetch()
    batch.solve()
    rb2 = batch.fetch()
    direct = S.solve(b)
    for r in (rb, rb2):
        assert np.array_equal(r.status, direct.status)
        assert np.array_equal(r.pivots, direct.pivots)
        assert np.array_equal(r.objective, direct.objective)
    assert not np.array_equal(ra.objective, rb.objective)
    with pytest.raises(sa.BlpgError):
        batch.repack(b[:10])  # a different LP count
    batch.free()
    S.close()


def test_repack_packed_copy_bit_exact():
    """After a repack the copy sends each pool range packed (bitmask of the entries that
    are not +0.0, the other values packed) when that is smaller: entries of -0.0 are
    values and stay -0.0, a bin without zeros goes dense; every result equals the solve of
    a fresh upload of the same LPs bit for bit."""
    rng = np.random.default_rng(9)
    base = sa.gen_config(2, 2, 50, 48, 32, shape_idx=0, batch_count=1024)
    new = []
    for lp in sa.gen_config(2, 8, 50, 48, 32, shape_idx=0, batch_count=1024):
        A = lp.A.copy()
        A[(A == 0) & (rng.random(A.shape) < 0.5)] = -0.0
        new.append(sa.LP(A, lp.b.copy(), lp.c.copy()))
    assert any(np.signbit(lp.A[lp.A == 0]).any() for lp in new)
    dense = [sa.LP(rng.uniform(0.5, 1.0, (74, 83)), rng.uniform(1, 10, 74),
                   rng.uniform(0, 1, 83)) for _ in range(16)]
    for precision in ("fp64", "fp32"):
        S = sa.Solver(precision=precision)
        batch = S.upload(base + dense)
        batch.solve()
        batch.fetch()
        batch.repack(new + dense)
        batch.solve()
        r = batch.fetch()
        direct = S.solve(new + dense)
        assert np.array_equal(r.status, direct.status)
        assert np.array_equal(r.pivots, direct.pivots)
        assert np.array_equal(r.objective, direct.objective)
        for k in range(len(new) + len(dense)):
            assert np.array_equal(r.x[k], direct.x[k]), k
        batch.free()
        S.close()
