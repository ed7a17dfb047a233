"""GPU tests of the learned warm start (BF16-MFMA network + crash/repair), parametric
family batches (A_f 105x103 fixed, b and c scaled by 1 + 0.05 N(0,1)):

  - forward logits and the top-m crash lists against the CPU reference with the kernels'
    bf16 rounding points (tests/mlp_ref.py, itself pinned by finite differences);
  - one backward pass against the same reference (gradient check) and against the exact
    float64 gradient within the bf16 bound;
  - training lowers the loss; on unseen instances the warm-started FP32 solve reaches the
    cold FP64 status and objective (<= 1e-6 relative) with fewer pivots, and its pivots
    and basis equal the oracle's replay of the same crash list.
"""
import numpy as np
import pytest

import oracle
import simplexprojects_amd as sa
from tests.mlp_ref import features, forward_backward, labels

pytestmark = pytest.mark.gpu

M_, N_ = 105, 103
H_ = 256


def family(f, first, count):
    return sa.gen_config(4, 4, M_, N_, count, shape_idx=f, first=first, batch_count=1024)


def cold_bases(S, lps):
    r = S.solve(lps)
    assert (r.status == sa.OPTIMAL).all()
    return r, [b.copy() for b in r.basis]


def test_forward_and_pick_match_cpu():
    S = sa.Solver()
    ws = sa.WarmStart(S, 2, M_, N_, seed=1)
    lps = family(0, 0, 70) + family(1, 0, 50)
    _, bases = cold_bases(S, lps)
    ws.gradient(lps, bases)  # fits the input normalisation
    logits, warm = ws.predict(lps)
    for f, lo, hi in ((0, 0, 70), (1, 70, 120)):
        theta = ws.params(f).astype(np.float64)
        mu, isd = ws.norm(f)
        grp = lps[lo:hi]
        x = features(np.stack([lp.b for lp in grp]), np.stack([lp.c for lp in grp]), mu, isd)
        Y = labels(bases[lo:hi], M_ + N_)
        _, _, Z = forward_backward(theta, x, Y, H_, bf16_emulation=True)
        gz = logits[lo:hi]
        assert np.linalg.norm(gz - Z) <= 1e-3 * np.linalg.norm(Z), f
        for k in range(hi - lo):
            want = np.argsort(-gz[k], kind="stable")[:M_]
            assert np.array_equal(warm[lo + k], want)
    ws.close()
    S.close()


def test_gradient_matches_cpu():
    S = sa.Solver()
    ws = sa.WarmStart(S, 2, M_, N_, seed=2)
    lps = family(0, 0, 64) + family(1, 0, 40)
    _, bases = cold_bases(S, lps)
    g, loss = ws.gradient(lps, bases)
    for f, lo, hi in ((0, 0, 64), (1, 64, 104)):
        theta = ws.params(f).astype(np.float64)
        mu, isd = ws.norm(f)
        grp = lps[lo:hi]
        x = features(np.stack([lp.b for lp in grp]), np.stack([lp.c for lp in grp]), mu, isd)
        Y = labels(bases[lo:hi], M_ + N_)
        l_em, g_em, _ = forward_backward(theta, x, Y, H_, bf16_emulation=True)
        l_ex, g_ex, _ = forward_backward(theta, x.astype(np.float64), Y, H_, bf16_emulation=False)
        assert abs(loss[f] - l_em) <= 1e-4 * l_em
        assert np.linalg.norm(g[f] - g_em) <= 5e-3 * np.linalg.norm(g_em), f
        assert np.linalg.norm(g[f] - g_ex) <= 5e-2 * np.linalg.norm(g_ex), f
        # every parameter block carries gradient (no block silently dropped)
        pos = 0
        for name, shape in ws.layout():
            k = int(np.prod(shape))
            blk, ref = g[f][pos:pos + k], g_em[pos:pos + k]
            assert np.linalg.norm(blk - ref) <= 1e-2 * np.linalg.norm(ref) + 1e-12, name
            pos += k
    ws.close()
    S.close()


def test_training_and_warm_start_gain():
    S64 = sa.Solver(precision="fp64")
    nf = 3
    train = sum((family(f, 0, 512) for f in range(nf)), [])
    _, bases = cold_bases(S64, train)
    ws = sa.WarmStart(S64, nf, M_, N_, seed=3, steps=300)
    loss = ws.train(train, bases)
    assert loss[-1] < 0.2 * loss[0]
    assert np.all(np.isfinite(loss))

    test = sum((family(f, 512, 64) for f in range(nf)), [])
    cold = S64.solve(test)
    S32 = sa.Solver(precision="fp32", use_graph=False)
    batch = S32.upload(test)
    batch.set_warmstart(ws)
    batch.solve()
    batch.sync()
    warm = batch.download()
    info = batch.info()
    assert info["warmstart_ms"] > 0
    # the same solve replayed through a captured graph (network included) is identical
    S32.set(use_graph=1)
    for _ in range(3):
        batch.solve()
        g = batch.fetch()
        assert np.array_equal(g.pivots, warm.pivots) and np.array_equal(g.objective, warm.objective)
    S32.set(use_graph=0)
    assert np.array_equal(warm.status, cold.status)
    rel = np.abs(warm.objective - cold.objective) / np.maximum(1, np.abs(cold.objective))
    assert rel.max() <= 1e-6
    assert warm.pivots.mean() < cold.pivots.mean()
    # the device's crash lists are predict()'s, and the solve replays them exactly
    _, lists = ws.predict(test[:12])
    for k in range(12):
        if warm.fallback[k]:
            continue
        o = oracle.solve(test[k].A, test[k].b, test[k].c, precision="fp32", warm=lists[k])
        assert warm.pivots[k] == o["pivots"] and np.array_equal(warm.basis[k], o["basis"])
    batch.free()
    ws.close()
    S32.close()
    S64.close()
