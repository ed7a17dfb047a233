"""CPU pins of the warm-start network's reference (tests/mlp_ref.py): the hand-written
backward equals central finite differences of the loss in float64, the bf16 rounding
helper rounds to nearest even, and the emulated forward stays within the bf16 bound of the
exact one."""
import numpy as np

from tests.mlp_ref import bf16, forward_backward, labels


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    d, H, rows = 7, 5, 6
    n_par = H * d + H + H * H + H + d * H + d
    theta = rng.normal(0, 0.7, n_par)
    x = rng.normal(0, 1, (rows, d))
    Y = labels([rng.choice(d, 3, replace=False) for _ in range(rows)], d)
    loss, g, _ = forward_backward(theta, x, Y, H, bf16_emulation=False)
    h = 1e-6
    fd = np.empty(n_par)
    for i in range(n_par):
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        fd[i] = (forward_backward(tp, x, Y, H, False)[0] - forward_backward(tm, x, Y, H, False)[0]) / (2 * h)
    assert np.max(np.abs(fd - g)) <= 1e-7 * max(1.0, np.max(np.abs(g)))
    assert loss > 0


def test_bf16_rounding():
    v = np.array([1.0, 1.0 + 2**-8, 1.0 + 3 * 2**-8, -2.5, 3.0e38, 1.0 + 2**-9], np.float32)
    r = bf16(v)
    # 1+2^-8 is a tie between 1 and 1+2^-7: to even (1); 1+3*2^-8 ties up to 1+2^-6
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == np.float32(1 + 2**-6)
    assert r[3] == -2.5 and r[5] == 1.0
    u = r.view(np.uint32)
    assert np.all((u & 0xFFFF) == 0)


def test_bf16_emulation_close_to_exact():
    rng = np.random.default_rng(1)
    d, H, rows = 40, 64, 32
    n_par = H * d + H + H * H + H + d * H + d
    theta = rng.normal(0, 0.2, n_par)
    x = rng.normal(0, 1, (rows, d))
    Y = labels([rng.choice(d, 10, replace=False) for _ in range(rows)], d)
    l0, g0, z0 = forward_backward(theta, x, Y, H, bf16_emulation=False)
    l1, g1, z1 = forward_backward(theta, x, Y, H, bf16_emulation=True)
    assert abs(l0 - l1) <= 1e-2 * l0
    assert np.linalg.norm(z1 - z0) <= 2e-2 * np.linalg.norm(z0)
    assert np.linalg.norm(g1 - g0) <= 5e-2 * np.linalg.norm(g0)
