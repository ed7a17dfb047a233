"""CPU reference of the warm-start network (test infrastructure).

Forward and backward of  x -> W1 (ReLU) -> W2 (ReLU) -> W3 -> logits  with the per-family
mean sigmoid cross-entropy, written out by hand so that the rounding points of the GPU
kernels can be mirrored exactly: with `bf16=True` the inputs, weights, hidden activations
and the backward signals that feed GEMMs are rounded to bfloat16 (round to nearest even)
where the kernels round them, and products accumulate in float32.  With `bf16=False` and
float64 arrays it is the exact gradient of the loss (checked by finite differences).
"""
import numpy as np


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def unpack(theta, H, d):
    """Split a flat parameter vector (the ABI's order) into W1, b1, W2, b2, W3, b3."""
    shapes = [(H, d), (H,), (H, H), (H,), (d, H), (d,)]
    out, pos = [], 0
    for s in shapes:
        k = int(np.prod(s))
        out.append(theta[pos:pos + k].reshape(s))
        pos += k
    assert pos == theta.size
    return out


def features(b, c, mu, inv_sd):
    raw = np.concatenate([b, c], axis=1).astype(np.float32)
    return ((raw - mu) * inv_sd).astype(np.float32)


def labels(bases, d):
    Y = np.zeros((len(bases), d))
    for k, bs in enumerate(bases):
        Y[k, bs] = 1.0
    return Y


def forward_backward(theta, x, Y, H, bf16_emulation=True):
    """Loss (mean over rows and outputs) and its gradient, flat in the ABI's order."""
    d = x.shape[1]
    W1, b1, W2, b2, W3, b3 = unpack(theta, H, d)
    if bf16_emulation:
        r = bf16
        ft = np.float32
        x, W1, W2, W3 = r(x), r(W1), r(W2), r(W3)
        b1, b2, b3 = b1.astype(ft), b2.astype(ft), b3.astype(ft)
        Y = Y.astype(ft)
    else:
        r = lambda a: a
    n_rows = x.shape[0]
    H1 = r(np.maximum(x @ W1.T + b1, 0))
    H2 = r(np.maximum(H1 @ W2.T + b2, 0))
    Z = H2 @ W3.T + b3
    scale = 1.0 / (n_rows * d)
    loss = float(np.sum(np.maximum(Z, 0) - Z * Y + np.log1p(np.exp(-np.abs(Z)))) * scale)
    dZ = (1.0 / (1.0 + np.exp(-Z)) - Y) * scale
    if bf16_emulation:
        dZ = dZ.astype(np.float32)
    dZb = r(dZ)
    gW3 = dZb.T @ H2
    gb3 = dZ.sum(0)
    dA2 = (dZb @ W3) * (H2 > 0)
    dA2b = r(dA2)
    gW2 = dA2b.T @ H1
    gb2 = dA2.sum(0)
    dA1 = (dA2b @ W2) * (H1 > 0)
    dA1b = r(dA1)
    gW1 = dA1b.T @ x
    gb1 = dA1.sum(0)
    grad = np.concatenate([g.reshape(-1) for g in (gW1, gb1, gW2, gb2, gW3, gb3)])
    return loss, grad, Z
