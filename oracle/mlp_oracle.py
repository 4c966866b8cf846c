"""CPU restatement of the candidate-proposer network (6 -> 128 -> 128 -> 4).

TEST INFRASTRUCTURE ONLY: the checker for the GPU forward / backward kernels
of the product library (helm-project_amd/csrc/mlp.hip).  Only tests/ use it.

Two modes:
  * emulate=True mirrors the kernels' arithmetic: inputs, weights,
    activations and back-propagated signals rounded to BF16 (round to nearest
    even) where the kernels round them, FP32 products and sums;
  * emulate=False is the same network in float64 without rounding, the form
    the finite-difference test pins.

The network has no counterpart in the reference repository; it follows the
layer sizes and loss the proposer uses: outputs o[0:3] are the cap-centre
offset, o[3] the log radius, trained with a squared error (weight wc) on the
centre and a pinball loss (quantile tau) on the log radius.
"""
import numpy as np

K_IN, K_H, K_OUT = 6, 128, 4
O_W1 = 0
O_B1 = O_W1 + K_H * K_IN
O_W2 = O_B1 + K_H
O_B2 = O_W2 + K_H * K_H
O_W3 = O_B2 + K_H
O_B3 = O_W3 + K_OUT * K_H
N_PARAMS = O_B3 + K_OUT


def bf16(a):
    """Round float32 values to the nearest bfloat16 (ties to even), as float32."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def unpack(p):
    p = np.asarray(p)
    return (p[O_W1:O_B1].reshape(K_H, K_IN), p[O_B1:O_W2], p[O_W2:O_B2].reshape(K_H, K_H), p[O_B2:O_W3],
            p[O_W3:O_B3].reshape(K_OUT, K_H), p[O_B3:])


def forward(p, x, emulate=True):
    """Outputs o [n, 4] and the intermediates (xb, h1, h2)."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    if emulate:
        f = np.float32
        xb = bf16(np.asarray(x, dtype=f))
        h1 = bf16(np.maximum(xb @ bf16(W1.astype(f)).T + b1.astype(f), 0))
        h2 = bf16(np.maximum(h1 @ bf16(W2.astype(f)).T + b2.astype(f), 0))
        o = h2 @ bf16(W3.astype(f)).T + b3.astype(f)
    else:
        xb = np.asarray(x, dtype=np.float64)
        h1 = np.maximum(xb @ W1.T + b1, 0)
        h2 = np.maximum(h1 @ W2.T + b2, 0)
        o = h2 @ W3.T + b3
    return o, (xb, h1, h2)


def loss_grad(p, x, t, wc, tau, emulate=True):
    """Mean loss over the samples and its gradient with respect to p."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    n = x.shape[0]
    o, (xb, h1, h2) = forward(p, x, emulate)
    d = o[:, :3] - t[:, :3]
    u = t[:, 3] - o[:, 3]
    loss = (wc * (d * d).sum(axis=1) + np.where(u > 0, tau * u, (tau - 1) * u)).sum() / n
    g = np.zeros_like(o)
    g[:, :3] = 2 * wc * d
    g[:, 3] = np.where(u > 0, -tau, 1 - tau)
    g = g / n
    if emulate:
        f = np.float32
        g = g.astype(f)
        gb = bf16(g)
        W3b, W2b = bf16(W3.astype(f)), bf16(W2.astype(f))
        dW3 = gb.T @ h2
        db3 = g.sum(axis=0)
        dz2 = (gb @ W3b) * (h2 > 0)
        db2 = dz2.sum(axis=0)
        dz2b = bf16(dz2)
        dW2 = dz2b.T @ h1
        dz1 = (dz2b @ W2b) * (h1 > 0)
        db1 = dz1.sum(axis=0)
        dW1 = bf16(dz1).T @ xb
    else:
        dW3 = g.T @ h2
        db3 = g.sum(axis=0)
        dz2 = (g @ W3) * (h2 > 0)
        db2 = dz2.sum(axis=0)
        dW2 = dz2.T @ h1
        dz1 = (dz2 @ W2) * (h1 > 0)
        db1 = dz1.sum(axis=0)
        dW1 = dz1.T @ xb
    grad = np.concatenate([dW1.ravel(), db1, dW2.ravel(), db2, dW3.ravel(), db3])
    return float(loss), grad
