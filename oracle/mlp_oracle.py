"""CPU oracle for the learned continuation-value network -- TEST INFRASTRUCTURE ONLY.

The network (2 -> 64 SiLU -> 64 SiLU -> 1, flat FP32 parameter vector laid out as W1[64][2],
b1[64], W2[64][64] (out, in), b2[64], W3[64], b3) has no reference counterpart; this module
states its forward pass and the gradient of the batch-mean squared error in numpy FP64.
With bf16=True it rounds exactly the operands the device feeds to its BF16 MFMA products
(H1, W2 and the layer-2 output gradient G2) to bfloat16 (round to nearest even), so device
and oracle differ only by FP32 accumulation order and hardware transcendentals.
"""
import numpy as np

H = 64
N_PARAMS = 2 * H + H + H * H + H + H + 1


def unpack(w):
    w = np.asarray(w, dtype=np.float64)
    o = 0
    W1 = w[o:o + 2 * H].reshape(H, 2); o += 2 * H
    b1 = w[o:o + H]; o += H
    W2 = w[o:o + H * H].reshape(H, H); o += H * H
    b2 = w[o:o + H]; o += H
    W3 = w[o:o + H]; o += H
    b3 = w[o]
    return W1, b1, W2, b2, W3, b3


def bf16(a):
    """Round to bfloat16 (nearest even) and return as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def silu(a):
    return a / (1.0 + np.exp(-a))


def silu_grad(a):
    s = 1.0 / (1.0 + np.exp(-a))
    return s * (1.0 + a * (1.0 - s))


def forward(w, x, t, bf=False):
    W1, b1, W2, b2, W3, b3 = unpack(w)
    X = np.stack([x, t], axis=1).astype(np.float64)
    A1 = X @ W1.T + b1
    H1 = silu(A1)
    if bf:
        H1 = bf16(H1)
        W2 = bf16(W2)
    A2 = H1 @ W2.T + b2
    H2 = silu(A2)
    y = H2 @ W3 + b3
    return y, (X, A1, H1, A2, H2, W2)


def gradients(w, x, t, y, bf=False):
    """Gradient of mean((f(x,t) - y)^2) over the batch, flat like the parameters."""
    W1, b1, W2f, b2, W3, b3 = unpack(w)
    yhat, (X, A1, H1, A2, H2, W2) = forward(w, x, t, bf)
    n = len(x)
    g3 = 2.0 * (yhat - y) / n
    dW3 = H2.T @ g3
    db3 = g3.sum()
    G2 = np.outer(g3, W3) * silu_grad(A2)
    db2 = G2.sum(axis=0)
    G2m = bf16(G2) if bf else G2
    dW2 = G2m.T @ H1
    G1 = (G2m @ W2) * silu_grad(A1)
    dW1 = G1.T @ X
    db1 = G1.sum(axis=0)
    return np.concatenate([dW1.ravel(), db1, dW2.ravel(), db2, dW3, [db3]]), yhat


def init(seed=0):
    rng = np.random.default_rng(seed)
    w = np.zeros(N_PARAMS)
    W1, b1, W2, b2, W3, b3 = unpack(w)
    o = 0
    w[o:o + 2 * H] = rng.normal(0, 1.0, 2 * H); o += 2 * H
    w[o:o + H] = rng.normal(0, 0.5, H); o += H
    w[o:o + H * H] = rng.normal(0, 1.0 / np.sqrt(H), H * H); o += H * H
    w[o:o + H] = rng.normal(0, 0.1, H); o += H
    w[o:o + H] = rng.normal(0, 1.0 / np.sqrt(H), H); o += H
    w[o] = 0.05
    return w.astype(np.float32)
