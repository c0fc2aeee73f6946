"""GPU checks of the BF16-MFMA network kernels against the numpy oracle (oracle/mlp_oracle.py).

The oracle itself is pinned by central finite differences of its FP64 loss. Device vs the
BF16-faithful oracle: forward within 2e-3 of the output scale, gradient within 2e-3 relative L2 (only
FP32 accumulation order and hardware exp/rcp differ); device vs the exact FP64 oracle within
the BF16 rounding of the MFMA operands (3e-2)."""
import os
import sys

import numpy as np
import pytest

import mcdx

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as M  # noqa: E402


def batch(seed=1, n=128):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.5, 1.2, n).astype(np.float32)
    t = rng.uniform(0.0, 1.0, n).astype(np.float32)
    y = (np.maximum(1.0 - x, 0.0) + 0.05 * rng.normal(size=n)).astype(np.float32)
    return x, t, y


def test_oracle_gradient_is_the_derivative():
    w = M.init(3).astype(np.float64)
    x, t, y = batch(2, 16)
    g, _ = M.gradients(w, x, t, y)

    def loss(v):
        return np.mean((M.forward(v, x, t)[0] - y) ** 2)
    rng = np.random.default_rng(4)
    for i in rng.choice(M.N_PARAMS, 40, replace=False):
        e = np.zeros_like(w)
        e[i] = 1e-6
        fd = (loss(w + e) - loss(w - e)) / 2e-6
        assert abs(fd - g[i]) < 1e-6 + 1e-5 * abs(g[i]), i


@pytest.mark.gpu
def test_forward_matches_oracle():
    w = M.init(5)
    x, t, _ = batch(6, 1000)
    got = mcdx.mlp_forward(w, x, t)
    want_bf, _ = M.forward(w, x, t, bf=True)
    want, _ = M.forward(w, x, t)
    # A BF16 operand can round the other way when the device's FP32 activation differs in
    # its last bits, so the bound is on the error relative to the output scale.
    scale = float(np.max(np.abs(want)))
    err_bf = float(np.max(np.abs(got - want_bf))) / scale
    err = float(np.max(np.abs(got - want))) / scale
    assert err_bf < 2e-3, err_bf
    assert err < 3e-2, err


@pytest.mark.gpu
def test_gradients_match_oracle():
    w = M.init(7)
    x, t, y = batch(8, 128)
    g, loss = mcdx.mlp_gradients(w, x, t, y)
    gb, yhat = M.gradients(w, x, t, y, bf=True)
    ge, _ = M.gradients(w, x, t, y)
    rel_b = np.linalg.norm(g - gb) / np.linalg.norm(gb)
    rel_e = np.linalg.norm(g - ge) / np.linalg.norm(ge)
    assert rel_b < 2e-3, rel_b
    assert rel_e < 3e-2, rel_e
    assert abs(loss.sum() - np.sum((yhat - y) ** 2)) < 1e-3 * np.sum((yhat - y) ** 2) + 1e-6
