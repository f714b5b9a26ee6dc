"""The cost-model MLP's HIP kernels against the double-precision oracle.

FP32 forward outputs and gradients are compared with the oracle's FP64
evaluation at a relative tolerance of 1e-4 (FP32 rounding over 64-term dot
products and a 2000-sample mean); training must reduce the loss.
"""
import ctypes

import numpy as np
import pytest

from tests.gpu_util import announce

pytestmark = pytest.mark.gpu


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _params(mgx, D, seed=3):
    P = mgx.lib().mgx_mlp_param_count(D)
    th = np.empty(P, dtype=np.float32)
    mgx._lib.call("mgx_mlp_init", th.ctypes.data, D, seed)
    return th


def test_forward_matches_oracle(mgx, orc):
    announce("mlp forward")
    r = np.random.default_rng(1)
    D, n = 9, 500
    th = _params(mgx, D)
    X = _f32(r.normal(0, 1, (n, D)))
    y = np.empty(n, dtype=np.float32)
    mgx._lib.call("mgx_mlp_predict", th.ctypes.data, D, X.ctypes.data, n, y.ctypes.data)
    ref = orc.mlp_forward(th.astype(np.float64), X.astype(np.float64))
    np.testing.assert_allclose(y, ref, rtol=1e-4, atol=1e-4 * np.abs(ref).max())


def test_gradient_matches_oracle(mgx, orc):
    announce("mlp gradient")
    r = np.random.default_rng(2)
    D, n = 9, 2000
    th = _params(mgx, D)
    X = _f32(r.normal(0, 1, (n, D)))
    t = _f32(r.normal(0, 1, n))
    P = th.size
    g = np.empty(P, dtype=np.float32)
    loss = np.zeros(1, dtype=np.float32)
    mgx._lib.call("mgx_mlp_gradient", th.ctypes.data, D, X.ctypes.data, t.ctypes.data, n,
                  g.ctypes.data, loss.ctypes.data)
    ref_loss, ref_g = orc.mlp_gradient(th.astype(np.float64), X.astype(np.float64),
                                       t.astype(np.float64))
    assert loss[0] == pytest.approx(ref_loss, rel=1e-4)
    scale = np.abs(ref_g).max()
    np.testing.assert_allclose(g, ref_g, rtol=1e-3, atol=1e-4 * scale)


def test_training_reduces_loss(mgx, orc):
    announce("mlp train")
    r = np.random.default_rng(3)
    D, n = 5, 1500
    X = _f32(r.uniform(-1, 1, (n, D)))
    t = _f32(np.sin(2 * X[:, 0]) + X[:, 1] * X[:, 2] + 0.5 * X[:, 3])
    th = _params(mgx, D)
    epochs = 400
    hist = np.empty(epochs, dtype=np.float32)
    mgx._lib.call("mgx_mlp_train", th.ctypes.data, D, X.ctypes.data, t.ctypes.data, n, epochs,
                  ctypes.c_float(3e-3), ctypes.c_float(0.0), hist.ctypes.data)
    assert hist[-1] < 0.1 * hist[0]
    # the trained parameters evaluate to the recorded loss on the CPU oracle
    final = np.mean((orc.mlp_forward(th.astype(np.float64), X.astype(np.float64)) - t) ** 2)
    assert final < 0.1 * hist[0]
