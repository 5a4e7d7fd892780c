"""Screener network kernels (forward, backward, Adam) against the CPU oracle
(oracle/mlp.py, FP64 with the same BF16 weight rounding).  Tolerances reflect
FP32 accumulation: 1e-5 relative on outputs and loss, 1e-4 of the largest
gradient entry on gradients."""
import numpy as np
import pytest

import mlp as M
from nwick.mlp import MLP

pytestmark = pytest.mark.gpu
NIN = 24


def _data(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, NIN)).astype(np.float32)
    t = (np.sin(X[:, 0]) + X[:, 1] * X[:, 2] - 0.5 * X[:, 3] ** 2).astype(np.float32)
    w = rng.uniform(0.5, 2.0, n).astype(np.float32)
    return X, t, w


def test_forward_matches_oracle():
    net = MLP(NIN, seed=1)
    X, t, w = _data(1000, 0)
    y = net.forward(X)
    yo, _ = M.forward(net.params.astype(np.float64), X, NIN)
    assert np.max(np.abs(y - yo)) <= 1e-5 * max(1.0, np.max(np.abs(yo)))


def test_gradient_matches_oracle():
    net = MLP(NIN, seed=2)
    X, t, w = _data(3000, 1)
    loss, g = net.grad(X, t, w)
    lo, go = M.loss_grad(net.params.astype(np.float64), X, t, w, NIN)
    print(f"mlp grad: loss gpu {loss:.8f} oracle {lo:.8f} max|dg| {np.max(np.abs(g - go)):.3e} "
          f"max|g| {np.max(np.abs(go)):.3e}", flush=True)
    assert abs(loss - lo) <= 1e-5 * abs(lo)
    assert np.max(np.abs(g - go)) <= 1e-4 * np.max(np.abs(go))


def test_adam_step_matches_oracle():
    net = MLP(NIN, seed=3)
    X, t, w = _data(512, 2)
    p0 = net.params.astype(np.float64)
    lr = 1e-2
    net.step(X, t, w, lr=lr)
    p1 = net.params.astype(np.float64)
    _, g = M.loss_grad(p0, X, t, w, NIN)
    po, _, _ = M.adam_step(p0, g, np.zeros_like(p0), np.zeros_like(p0), 1, lr)
    big = np.abs(g) > 1e-5
    assert np.max(np.abs(p1 - po)[big]) <= 1e-3 * lr + 1e-6
    assert np.max(np.abs(p1 - po)) <= 2 * lr


def test_training_reduces_loss():
    net = MLP(NIN, seed=4)
    X, t, w = _data(20000, 3)
    l0, _ = net.grad(X, t)
    el = net.train(X, t, epochs=15, batch=1024, lr=2e-3, seed=5)
    l1, _ = net.grad(X, t)
    print(f"mlp training: loss {l0:.4f} -> {l1:.4f}, epochs {np.round(el, 4).tolist()}", flush=True)
    assert l1 < 0.1 * l0
