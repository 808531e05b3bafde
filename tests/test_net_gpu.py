"""The value network's GPU kernels (BF16 MFMA forward, backward, Adam)
against the CPU reference of the same arithmetic (oracle/net_ref.py)."""
import numpy as np
import pytest

import net_ref as R
from pts.net import ValueNet

pytestmark = pytest.mark.gpu


def batch(rows, seed):
    rng = np.random.default_rng(seed)
    x = np.zeros((rows, R.IN), np.float32)
    x[:, :52] = rng.random((rows, 52)) < 0.1
    x[:, 52:54] = rng.random((rows, 2))
    r = rng.random((rows, 2, 1326)) * (rng.random((rows, 2, 1326)) < 0.5)
    r /= r.sum(axis=2, keepdims=True)
    x[:, 54:] = r.reshape(rows, -1)
    target = (rng.standard_normal((rows, R.OUT)) * 0.6).astype(np.float32)
    mask = (rng.random((rows, R.OUT)) < 0.7).astype(np.float32)
    return x, target, mask


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30)


def test_forward_matches_the_reference():
    net = ValueNet(128, seed=3)
    p = net.params().astype(np.float64)
    x, _, _ = batch(128, 4)
    y = net.forward(x)
    _, yr = R.forward(p, x.astype(np.float64))
    assert y.shape == yr.shape
    # float32 accumulation vs float64 with identical bf16 rounding points
    assert rel(y, yr) < 2e-2, rel(y, yr)
    assert np.corrcoef(y.ravel(), yr.ravel())[0, 1] > 0.9999


def test_gradients_match_the_reference():
    net = ValueNet(64, seed=5)
    p = net.params().astype(np.float64)
    x, t, m = batch(64, 6)
    loss = net.gradients(x, t, m)
    g = net.grads().astype(np.float64)
    lr, gr = R.gradients(p, x.astype(np.float64), t.astype(np.float64), m.astype(np.float64))
    assert abs(loss - lr) <= 1e-3 * abs(lr)
    q = 0
    for (o, i) in R.shapes():
        for size in (o * i, o):
            a, b = g[q:q + size], gr[q:q + size]
            assert rel(a, b) < 3e-2, (o, i, size, rel(a, b))
            assert np.dot(a, b) / (np.linalg.norm(a) * np.linalg.norm(b) + 1e-300) > 0.999
            q += size
    assert q == len(g)


def test_adam_training_fits_a_batch():
    net = ValueNet(128, seed=7)
    x, t, m = batch(128, 8)
    t = (np.tanh(x[:, 54:54 + R.OUT] * 400.0) * 0.5).astype(np.float32)  # a learnable target
    first = net.train_step(x, t, m, 1e-3)
    for _ in range(60):
        last = net.train_step(x, t, m, 1e-3)
    assert last < 0.3 * first, (first, last)


def test_rows_must_be_a_multiple_of_64():
    net = ValueNet(64, seed=1)
    with pytest.raises(Exception):
        net.forward(np.zeros((10, R.IN), np.float32))


def test_adam_steps_match_the_reference():
    # two training steps; each Adam update against the textbook step
    # (oracle/net_ref.py) on the gradients the GPU computed, from the same
    # moments: the bias corrections 1 - b^t at t = 1 and 2, the moment
    # recursions and eps's place in the denominator
    net = ValueNet(64, seed=11)
    x, t, m = batch(64, 12)
    p = net.params().astype(np.float64)
    mom, vel = np.zeros_like(p), np.zeros_like(p)
    lr = 1e-3
    for step in (1, 2):
        net.train_step(x, t, m, lr)
        g = net.grads().astype(np.float64)
        want, mom, vel = R.adam(p, g, mom, vel, step, lr)
        got = net.params().astype(np.float64)
        du, dw = got - p, want - p
        live = np.abs(dw) > 1e-3 * lr
        assert np.count_nonzero(live) > 0.5 * len(p)
        assert np.max(np.abs(du - dw)) <= 2e-3 * lr, (step, np.max(np.abs(du - dw)))
        assert np.max(np.abs(du[live] - dw[live]) / np.abs(dw[live])) < 1e-2, step
        p = got
