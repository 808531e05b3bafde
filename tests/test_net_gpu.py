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
    # Float inputs: a pre-activation within float32 rounding of zero can
    # take the other side of a ReLU mask than in the float64 reference and
    # move one row of a weight gradient, so the gradients are held to a
    # relative L2 error and a cosine, not to their largest element; the exact
    # integer problem below pins the arithmetic bit for bit.
    q = 0
    for (o, i) in R.shapes():
        for size in (o * i, o):
            a, b = g[q:q + size], gr[q:q + size]
            l2 = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)
            print(f"({o}, {i}) size {size}: relative L2 {l2:.3e}, max-norm {rel(a, b):.3e}")
            assert l2 < 3e-2, (o, i, size, l2, rel(a, b))
            assert np.dot(a, b) / (np.linalg.norm(a) * np.linalg.norm(b) + 1e-300) > 0.999
            q += size
    assert q == len(g)


def integer_problem(rows, seed):
    """Parameters, inputs and targets on which every sum the network forms is
    exact in float32 whatever its order: sparse weights in {-1, 0, 1},
    integer biases and inputs, targets an integer away from the outputs
    (Huber derivatives -1, 0 or 1) and a loss mask of exactly 2^17 cells, so
    every gradient is a multiple of 2^-17. The bounds below check that no
    partial sum can leave float32's exact range."""
    rng = np.random.default_rng(seed)
    parts = []
    for (o, i) in R.shapes():
        w = rng.choice([-1.0, 0.0, 1.0], size=(o, i), p=[1 / 128, 63 / 64, 1 / 128])
        parts += [w.ravel(), rng.integers(-2, 3, size=o).astype(np.float64)]
    p = np.concatenate(parts)
    x = ((rng.random((rows, R.IN)) < 0.2) * rng.integers(1, 4, (rows, R.IN))).astype(np.float64)
    acts, y = R.forward(p, x)
    layers = R.unpack(p, R.shapes())
    for l, (w, b) in enumerate(layers):  # |x| |W|^T + |b| bounds every partial sum of a layer
        assert np.max(np.abs(acts[l]) @ np.abs(w).T + np.abs(b)) < 2.0 ** 24
    t = y + rng.integers(-2, 3, size=y.shape)
    m = np.zeros(y.size)
    m[rng.choice(y.size, 1 << 17, replace=False)] = 1.0
    m = m.reshape(y.shape)
    # backward bounds, in units of 2^-17: dz, then dz^T A, the column sums
    # and (dz |W|) masked by the ReLU
    dz = np.abs(m * np.clip(y - t, -1, 1))
    for l in range(len(layers) - 1, -1, -1):
        w, _ = layers[l]
        assert np.max(dz.T @ np.abs(acts[l])) < 2.0 ** 24 and np.max(dz.sum(axis=0)) < 2.0 ** 24
        if l > 0:
            dz = (dz @ np.abs(w)) * (acts[l] > 0)
            assert np.max(dz) < 2.0 ** 24
    return p, x.astype(np.float32), t.astype(np.float32), m.astype(np.float32)


def test_integer_problem_is_exact():
    # on exact sums the GEMM's tiling, its K split across wave groups and its
    # epilogues are checked bit for bit, independently of summation order
    p, x, t, m = integer_problem(64, 11)
    net = ValueNet(64, seed=1)
    net.set_params(p.astype(np.float32))
    _, yr = R.forward(p, x.astype(np.float64))
    y = net.forward(x)
    assert np.array_equal(y.astype(np.float64), yr)
    loss = net.gradients(x, t, m)
    lr, gr = R.gradients(p, x.astype(np.float64), t.astype(np.float64), m.astype(np.float64))
    assert loss == lr
    g = net.grads().astype(np.float64)
    assert np.array_equal(g, gr), np.flatnonzero(g != gr)[:10]


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
