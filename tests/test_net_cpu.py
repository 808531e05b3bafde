"""The value network's CPU reference (oracle/net_ref.py) pinned by
mathematics: its float64 gradients equal central finite differences of its
loss on a small network of the same structure, and its bf16 rounding is
round-to-nearest-even."""
import numpy as np

import net_ref as R


def test_bf16_rounding_is_nearest_even():
    assert R.bf16(1.0) == 1.0
    assert R.bf16(1.0 + 2.0 ** -8) == 1.0          # a tie rounds to the even mantissa
    assert R.bf16(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6
    assert R.bf16(1.0 + 2.0 ** -7) == 1.0 + 2.0 ** -7
    assert R.bf16(-3.14159) == -3.140625
    x = np.random.default_rng(0).standard_normal(1000)
    assert np.all(np.abs(R.bf16(x) - x) <= np.abs(x) * 2.0 ** -8)


def test_reference_gradients_match_finite_differences():
    rng = np.random.default_rng(1)
    shp = R.shapes(din=7, hid=6, dout=5, layers=5)
    n = sum(o * i + o for o, i in shp)
    params = rng.standard_normal(n) * 0.7
    x = rng.standard_normal((4, 7))
    target = rng.standard_normal((4, 5)) * 2.0   # some entries beyond the Huber knee
    mask = (rng.random((4, 5)) < 0.8).astype(float)
    loss, g = R.gradients(params, x, target, mask, shp, quantize=False)
    eps = 1e-6
    num = np.zeros(n)
    for i in range(n):
        p1, p2 = params.copy(), params.copy()
        p1[i] += eps
        p2[i] -= eps
        num[i] = (R.gradients(p1, x, target, mask, shp, quantize=False)[0]
                  - R.gradients(p2, x, target, mask, shp, quantize=False)[0]) / (2 * eps)
    assert np.allclose(g, num, rtol=1e-5, atol=1e-7), np.max(np.abs(g - num))
    assert loss > 0


def test_reference_shapes_match_the_network():
    shp = R.shapes()
    assert shp[0] == (1024, 2706) and shp[-1] == (2652, 1024) and len(shp) == 5
    assert sum(o * i + o for o, i in shp) == 8639068


def test_reference_adam_closed_forms():
    # t = 1 from zero moments: m_hat = g, v_hat = g^2, so the step is
    # lr * g / (|g| + eps); at t = 2 with the same gradient the bias
    # corrections make the step the same again
    g = np.array([1e-9, 1e-8, 0.5, -2.0, 0.0])
    w, m, v = R.adam(np.zeros(5), g, 0.0, 0.0, 1, 1e-3)
    assert np.allclose(w, -1e-3 * g / (np.abs(g) + 1e-8), rtol=1e-12, atol=0)
    w2, _, _ = R.adam(w, g, m, v, 2, 1e-3)
    assert np.allclose(w2 - w, w, rtol=1e-12, atol=0)


def test_resolve_gives_up_when_the_target_projects_beyond_the_budget():
    from pts.valuenet import projects_beyond
    assert not projects_beyond([(20, 130.0)], 0.5, 300)  # one check: no rate yet
    # 130 -> 124 over 20 iterations: 0.3 % pot per iteration, 412 more to 0.5%
    assert projects_beyond([(20, 130.0), (40, 124.0)], 0.5, 300)
    assert not projects_beyond([(20, 130.0), (40, 124.0)], 0.5, 500)
    # 3.0 -> 1.0 over 20: 0.1 per iteration, 5 more to 0.5%
    assert not projects_beyond([(20, 3.0), (40, 1.0)], 0.5, 50)
    assert projects_beyond([(20, 3.0), (40, 1.0)], 0.5, 44)
    assert projects_beyond([(20, 3.0), (40, 3.0)], 0.5, 10 ** 6)  # flat
    assert projects_beyond([(20, 3.0), (40, 3.5)], 0.5, 10 ** 6)  # rising
