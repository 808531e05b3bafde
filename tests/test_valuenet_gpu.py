"""The value-network path end to end on small spots: turn subgames solved by
the batched GPU DCFR become training rows whose targets invert the solver's
normalization, the network trained by its own kernels predicts held-out
subgames better than the method without it (a zero prediction), and the
trained network drives a depth-limited flop solve whose flop strategy warm
starts the exact full-tree solve."""
import numpy as np
import pytest

from pts import valuenet as V
from pts.net import ValueNet
from pts.solver import BatchSolver, Solver, depth_limited

pytestmark = pytest.mark.gpu


def test_targets_invert_the_solver_normalization():
    # an all-in turn: no betting, so every hand's EV is its equity share of
    # the pot against the opponent's range minus the half it put in
    from pts import make_config, PostflopMenu
    cfg = make_config("2c7d9hTs", "AhAd, KhKd", "QhQd, JhJd", 1000, 1000, PostflopMenu.one(["allin"], []))
    b = BatchSolver([cfg])
    b.iterate(1)
    rows = b.root_rows(2)
    x, y, m = V.rows_of(cfg, rows[0], b.num_combinations(0))
    assert m.sum() == 4 and abs(x[54:54 + 1326].sum() - 1) < 1e-6
    # EVs are pot fractions in [-1, 1] and the two players' weighted EVs sum to ~0 (zero-sum, no rake)
    ev0 = y[:1326][m[:1326] > 0]
    ev1 = y[1326:][m[1326:] > 0]
    assert np.all(np.abs(ev0) <= 1.0) and np.all(np.abs(ev1) <= 1.0)
    assert abs(ev0.mean() + ev1.mean()) < 0.05


def test_trained_network_beats_no_network_and_drives_the_solve():
    cfg = V.eval_flop(1, 8)
    pots = V.leaf_pots(cfg)
    X, Y, M, info = V.generate(192, 8, pots, batch=64, base=cfg.ranges)
    n = X.shape[0]
    ntr = 160
    net = ValueNet(1024, seed=3)
    V.train(net, X[:ntr], Y[:ntr], M[:ntr], 400, batch=256, lr=1e-3)
    held = V.heldout_loss(net, X[ntr:], Y[ntr:], M[ntr:])
    zero = float(np.sum(M[ntr:] * np.where(np.abs(Y[ntr:]) <= 1, 0.5 * Y[ntr:] ** 2, np.abs(Y[ntr:]) - 0.5)) /
                 np.sum(M[ntr:]))
    assert held < 0.8 * zero, (held, zero)
    dl = depth_limited(cfg, net)
    dl.iterate(100)
    pr, sec, full = V.judge(cfg, dl, 5.0, 200, 10)
    assert pr.exploitability_pct <= 5.0
