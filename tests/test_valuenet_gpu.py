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


def test_value_net_workload_shapes_on_the_full_flop():
    # The value-net workload's shapes: a flop of the flop workload (its tree
    # and 169-class ranges, 1176 hands per player), the 2706-4x1024-2652
    # network trained with batches of 1024 rows on GPU-solved turn subgames
    # that start from that flop's ranges, the depth-limited solve, and the
    # judge: the full 16.19 G-cell tree with the flop pinned to the
    # depth-limited average strategy and the later streets solved by exact
    # DCFR, whose exploitability the exact FP64 best response measures.
    from pts import spots
    cfg = spots.flop_spot(1)
    X, Y, M, info = V.generate(128, 0, V.leaf_pots(cfg), batch=64, base=cfg.ranges, target_pct=1.0)
    assert X.shape == (128, 2706) and Y.shape == (128, 2652) and np.all(np.isfinite(Y))
    net = ValueNet(1024, seed=3)
    losses = V.train(net, X, Y, M, 30, batch=1024, lr=1e-3)
    assert np.all(np.isfinite(losses)) and losses[-1] < losses[0]
    dl = depth_limited(cfg, net)
    dl.iterate(50)
    assert np.isfinite(dl.exploitability()) and dl.exploitability() >= 0.0
    hist, sec, full, reached = V.resolve(cfg, dl, target_pct=0.0, max_iterations=12, check_every=6)
    assert full.stats().storage_cells == 16188966528
    assert [t for t, _ in hist] == [6, 12]
    assert 0.0 <= hist[1][1] < hist[0][1] and np.isfinite(hist[0][1])
    # the pinned flop root plays the depth-limited solve's average strategy
    qd, qf = dl.query([]), full.query([])
    assert qd.actions == qf.actions
    live = qd.reach > 0
    assert np.allclose(qf.strategy[:, live], qd.strategy[:, live], rtol=0, atol=1e-5)
