"""Fast precision (the benched mode: terminal sums as parallel f64 scans)
against the oracle at the workloads' own iteration counts, on the tolerance
the path is held to: per information set (decision node x hand) the average
strategies within 1e-3 in L1 (sum over actions of |p - q|), and
exploitability within 0.01% of the pot. The metric follows the reference's
strategy comparison (crates/solver/examples/compare.rs:55-94), taken at every
information set instead of a few nodes.

The full flop (16.19 G cells) is too big for the oracle; it is checked by
properties of the solution after a bounded number of iterations (zero-sum
root values, best responses at least the average values, exploitability
falling, strategies that sum to one, finite values), and fast precision
against exact precision on its first iterations."""
import numpy as np
import pytest

import pyoracle as O
from pts import spots
from pts.compare import infoset_l1, node_diff
from pts.native import EXACT, FAST
from pts.solver import Solver

pytestmark = pytest.mark.gpu

CASES = [("river0", lambda: spots.river_spot(0), 1000), ("river1", lambda: spots.river_spot(1), 1000),
         ("river2", lambda: spots.river_spot(2), 1000), ("turn0", lambda: spots.turn_spot(0), 500)]


@pytest.mark.parametrize("name,make,iters", CASES, ids=[c[0] for c in CASES])
def test_fast_precision_within_the_tolerance(name, make, iters):
    cfg = make()
    c = O.OracleSolver(cfg)
    c.iterate(iters)
    g = Solver(cfg, precision=FAST)
    g.iterate(iters)
    _, sg = g.storage()
    _, sc = c.storage()
    d = infoset_l1(g.nodes(), g.n, sg, sc, tol=1e-3)
    print(f"{name}: {d.infosets} information sets, max L1 {d.max_l1:.3e}, mean {d.mean_l1:.3e}, "
          f"{int(np.sum(sg != sc))} of {sg.size} strategy-sum cells differ")
    assert d.over == 0, f"{d.over} information sets above 1e-3 (max {d.max_l1:.3e} at node {d.worst_node})"
    eg, ec = g.exploitability(), c.exploitability_f64()
    assert abs(eg - ec) <= 1e-4 * cfg.root.pot, (eg, ec)
    nd = node_diff(c.query([]), g.query([]))
    assert nd.strat_l1_pct <= 0.1 and nd.worst_loss_pct <= 0.01, nd


def test_full_flop_properties():
    cfg = spots.flop_spot(0)
    pot = cfg.root.pot
    s = Solver(cfg, precision=FAST)
    assert s.stats().storage_cells == 16188966528
    # every segment with a walk but the root (671,448 segments: the root,
    # 319,872 one-member chance children, 351,575 walks of 7 shapes) runs in
    # a shape-specialised kernel
    assert s.jit_segments() == 351085
    s.iterate(4)
    e4 = s.exploitability()
    s.iterate(12)
    e16 = s.exploitability()
    assert 0.0 <= e16 < e4, (e4, e16)
    avg = s.root_values(2)
    br = s.root_values(1)
    assert np.all(np.isfinite(avg)) and np.all(np.isfinite(br))
    assert abs(avg[0] + avg[1]) <= 1e-9 * pot, avg  # no rake: the average profile is zero-sum
    assert br[0] >= avg[0] and br[1] >= avg[1], (br, avg)
    assert abs((br[0] + br[1]) / 2.0 - e16) <= 1e-9 * pot  # exploitability is the mean best-response gain
    root = s.query([])
    paths = [[], ["check"], ["check", "check", ("card", _free(cfg, 0))],
             ["check", "check", ("card", _free(cfg, 0)), "check", "check", ("card", _free(cfg, 1))]]
    for path in paths:
        q = s.query(path)
        assert np.all(np.isfinite(q.strategy)) and np.all(np.isfinite(q.ev)), path
        live = q.reach > 0
        assert np.allclose(q.strategy.sum(axis=0)[live], 1.0, atol=1e-5), path
    assert len(root.actions) == 4  # check, 33%, 75%, all-in
    fast = (e16, avg.copy(), root.strategy.copy())
    del s
    # exact precision on the same iterations: the fast sums differ only in f64 order
    x = Solver(cfg, precision=EXACT)
    x.iterate(16)
    ex = x.exploitability()
    assert abs(ex - fast[0]) <= 1e-4 * pot, (ex, fast[0])
    assert np.allclose(x.root_values(2), fast[1], rtol=0, atol=1e-4 * pot)
    assert np.max(np.abs(x.query([]).strategy - fast[2])) <= 1e-3


def _free(cfg, k):
    return [c for c in range(52) if c not in cfg.board][k]
