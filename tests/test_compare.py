"""The strategy-comparison harness (pts.compare) on the oracle, no GPU:
its average strategy equals the oracle's query strategy bit for bit (the
normalization of crates/solver/src/cfr.rs:124-154), the infoset L1 of a
storage with itself is zero and a planted difference is found where it was
planted, and node_diff follows compare.rs:55-94 on a hand-built pair."""
import numpy as np

import pyoracle as O
from pts import spots
from pts.compare import average_strategy, infoset_l1, node_diff
from pts.solver import NodeInfo


def test_average_strategy_is_the_oracle_query_strategy():
    cfg = spots.river_spot(2)
    o = O.OracleSolver(cfg)
    o.iterate(40)
    _, s = o.storage()
    nodes = o.nodes()
    q = o.query([])
    na, n = len(q.actions), len(q.combos)
    off = int(nodes["offset"][0])
    mine = average_strategy(s[off:off + na * n], na, n)
    assert np.array_equal(mine, q.strategy)


def test_infoset_l1_finds_a_planted_difference():
    cfg = spots.river_spot(3)
    o = O.OracleSolver(cfg)
    o.iterate(30)
    _, s = o.storage()
    nodes = o.nodes()
    same = infoset_l1(nodes, o.n, s, s)
    assert same.max_l1 == 0.0 and same.over == 0
    # plant: node 0, hand 5 plays its first action only in the copy
    t = s.copy()
    na, n, off = int(nodes["num_children"][0]), o.n[0], int(nodes["offset"][0])
    for a in range(na):
        t[off + a * n + 5] = 1.0 if a == 0 else 0.0
    d = infoset_l1(nodes, o.n, s, t)
    p = average_strategy(s[off:off + na * n], na, n)[:, 5].astype(np.float64)
    want = abs(1.0 - p[0]) + float(np.sum(p[1:]))
    assert d.worst_node == 0 and d.worst_hand == 5 and abs(d.max_l1 - want) < 1e-6 and d.over == 1


def test_node_diff_on_a_hand_built_pair():
    # two hands, two actions, pot 100: hand 0 identical, hand 1 swaps its best action
    a = NodeInfo(player=0, actions=["check", ("raise", 50)], combos=np.array([7, 9]), pot=100,
                 strategy=np.array([[1.0, 0.2], [0.0, 0.8]], np.float32),
                 ev=np.array([[10.0, 5.0], [4.0, 8.0]], np.float32), reach=np.array([1.0, 3.0], np.float32))
    b = NodeInfo(player=0, actions=["check", ("raise", 50)], combos=np.array([7, 9]), pot=100,
                 strategy=np.array([[1.0, 0.6], [0.0, 0.4]], np.float32),
                 ev=np.array([[10.0, 7.0], [4.0, 6.0]], np.float32), reach=np.array([1.0, 3.0], np.float32))
    d = node_diff(a, b)
    # hand 1 (EVs check 5 / raise 8 against 7 / 6): L1/2 = 0.4, mean EV difference (2 + 2) / 2 / 100 = 2%,
    # and following the second solve's best action (check) loses 8 - 5 = 3% of the pot
    assert abs(d.strat_l1_pct - 100 * 3 * 0.4 / 4) < 1e-5  # f32 strategies
    assert abs(d.ev_abs_pct - 3 * 2.0 / 4) < 1e-6
    assert abs(d.loss_pct - 3 * 3.0 / 4) < 1e-6
    assert abs(d.disagree_pct - 75.0) < 1e-6 and abs(d.worst_loss_pct - 3.0) < 1e-6
