"""The product's hand evaluator (pts_evaluate, the one the tree builder's
strength tables use) against the oracle's pinned evaluator on every 5-card
and every 7-card hand, and the tree builder's node counts against the
independent counts of the workloads (a throwaway restatement of count_tree
and postflop_actions, crates/solver/src/tree.rs:337-401,
crates/holdem/src/menu.rs:264-392)."""
import numpy as np
import pytest

import pyoracle as O
from pts import native, spots
from pts.solver import Solver

# 7-card category totals over all C(52, 7) hands (textbook closed forms):
# high card .. straight flush (eval.rs:70-198 numbers categories 0..8)
SEVEN_CARD_TOTALS = [23294460, 58627800, 31433400, 6461620, 6180020, 4047644, 3473184, 224848, 41584]
# eval.rs:245-265: 5-card category counts
FIVE_CARD_TOTALS = [1302540, 1098240, 123552, 54912, 10200, 5108, 3744, 624, 40]


def _exhaustive(ncards, total, chunk):
    counts = np.zeros(9, np.int64)
    distinct = set()
    first = 0
    while first < total:
        hands = O.enumerate_hands(ncards, first, chunk)
        assert len(hands) == min(chunk, total - first)
        mine = native.evaluate(hands)
        ref = O.evaluate_batch(hands)
        bad = np.nonzero(mine != ref)[0]
        assert len(bad) == 0, f"hand {hands[bad[0]].tolist()}: product {mine[bad[0]]:#x}, oracle {ref[bad[0]]:#x}"
        counts += np.bincount(mine >> 26, minlength=9)
        distinct.update(np.unique(mine).tolist())
        first += len(hands)
    assert len(O.enumerate_hands(ncards, total, 1)) == 0  # the enumeration ends where C(52, n) says
    return counts.tolist(), len(distinct)


def test_every_five_card_hand():
    counts, distinct = _exhaustive(5, 2598960, 1 << 20)
    assert counts == FIVE_CARD_TOTALS
    assert distinct == 7462


def test_every_seven_card_hand():
    counts, distinct = _exhaustive(7, 133784560, 1 << 23)
    assert counts == SEVEN_CARD_TOTALS
    assert distinct == 4824


def test_evaluator_rejects_bad_hands():
    with pytest.raises(native.PtsError, match="repeats a card"):
        native.evaluate(np.array([[0, 1, 2, 3, 3]], np.uint8))
    with pytest.raises(native.PtsError, match="above 51"):
        native.evaluate(np.array([[0, 1, 2, 3, 52]], np.uint8))
    with pytest.raises(native.PtsError, match="5 to 7 cards"):
        native.evaluate(np.array([[0, 1, 2, 3]], np.uint8))


@pytest.mark.parametrize("maker,player,chance,terminal", [
    ("river_spot", 16, 0, 29),
    ("turn_spot", 5776, None, 10526),
    ("flop_spot", 5056824, 13988, 9366764),
])
def test_workload_tree_counts(maker, player, chance, terminal):
    cfg = getattr(spots, maker)(0)
    a = Solver.estimate(cfg)
    b = O.OracleSolver.estimate(cfg)
    assert (a.player_nodes, a.terminal_nodes) == (player, terminal)
    if chance is not None:
        assert a.chance_nodes == chance
    assert (a.player_nodes, a.chance_nodes, a.terminal_nodes, a.storage_cells) == \
        (b.player_nodes, b.chance_nodes, b.terminal_nodes, b.storage_cells)
