"""Pins of the oracle's primitives: cards/combos, evaluator, ranges, menu,
DCFR factors and the check schedule, against the reference's own test
values (file:line cited) and closed forms."""
import math

import numpy as np
import pytest

import pyoracle as O
from pts import PostflopMenu, parse_cards, combo_index, combo_cards, HandClass, parse_range, StreetSizes
from pts.config import sizes


def test_cards_and_combos():
    # card.rs:254-255: As = (12, 3), 2c = 0
    assert parse_cards("As") == [12 * 4 + 3]
    assert parse_cards("2c") == [0]
    seen = set()
    for a in range(52):
        for b in range(52):
            if a == b:
                continue
            i = combo_index(a, b)
            assert i == O.lib().orc_combo_index(a, b)
            x, y = combo_cards(i)
            assert {x, y} == {a, b} and x >> 2 >= y >> 2
            seen.add(i)
    assert seen == set(range(1326))  # combo.rs:213-230
    assert combo_index(0, 1) == 0  # properties.rs:193 (2c2d)
    # combo.rs:246-248
    assert HandClass(12, 11, True).grid_index() == 1
    assert HandClass(12, 11, False).grid_index() == 13
    assert HandClass(0, 0, False).grid_index() == 168
    for g in range(169):
        for c in HandClass.from_grid_index(g).combos():
            assert O.lib().orc_combo_class(c) == g


def test_evaluator_known_orderings():
    ev = lambda s: O.evaluate(parse_cards(s))  # noqa: E731
    cat = lambda s: ev(s) >> 26  # noqa: E731
    # eval.rs:210-222
    assert cat("AsKsQsJsTs2d3c") == 8
    assert cat("As2s3s4s5s9d9c") == 8
    assert cat("AsAhAdAc5s9d9c") == 7
    assert cat("AsAhAd9c5s9d2c") == 6
    assert cat("AsAhAd9c9s9d2c") == 6
    assert cat("As9s5s3s2sKdKc") == 5
    assert cat("As2d3h4c5s9d9c") == 4
    assert cat("7s7d7h4c5sKdJc") == 3
    assert cat("7s7d4h4c5sKdJc") == 2
    assert cat("7s7d4h3c5sKdJc") == 1
    assert cat("7s8d4h3c2sKdJc") == 0
    # eval.rs:225-243
    assert ev("As2d3h4c5s") < ev("6s2d3h4c5s") < ev("AsKdQhJcTs")
    assert ev("AsAd9h4c2s") < ev("AsAdTh4c2s")
    assert ev("AsAd9h4c2s") > ev("KsKdQh4c2s")
    assert ev("AsAd2h2c3s") > ev("KsKdQhQc3s")
    assert ev("AsAd3h3c2s") > ev("AsAd2h2cKs")
    assert ev("AsAd3h3c5s") > ev("AsAd3h3c4s")
    assert ev("3s3d3hAcAs") > ev("2s2d2hAcAs")
    assert ev("3s3d3hAcAs") > ev("3s3d3hKcKs")
    assert ev("AsKdQhJcTs2d3c") == ev("AsKdQhJcTs4d5c")
    assert ev("AsAhAd9c9s9d2c") == ev("AsAhAd9c9s2d3c")


def test_five_card_category_counts():
    # eval.rs:245-265, plus the textbook 7,462 distinct 5-card hand values.
    counts, distinct = O.count_categories(5)
    assert counts == [1302540, 1098240, 123552, 54912, 10200, 5108, 3744, 624, 40]
    assert distinct == 7462


@pytest.mark.slow
def test_seven_card_category_counts():
    # Textbook totals over all 133,784,560 seven-card hands and the 4,824
    # distinct best-five values they reach.
    counts, distinct = O.count_categories(7)
    assert counts == [23294460, 58627800, 31433400, 6461620, 6180020, 4047644, 3473184, 224848, 41584]
    assert distinct == 4824


def test_seven_card_is_best_five_brute_force():
    # eval.rs:267-288 on a seeded sample
    rng = np.random.default_rng(7)
    from itertools import combinations
    for _ in range(3000):
        seven = list(rng.permutation(52)[:7])
        best = max(O.evaluate(list(f)) for f in combinations(seven, 5))
        assert O.evaluate(seven) == best


@pytest.mark.parametrize("s,count", [("AA", 6), ("AKs", 4), ("AKo", 12), ("AK", 16), ("22+", 78), ("TT+", 30),
                                     ("A2s+", 48), ("KTo+", 36), ("99-55", 30), ("A9s-A6s", 16), ("AsKs", 1),
                                     ("any", 1326), ("AA, KK:0.5", 9), ("A5+", 144)])
def test_range_notation(s, count):
    # range.rs:275-293; the product's parser and the oracle's must agree
    w = O.parse_range(s)
    assert float(w.sum()) == count
    assert np.array_equal(np.asarray(parse_range(s), np.float32), w)


def test_range_errors():
    for bad in ("AKx", "AA-KQs"):
        with pytest.raises(ValueError):
            O.parse_range(bad)
        with pytest.raises(ValueError):
            parse_range(bad)


def _sit(street, pot, to_call, current_bet, max_to):
    # menu.rs:398-410
    return dict(street=street, pot=pot, to_call=to_call, current_bet=current_bet, max_to=max_to,
                min_raise_to=100 if current_bet == 0 else current_bet * 2, num_raises=0 if current_bet == 0 else 1)


def _acts(s, menu=None, **over):
    d = dict(s)
    d.update(over)
    return O.postflop_actions(d["street"], d["pot"], d["to_call"], d["current_bet"], d["max_to"], d["min_raise_to"],
                              d["num_raises"], menu or PostflopMenu())


def test_menu_cases():
    # menu.rs:421-466
    assert _acts(_sit(1, 600, 0, 0, 9750)) == ["check", ("raise", 198), ("raise", 450)]
    assert _acts(_sit(1, 1050, 450, 450, 9750)) == ["fold", "call", ("raise", 1350)]
    assert _acts(_sit(3, 2000, 0, 0, 1500)) == ["check", ("raise", 1500)]
    assert _acts(_sit(2, 2000, 0, 0, 2900)) == ["check", ("raise", 1000), ("raise", 2900)]
    assert _acts(_sit(2, 5000, 1000, 2000, 9000), num_raises=3) == ["fold", "call"]


def test_menu_explicit_no_thresholds():
    m = PostflopMenu.one(["33%", "75%", "allin"], ["2.5x", "allin"], max_raises=2)
    # pot 2000, 9000 behind: 33% = 660, 75% = 1500, all-in 9000
    assert _acts(dict(street=2, pot=2000, to_call=0, current_bet=0, max_to=9000, min_raise_to=100, num_raises=0), m) == \
        ["check", ("raise", 660), ("raise", 1500), ("raise", 9000)]
    # facing 660: 2.5x = 1650, all-in
    assert _acts(dict(street=2, pot=2660, to_call=660, current_bet=660, max_to=9000, min_raise_to=1320, num_raises=1), m) == \
        ["fold", "call", ("raise", 1650), ("raise", 9000)]


def test_discount_closed_forms():
    # cfr.rs:86-106: pos = (t-1)^a/((t-1)^a+1), neg = 0.5 for beta 0 (t>1),
    # strat = ((t-1)/t)^gamma without restarts
    for t in (1, 2, 3, 10, 77, 500):
        pos, neg, strat = O.discount(t, 1.5, 0.0, 2.0)
        prev = t - 1
        pa = prev ** 1.5
        assert pos == np.float32(pa / (pa + 1.0))
        assert neg == np.float32((prev ** 0.0) / (prev ** 0.0 + 1.0))
        assert strat == np.float32((prev / t) ** 2.0)
    # restarts at powers of 4: t=17 -> since = 16 - 16 = 0
    assert O.discount(17, 1.5, 0.0, 3.0, restart=True)[2] == 0.0
    assert O.discount(20, 1.5, 0.0, 3.0, restart=True)[2] == np.float32((3 / 4) ** 3)


def test_check_schedule():
    # tests.rs:347-359
    g = O.next_check_gap
    assert g(None, (20, 2.0), 0.5, 20) == 20
    assert g((20, 2.0), (40, 1.0), 0.5, 20) == 20
    assert g((20, 1.0), (40, 0.55), 0.5, 20) == 5
    assert 5 <= g((20, 1.0), (40, 0.6), 0.5, 20) < 20
    assert g((20, 1.0), (40, 1.2), 0.5, 20) == 20
