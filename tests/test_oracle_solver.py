"""Pins of the oracle's solver: the reference's known-answer solves
(crates/solver/src/tests.rs), brute-force card removal
(crates/solver/src/proptests.rs:94-184) and solution invariants
(proptests.rs:224-302)."""
import numpy as np
import pytest

import pyoracle as O
from pts import BettingState, DcfrParams, PostflopMenu, make_config, parse_cards, combo_index

REF_DCFR = DcfrParams.reference_default()


def cfg(board, oop, ip, pot, stack, menu, **kw):
    kw.setdefault("dcfr", REF_DCFR)
    kw.setdefault("isomorphism", True)
    return make_config(board, oop, ip, pot, stack, menu, **kw)


def combo(s):
    c = parse_cards(s)
    return combo_index(c[0], c[1])


def solve(c, iters, target, check=25):
    s = O.OracleSolver(c)
    p, _ = s.solve(iters, target, check)
    return s, p


def test_polarized_river_toy_game():
    # tests.rs:42-88: closed-form equilibrium of the polarized river game
    c = cfg("Kc8d5h3s2c", "AA, JTs, J9s:0.5", "KQ", 1000, 5000, PostflopMenu.one(["100%"], []))
    s, p = solve(c, 2000, 0.05)
    assert p.exploitability_pct < 0.2
    root = s.query([])
    assert root.actions == ["check", ("raise", 1000)]
    aa = root.hand_index(combo("AsAh"))
    assert root.hand_strategy(aa)[1] > 0.98
    ev = root.hand_ev(aa)
    assert abs(ev[1] - 1500) < 20 and abs(ev[0] - 1000) < 20
    from pts import HandClass
    bet = tot = 0.0
    for i in range(len(root.combos)):
        if HandClass(*[x >> 2 for x in __import__("pts").combo_cards(int(root.combos[i]))], False).hi == 9:
            bet += root.reach[i] * root.strategy[1, i]
            tot += root.reach[i]
            e = root.hand_ev(i)
            assert abs(e[0]) < 20 and abs(e[1]) < 40
    assert abs(bet / tot - 0.5) < 0.05
    facing = s.query([("raise", 1000)])
    assert facing.actions == ["fold", "call"]
    assert abs(facing.range_strategy()[1] - 0.5) < 0.05
    kq = facing.hand_index(combo("KdQd"))
    e = facing.hand_ev(kq)
    assert abs(e[0]) < 1e-3 and abs(e[1]) < 30
    r0, _ = s.ranges_after([("raise", 1000)])
    assert r0[combo("AsAh")] > 0.98
    assert 0.2 < r0[combo("JsTs")] < 0.8


def test_all_in_equity_on_turn():
    # tests.rs:92-104: KK has 2 outs among 44 rivers
    c = cfg("2c7d9hTs", "AhAd", "KhKd", 1000, 1000, PostflopMenu.one(["allin"], []))
    s, _ = solve(c, 50, 0.0)
    facing = s.query([("raise", 1000)])
    assert facing.actions == ["fold", "call"]
    ev = facing.hand_ev(0)
    expected = 2.0 / 44.0 * 3000.0 - 1000.0
    assert abs(ev[1] - expected) < 0.5
    assert abs(ev[0]) < 1e-3
    assert facing.hand_strategy(0)[0] > 0.99


def _mid_street(c):
    c.root = BettingState(3, 1500, [0, 500], [5000, 4500], 500, 500, 1, 0, [True, False])
    return c


def test_mid_street_root_with_ties():
    # tests.rs:106-127
    c = _mid_street(cfg("AsKsQsJsTs", "22", "33", 1500, 5000, PostflopMenu.one(["100%"], ["3x"])))
    s, _ = solve(c, 200, 0.0)
    root = s.query([])
    assert root.actions[:2] == ["fold", "call"]
    ev = root.hand_ev(0)
    assert abs(ev[0]) < 1e-3 and abs(ev[1] - 500) < 1.0


def test_rake_reduces_chops():
    # tests.rs:305-328
    c = _mid_street(cfg("AsKsQsJsTs", "22", "33", 1500, 5000, PostflopMenu.one(["100%"], ["3x"]),
                        rake_rate=0.05, rake_cap=300.0))
    s, _ = solve(c, 200, 0.0)
    ev = s.query([]).hand_ev(0)
    assert abs(ev[0]) < 1e-3 and abs(ev[1] - 450) < 1.0


def test_rake_on_folds():
    # tests.rs:331-345
    c = cfg("Kc8d5h3s2c", "AA", "QJs", 1000, 5000, PostflopMenu.one(["100%"], []), rake_rate=0.05, rake_cap=1e9)
    s, _ = solve(c, 300, 0.0)
    root = s.query([])
    ev = root.hand_ev(root.hand_index(combo("AsAh")))
    assert abs(ev[1] - 950) < 2 and abs(ev[0] - 950) < 2


def test_turn_spot_converges():
    # tests.rs:131-149
    c = cfg("Qh9d4c2s", "AQ, KQ, QJs, JTs, 99, 44, T9s, 87s, A5s, KJs",
            "AA, KK, AQ, QT, 99, JJ, TT, 98s, KJs, 65s, AJs", 600, 4000, PostflopMenu())
    s = O.OracleSolver(c)
    p, hist = s.solve(300, 0.3, 20)
    assert p.exploitability_pct < 0.5
    assert hist[0].exploitability_pct > hist[-1].exploitability_pct


def test_too_large_is_rejected():
    c = cfg("Kh7d2c", "any", "any", 600, 9700, PostflopMenu(), max_memory_bytes=1 << 20)
    with pytest.raises(ValueError, match="GB but the limit"):
        O.OracleSolver(c)


def _reach_vec(n, seed):
    # proptests.rs:68-83
    M = (1 << 64) - 1
    x = (seed * 6364136223846793005 + 1442695040888963407) & M
    out = []
    for _ in range(n):
        x = (x * 6364136223846793005 + 1442695040888963407) & M
        v = np.float32(((x >> 33) % 1000)) / np.float32(1000.0)
        out.append(0.0 if v < 0.2 else v)
    return np.asarray(out, np.float32)


def _rand_spot(rng):
    board = list(rng.permutation(52)[:5])
    rs = []
    for _ in range(2):
        w = np.zeros(1326, np.float32)
        from pts import HandClass
        for g in rng.choice(169, size=rng.integers(3, 14), replace=False):
            for c in HandClass.from_grid_index(int(g)).combos():
                w[c] = rng.integers(1, 5) / 4.0
        rs.append(w)
    return board, rs


def _conflict(a, b):
    return len({int(a[0]), int(a[1])} & {int(b[0]), int(b[1])}) > 0


@pytest.mark.parametrize("seed", range(12))
def test_card_removal_matches_brute_force(seed):
    # proptests.rs:94-184: compat / fold / showdown against O(n^2) sums
    rng = np.random.default_rng(seed)
    board, (r0, r1) = _rand_spot(rng)
    hs = O.OracleHandSet(r0, r1, board)
    if min(hs.n) == 0:
        return
    bset = set(int(c) for c in board)
    for p in (0, 1):
        o = p ^ 1
        mine, theirs = hs.cards(p), hs.cards(o)
        r = _reach_vec(hs.n[o], seed * 7 + p)
        compat = hs.compatible_reach(p, r)
        strength = lambda h: O.evaluate(list(board) + [int(h[0]), int(h[1])])  # noqa: E731
        so = [strength(t) for t in theirs]
        w, l, t = 1.7, -2.3, (0.0 if seed % 2 else 0.9)
        sd = hs.eval_showdown(board, p, r, w, l, t)
        sdb, fb = hs.showdown_batch(board, p, r, w, l, t, fold_pay=0.37)
        fold = hs.eval_fold(p, r, 0.37)
        foldb = hs.fold_batch(p, r, 0.37)
        for i, h in enumerate(mine):
            brute_c = sum(float(r[j]) for j, oh in enumerate(theirs) if not _conflict(h, oh))
            assert abs(compat[i] - brute_c) <= 1e-4 * (1 + brute_c)
            assert abs(fold[i] - 0.37 * brute_c) <= 1e-4 * (1 + brute_c)
            assert abs(foldb[i] - 0.37 * brute_c) <= 1e-4 * (1 + brute_c)
            if int(h[0]) in bset or int(h[1]) in bset:
                continue
            sh = strength(h)
            v = 0.0
            for j, oh in enumerate(theirs):
                if int(oh[0]) in bset or int(oh[1]) in bset or _conflict(h, oh):
                    continue
                v += float(r[j]) * (w if sh > so[j] else l if sh < so[j] else t)
            assert abs(sd[i] - v) <= 1e-4 * (1 + abs(v))
            assert abs(sdb[i] - v) <= 1e-4 * (1 + abs(v))
            assert abs(fb[i] - 0.37 * brute_c) <= 1e-4 * (1 + brute_c)


def _river_cfg(rng, two, raise_):
    board, (r0, r1) = _rand_spot(rng)
    bet = ["50%", "100%"] if two else ["75%"]
    rs = ["3x"] if raise_ else []
    from pts import SolveConfig
    pot = int(rng.integers(200, 3000))
    stack = int(rng.integers(100, 6000))
    return SolveConfig(board, [list(r0), list(r1)], BettingState.street_start(3, pot, stack, stack),
                       PostflopMenu.one(bet, rs), dcfr=REF_DCFR, isomorphism=True)


@pytest.mark.parametrize("seed", range(8))
def test_river_solutions_are_well_formed(seed):
    # proptests.rs:224-259
    rng = np.random.default_rng(100 + seed)
    c = _river_cfg(rng, seed % 2 == 0, seed % 3 != 0)
    try:
        s = O.OracleSolver(c)
    except ValueError:
        return
    pr, _ = s.solve(30, 0.0, 30)
    pot = c.root.pot
    assert pr.exploitability >= -1e-3 * pot
    v = s.root_values(2)
    assert abs(v[0] + v[1]) <= 1e-3 * pot
    info = s.query([])
    assert np.all(np.abs(info.strategy.sum(axis=0) - 1) < 1e-3)
    for a, act in enumerate(info.actions):
        assert np.all(np.isfinite(info.ev[a]))
        if act == "fold":
            assert np.all(np.abs(info.ev[a]) < 1e-3)


@pytest.mark.parametrize("seed", range(6))
def test_batched_equals_recursive(seed):
    # proptests.rs:266-302: from identical state, batched and recursive
    # traversals agree to 1e-4 relative
    rng = np.random.default_rng(200 + seed)
    c = _river_cfg(rng, True, True)
    try:
        a = O.OracleSolver(c)
        b = O.OracleSolver(c)
    except ValueError:
        return
    b.set_batch_river(False)
    a.solve(2, 0.0, 2)
    for p in (0, 1):
        r, s = a.storage()
        b.set_storage(r, s)
        for mode in (2, 1, 0):
            va = a.traverse_root(p, mode, 3)
            vb = b.traverse_root(p, mode, 3)
            scale = max(np.abs(va).max(initial=0), np.abs(vb).max(initial=0))
            assert np.all(np.abs(va - vb) <= 1e-4 * (1 + scale))
        ra, sa = a.storage()
        rb, sb = b.storage()
        for x, y in ((ra, rb), (sa, sb)):
            scale = max(np.abs(x).max(initial=0), np.abs(y).max(initial=0))
            assert np.all(np.abs(x - y) <= 1e-4 * (1 + scale))


def test_iso_is_lossless_on_a_monotone_turn():
    # tests.rs:234-271 / proptests.rs:305-327 (short version)
    make = lambda iso: cfg("Ks9s5s2s", "AA, KQ, QJ, T9, 87, 66, A5", "AK, KJ, JT, 98, 76, 55, Q2", 1000, 3000,  # noqa: E731
                           PostflopMenu.one(["66%"], ["3x"]), isomorphism=iso)
    a, b = O.OracleSolver(make(True)), O.OracleSolver(make(False))
    assert a.stats().player_nodes < b.stats().player_nodes
    pa, _ = a.solve(20, 0.0, 20)
    pb, _ = b.solve(20, 0.0, 20)
    assert abs(pa.exploitability - pb.exploitability) < 1e-2 * (1 + abs(pb.exploitability))
    qa, qb = a.query([]), b.query([])
    assert np.max(np.abs(qa.strategy - qb.strategy)) < 2e-3
    for river in ("7c", "7d", "7h", "As", "Ah"):
        path = ["check", "check", ("card", parse_cards(river)[0])]
        qa, qb = a.query(path), b.query(path)
        assert np.max(np.abs(qa.strategy - qb.strategy)) < 2e-3
        assert np.max(np.abs(qa.ev - qb.ev)) < 0.5


def test_f64_best_response_close_to_f32():
    c = cfg("Qh9d4c2sKs", "AQ, KQ, QJs, JTs, 99, 44", "AA, KK, AQ, QT, 99, 98s", 600, 4000, PostflopMenu())
    s = O.OracleSolver(c)
    s.solve(40, 0.0, 40)
    e32, e64 = s.exploitability(), s.exploitability_f64()
    assert abs(e32 - e64) < 1e-4 * c.root.pot
