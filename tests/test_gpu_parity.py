"""GPU parity: the HIP engine against the oracle (the reference CPU solve
restated), on the same seeded inputs. In exact precision the engine repeats
the reference's operation order, so regrets, strategy sums, values and
exploitability must be bit-identical (compared as floats, so +0 == -0). Fast
precision (block-scan terminal sums) is checked against the same oracle with
tolerances."""
import numpy as np
import pytest

import pyoracle as O
from pts import spots, PostflopMenu, make_config, parse_cards, combo_index, DcfrParams
from pts.solver import Engine, Solver, SolveOptions
from pts.native import EXACT, FAST

from test_host_cpu import iso_flop, iso_turn, oracle_tree

pytestmark = pytest.mark.gpu


def random_storage(n, seed):
    rng = np.random.default_rng(seed)
    r = (rng.standard_normal(n) * 3).astype(np.float32)
    r[rng.random(n) < 0.3] = 0.0
    s = (rng.random(n) * 5).astype(np.float32)
    s[rng.random(n) < 0.2] = 0.0
    return r, s


def same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and np.array_equal(a, b)


def describe(a, b):
    d = np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64))
    return f"{int((d > 0).sum())} of {d.size} differ, max {d.max(initial=0):.3e}"


SPOTS = [("river0", lambda: spots.river_spot(0)), ("river5", lambda: spots.river_spot(5)),
         ("turn0", lambda: spots.turn_spot(0))]


@pytest.mark.parametrize("name,make", SPOTS, ids=[s[0] for s in SPOTS])
def test_traversals_from_random_storage(name, make):
    cfg = make()
    o, t = oracle_tree(cfg)
    r, s = random_storage(o.storage_len, 11)
    o.set_storage(r, s)
    e = Engine(t, r, s, precision=EXACT)
    for p in (0, 1):
        for mode in (1, 2):  # FP64 best response and average
            g = e.traverse(p, mode)
            c = o.traverse_root_f64(p, mode)
            assert same(g, c), f"p{p} mode{mode}: {describe(g, c)}"
    # one CFR traversal per player at t = 3 (updates storage)
    d = O.discount(3)
    for p in (0, 1):
        g = e.traverse(p, 0, (d[0], d[1], d[2], 1))
        c = o.traverse_root(p, 0, 3)
        assert same(g.astype(np.float32), c), f"cfr p{p}: {describe(g, c)}"
        rg, sg = e.download()
        rc, sc = o.storage()
        assert same(rg, rc), f"regrets after p{p}: {describe(rg, rc)}"
        assert same(sg, sc), f"sums after p{p}: {describe(sg, sc)}"


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_river_solves_bit_exact(seed):
    cfg = spots.river_spot(seed)
    g, c = Solver(cfg, precision=EXACT), O.OracleSolver(cfg)
    g.iterate(40)
    c.iterate(40)
    rg, sg = g.storage()
    rc, sc = c.storage()
    assert same(rg, rc), describe(rg, rc)
    assert same(sg, sc), describe(sg, sc)
    assert g.exploitability() == c.exploitability_f64()


def test_turn_solve_bit_exact():
    cfg = spots.turn_spot(0)
    g, c = Solver(cfg, precision=EXACT), O.OracleSolver(cfg)
    g.iterate(6)
    c.iterate(6)
    rg, sg = g.storage()
    rc, sc = c.storage()
    assert same(rg, rc), describe(rg, rc)
    assert same(sg, sc), describe(sg, sc)
    eg, ec = g.exploitability(), c.exploitability_f64()
    assert eg == ec, (eg, ec)
    # the reference's own f32 best response agrees to far inside 0.01% of the pot
    assert abs(eg - c.exploitability()) < 1e-4 * cfg.root.pot


@pytest.mark.parametrize("kind,index", [("river", 11), ("river", 1), ("turn", 1)])
def test_large_segments_bit_exact(kind, index):
    # Segment sizes select the walk: up to 32, 48 or 64 members in registers,
    # more than 64 in LDS. The river trees here are single segments of 57 and
    # 99 members; the turn tree's river segments span all four sizes (672 of
    # at most 32, 192 of 33-48, 192 of 49-64, 49 of 65-69 members).
    cfg = spots.batch_spot(kind, index)
    g, c = Solver(cfg, precision=EXACT), O.OracleSolver(cfg)
    n = 8 if kind == "river" else 3
    g.iterate(n)
    c.iterate(n)
    rg, sg = g.storage()
    rc, sc = c.storage()
    assert same(rg, rc), describe(rg, rc)
    assert same(sg, sc), describe(sg, sc)
    assert g.exploitability() == c.exploitability_f64()


@pytest.mark.parametrize("make", [iso_turn, iso_flop], ids=["iso-turn", "iso-flop"])
def test_isomorphism_bit_exact(make):
    # chance nodes merge suit-isomorphic cards (iso.rs): the chance sums add
    # the canonical child's row at the permuted hands after the canonical
    # children, in the oracle's order
    cfg = make()
    g, c = Solver(cfg, precision=EXACT), O.OracleSolver(cfg)
    g.iterate(10)
    c.iterate(10)
    rg, sg = g.storage()
    rc, sc = c.storage()
    assert same(rg, rc), describe(rg, rc)
    assert same(sg, sc), describe(sg, sc)
    assert g.exploitability() == c.exploitability_f64()
    f = Solver(cfg)  # fast precision: the same chance kernel
    f.iterate(10)
    assert abs(f.exploitability() - c.exploitability_f64()) <= 1e-4 * cfg.root.pot


def compressed(cfg):
    cfg.storage_mode = 1  # 16-bit storage (storage.rs:87-106)
    return cfg


Q_SPOTS = [("river0", lambda: spots.river_spot(0), 12), ("turn0", lambda: spots.turn_spot(0), 4),
           ("iso-turn", iso_turn, 6), ("river-big", lambda: spots.batch_spot("river", 11), 6)]


@pytest.mark.parametrize("name,make,iters", Q_SPOTS, ids=[s[0] for s in Q_SPOTS])
def test_compressed_solves_bit_exact(name, make, iters):
    # i16 regrets and u16 sums with per-node scales: every node's new scale is
    # its block's largest |value| over all hands (the down pass reduces over
    # its workgroup, the up pass over the hand tiles in a first walk), then
    # the block is re-encoded in a second walk
    cfg = compressed(make())
    g, c = Solver(cfg, precision=EXACT), O.OracleSolver(cfg)
    assert g.compressed and c.compressed
    g.iterate(iters)
    c.iterate(iters)
    for x, y, what in zip(g.storage_q(), c.storage_q(), ("regrets", "sums", "regret scales", "sum scales")):
        assert same(x, y), f"{what}: {describe(x, y)}"
    assert g.exploitability() == c.exploitability_f64()
    rg, sg = g.storage()  # decoded (storage.rs:209-246)
    rc, sc = c.storage()
    assert same(rg, rc) and same(sg, sc)


def test_compressed_traversals_from_random_storage():
    # every code path of the codec from arbitrary 16-bit storage: integers
    # over their whole range, some blocks with scale 0
    cfg = compressed(spots.turn_spot(1))
    g, c = Solver(cfg, precision=EXACT), O.OracleSolver(cfg)
    n, k = c.storage_len, c.stats().player_nodes
    rng = np.random.default_rng(5)
    r = rng.integers(-32768, 32768, n).astype(np.int16)
    s = rng.integers(0, 65536, n).astype(np.uint16)
    rs = (rng.random(k) * 50).astype(np.float32)
    ss = (rng.random(k) * 20).astype(np.float32)
    rs[rng.random(k) < 0.1] = 0.0
    ss[rng.random(k) < 0.1] = 0.0
    g.set_storage_q(r, s, rs, ss)
    c.set_storage_q(r, s, rs, ss)
    assert g.exploitability() == c.exploitability_f64()
    g.iterations = 5
    c.iterations = 5
    g.iterate(2)
    c.iterate(2)
    for x, y, what in zip(g.storage_q(), c.storage_q(), ("regrets", "sums", "regret scales", "sum scales")):
        assert same(x, y), f"{what}: {describe(x, y)}"
    qa, qb = g.query([]), c.query([])
    assert same(qa.strategy, qb.strategy) and same(qa.ev, qb.ev), describe(qa.ev, qb.ev)


def test_compressed_engine_layer_traversals():
    # pts_engine_create_q (HostStorage::Compressed behind the engine boundary):
    # one traversal per mode from random 16-bit storage, against the oracle
    cfg = compressed(spots.turn_spot(0))
    o, t = oracle_tree(cfg)
    nodes = o.nodes()
    slot = np.where(nodes["kind"] <= 1, nodes["slot"], 0).astype(np.uint32)
    n, k = o.storage_len, o.stats().player_nodes
    rng = np.random.default_rng(9)
    r = rng.integers(-32768, 32768, n).astype(np.int16)
    s = rng.integers(0, 65536, n).astype(np.uint16)
    rs = (rng.random(k) * 30).astype(np.float32)
    ss = (rng.random(k) * 10).astype(np.float32)
    ss[rng.random(k) < 0.1] = 0.0
    o.set_storage_q(r, s, rs, ss)
    e = Engine(t, None, None, precision=EXACT, q16=(slot, r, s, rs, ss))
    for p in (0, 1):
        for mode in (1, 2):
            assert same(e.traverse(p, mode), o.traverse_root_f64(p, mode)), f"p{p} mode{mode}"
    d = O.discount(4)
    for p in (0, 1):
        g = e.traverse(p, 0, (d[0], d[1], d[2], 1))
        c = o.traverse_root(p, 0, 4)
        assert same(g.astype(np.float32), c), f"cfr p{p}: {describe(g, c)}"
        for x, y in zip(e.download_q(), o.storage_q()):
            assert same(x, y), describe(x, y)


def test_compressed_queries_and_fast_precision():
    cfg = compressed(spots.turn_spot(2))
    g, c = Solver(cfg, precision=EXACT), O.OracleSolver(cfg)
    g.iterate(8)
    c.iterate(8)
    path = [c.query([]).actions[0]]
    path = path + [c.query(path).actions[0]]
    card = next(x for x in range(52) if x not in cfg.board)
    for pth in ([], path[:1], path + [("card", card)]):
        a, b = g.query(pth), c.query(pth)
        assert a.actions == b.actions
        assert same(a.strategy, b.strategy) and same(a.reach, b.reach)
        assert same(a.ev, b.ev), describe(a.ev, b.ev)
        ra, rb = g.ranges_after(pth), c.ranges_after(pth)
        assert same(ra[0], rb[0]) and same(ra[1], rb[1])
    f = Solver(cfg)  # fast precision terminal sums, same storage kernels
    f.iterate(8)
    assert f.compressed
    assert abs(f.exploitability() - c.exploitability_f64()) <= 1e-3 * cfg.root.pot
    # the f32 solve it approximates (tests.rs:273-302)
    plain = spots.turn_spot(2)
    p32 = Solver(plain, precision=EXACT)
    p32.iterate(8)
    assert np.max(np.abs(p32.query([]).range_strategy() - g.query([]).range_strategy())) < 0.03


def test_dumped_trainer_spot_bit_exact():
    # a spot in the trainer's POKER_TRAINER_DUMP format (review.rs:1176-1213),
    # facing a bet mid-street with rake, replayed
    from pts.dump import read_dump
    from test_dump import TURN_FACING
    _, cfg = read_dump(TURN_FACING)
    g, c = Solver(cfg, precision=EXACT), O.OracleSolver(cfg)
    g.iterate(6)
    c.iterate(6)
    rg, sg = g.storage()
    rc, sc = c.storage()
    assert same(rg, rc) and same(sg, sc), describe(rg, rc)
    assert g.exploitability() == c.exploitability_f64()


@pytest.mark.parametrize("storage", [0, 1], ids=["f32", "16bit"])
def test_sampled_flop_bit_exact(storage):
    # chance-sampled mode (sampling.rs; gpu/testutil.rs:120-131 and its
    # compressed variant): 3 turn and 2 river cards per branch, all-in
    # runouts exact with isomorphism
    from test_sampled import sampled_flop
    cfg = sampled_flop()
    cfg.storage_mode = storage
    g, c = Solver(cfg, precision=EXACT), O.OracleSolver(cfg)
    g.iterate(8)
    c.iterate(8)
    if storage:
        for x, y in zip(g.storage_q(), c.storage_q()):
            assert same(x, y), describe(x, y)
    else:
        rg, sg = g.storage()
        rc, sc = c.storage()
        assert same(rg, rc) and same(sg, sc), describe(rg, rc)
    assert g.exploitability() == c.exploitability_f64()
    q = [c.query([]).actions[0]]
    q = q + [c.query(q).actions[0]]
    assert q == ["check", "check"]
    nodes = c.nodes()
    chance = nodes["first_child"][nodes["first_child"][0]]  # check, check: the turn's chance node
    turn = int(nodes["edge_card"][nodes["first_child"][chance]])  # its first sampled card
    for pth in ([], q + [("card", turn)]):
        a, b = g.query(pth), c.query(pth)
        assert same(a.strategy, b.strategy) and same(a.ev, b.ev), describe(a.ev, b.ev)


def test_solve_loop_matches_the_oracle():
    cfg = spots.river_spot(4)
    g, c = Solver(cfg, precision=EXACT), O.OracleSolver(cfg)
    hg = []
    pg = g.solve(SolveOptions(200, 0.5, 10), hg)
    pc, hc = c.solve(200, 0.5, 10)
    # the checks land on the same iterations; exploitability is FP64 on the GPU
    assert [h.iteration for h in hg] == [h.iteration for h in hc]
    for a, b in zip(hg, hc):
        assert abs(a.exploitability - b.exploitability) < 1e-4 * cfg.root.pot


@pytest.mark.parametrize("name,make", SPOTS[::2], ids=[s[0] for s in SPOTS[::2]])
def test_traversals_from_extreme_storage(name, make):
    # gpu/tests.rs:101-120,143-150: regret matching and averaging where a
    # hand's positive regrets (or sums) are all zero, sum to a subnormal whose
    # reciprocal overflows, to a subnormal that still inverts, or overflow to
    # infinity; the guard (cfr.rs:113-121) makes each of these uniform or exact.
    # An own node's value is the sum of r+ * cfv times the reciprocal
    # (cfr.rs:315-382), so an f32::MAX regret overflows it to infinity and
    # the regret update then writes infinities and NaNs: both sides must agree
    # on those too.
    cfg = make()
    o, t = oracle_tree(cfg)
    n = o.storage_len
    rng = np.random.default_rng(5)
    sub, sub2, big = (np.array([1, 0x00400000, 0x7F7FFFFF], np.uint32).view(np.float32))
    rpool = np.array([-1.0, -0.0, 0.0, sub, sub2, 1.0, big, 1e30, 1e-40, -1e-30, -3.0, 2.0], np.float32)
    spool = np.array([0.0, sub, sub2, big, 1.0, 1e30, 3.0], np.float32)
    r, s = rpool[rng.integers(0, len(rpool), n)], spool[rng.integers(0, len(spool), n)]
    o.set_storage(r, s)
    e = Engine(t, r, s, precision=EXACT)

    def eq(a, b):
        a, b = np.asarray(a), np.asarray(b)
        return a.shape == b.shape and np.array_equal(a, b, equal_nan=True)
    for p in (0, 1):
        for mode in (1, 2):  # FP64 best response and average
            g, c = e.traverse(p, mode), o.traverse_root_f64(p, mode)
            assert eq(g, c), f"p{p} mode{mode}: {describe(g, c)}"
    d = O.discount(3)
    for p in (0, 1):
        g = e.traverse(p, 0, (d[0], d[1], d[2], 1))
        c = o.traverse_root(p, 0, 3)
        assert eq(g.astype(np.float32), c), f"cfr p{p}: {describe(g, c)}"
        (rg, sg), (rc, sc) = e.download(), o.storage()
        assert eq(rg, rc) and eq(sg, sc), f"storage after p{p}: {describe(rg, rc)}"


def test_solves_continue_across_calls_and_from_the_oracle():
    # gpu/engine/tests.rs:700-711: solving to 3 and then to 6 iterations leaves
    # the storage of one solve to 6; gpu/engine/tests.rs:274-286: a GPU solve
    # continues a CPU solve (here the oracle's, bit for bit)
    cfg = spots.turn_spot(1)
    once, twice = Solver(cfg, precision=EXACT), Solver(cfg, precision=EXACT)
    once.solve(SolveOptions(6, 0.0, 25))
    twice.solve(SolveOptions(3, 0.0, 25))
    p = twice.solve(SolveOptions(6, 0.0, 25))
    assert p.iteration == 6 and twice.iterations == 6
    (r1, s1), (r2, s2) = once.storage(), twice.storage()
    assert same(r1, r2) and same(s1, s2), describe(r1, r2)
    c = O.OracleSolver(cfg)
    c.iterate(3)
    g = Solver(cfg, precision=EXACT)
    g.set_storage(*c.storage())
    g.iterations = 3
    g.solve(SolveOptions(6, 0.0, 25))
    c.iterate(3)
    (rg, sg), (rc, sc) = g.storage(), c.storage()
    assert g.iterations == 6
    assert same(rg, rc) and same(sg, sc), describe(rg, rc)
    assert same(rg, r1) and same(sg, s1)


def test_queries_match_the_oracle():
    cfg = spots.turn_spot(2)
    g, c = Solver(cfg, precision=EXACT), O.OracleSolver(cfg)
    g.iterate(8)
    c.iterate(8)
    qg, qc = g.query([]), c.query([])
    assert qg.actions == qc.actions
    assert same(qg.strategy, qc.strategy)
    assert same(qg.ev, qc.ev), describe(qg.ev, qc.ev)
    path = [qc.actions[0]]
    qc2 = c.query(path)
    path2 = path + [qc2.actions[0]]
    card = next(x for x in range(52) if x not in cfg.board)
    for pth in (path, path2 + [("card", card)]):
        a, b = g.query(pth), c.query(pth)
        assert a.actions == b.actions
        assert same(a.strategy, b.strategy) and same(a.reach, b.reach)
        assert same(a.ev, b.ev), describe(a.ev, b.ev)
        ra, rb = g.ranges_after(pth), c.ranges_after(pth)
        assert same(ra[0], rb[0]) and same(ra[1], rb[1])


def test_queries_through_isomorphic_cards():
    # lib.rs:746-825 with iso.rs: a river card merged onto its suit-swapped
    # canonical card is answered from the canonical subtree through the hand
    # permutation, as the oracle does; and it agrees with a solve that merges
    # nothing (tests.rs:234-271)
    cfg = iso_turn()
    g, c = Solver(cfg, precision=EXACT), O.OracleSolver(cfg)
    g.iterate(8)
    c.iterate(8)
    cfg_plain = iso_turn()
    cfg_plain.isomorphism = False
    plain = Solver(cfg_plain, precision=EXACT)
    plain.iterate(8)
    path = [c.query([]).actions[0]]
    path = path + [c.query(path).actions[0]]
    assert path[0] == "check" and path[1] == "check"
    # spades are merged onto diamonds at the river: the merged tree is smaller
    assert g.stats().player_nodes < plain.stats().player_nodes
    for card in range(52):
        if card in cfg.board:
            continue
        pth = path + [("card", card)]
        a, b = g.query(pth), c.query(pth)
        assert a.actions == b.actions
        assert same(a.strategy, b.strategy) and same(a.reach, b.reach)
        assert same(a.ev, b.ev), describe(a.ev, b.ev)
        ra, rb = g.ranges_after(pth), c.ranges_after(pth)
        assert same(ra[0], rb[0]) and same(ra[1], rb[1])
        p = plain.query(pth)
        assert np.max(np.abs(a.strategy - p.strategy)) < 2e-3
        assert np.max(np.abs(a.ev - p.ev)) < 0.5
        pr = plain.ranges_after(pth)
        assert np.max(np.abs(ra[0] - pr[0])) < 1e-6 and np.max(np.abs(ra[1] - pr[1])) < 1e-6


REF = DcfrParams.reference_default()


def combo(s):
    c = parse_cards(s)
    return combo_index(c[0], c[1])


def test_polarized_toy_game_on_the_gpu():
    # crates/solver/src/tests.rs:42-88
    cfg = make_config("Kc8d5h3s2c", "AA, JTs, J9s:0.5", "KQ", 1000, 5000, PostflopMenu.one(["100%"], []), dcfr=REF)
    s = Solver(cfg, precision=EXACT)
    p = s.solve(SolveOptions(2000, 0.05, 25))
    assert p.exploitability_pct < 0.2
    root = s.query([])
    ev = root.hand_ev(root.hand_index(combo("AsAh")))
    assert abs(ev[1] - 1500) < 20 and abs(ev[0] - 1000) < 20
    facing = s.query([("raise", 1000)])
    assert abs(facing.range_strategy()[1] - 0.5) < 0.05


def test_all_in_equity_on_the_gpu():
    # crates/solver/src/tests.rs:92-104
    cfg = make_config("2c7d9hTs", "AhAd", "KhKd", 1000, 1000, PostflopMenu.one(["allin"], []), dcfr=REF)
    s = Solver(cfg, precision=EXACT)
    s.solve(SolveOptions(50, 0.0, 25))
    ev = s.query([("raise", 1000)]).hand_ev(0)
    assert abs(ev[1] - (2.0 / 44.0 * 3000.0 - 1000.0)) < 0.5 and abs(ev[0]) < 1e-3


def test_sampled_flop_solve_on_the_gpu():
    # crates/solver/src/tests.rs:150-172: 8 turn and 6 river cards per branch
    # (seed 3) under the default menu, one 66% bet and 3x raise after the flop;
    # the solve converges and the root's range strategy is a distribution
    cfg = make_config("Kh7d2c", "AK, KQ, 77, 22, AJs, QJs, T9s, 65s, A5s", "AK, KJ, QQ, JJ, 88, A7s, 98s, JTs, 76s",
                      600, 9700, PostflopMenu(), future_menu=PostflopMenu.one(["66%"], ["3x"]), isomorphism=True,
                      dcfr=REF)
    cfg.sampled = (8, 6, 3)
    assert Solver.estimate(cfg).player_nodes > 0
    s, c = Solver(cfg, precision=EXACT), O.OracleSolver(cfg)
    hg = []
    p = s.solve(SolveOptions(300, 0.5, 25), hg)
    assert p.exploitability_pct < 1.0, p.exploitability_pct
    root = s.query([])
    assert abs(float(np.sum(root.range_strategy())) - 1.0) < 1e-4
    # the exact solve is the oracle's, check for check
    _, hc = c.solve(300, 0.5, 25)
    assert [h.iteration for h in hg] == [h.iteration for h in hc]
    assert same(root.strategy, c.query([]).strategy)


def test_ties_and_rake_on_the_gpu():
    # crates/solver/src/tests.rs:106-127,305-328
    from pts import BettingState
    for rake, expect in ((0.0, 500.0), (0.05, 450.0)):
        cfg = make_config("AsKsQsJsTs", "22", "33", 1500, 5000, PostflopMenu.one(["100%"], ["3x"]), dcfr=REF,
                          rake_rate=rake, rake_cap=300.0 if rake else 0.0)
        cfg.root = BettingState(3, 1500, [0, 500], [5000, 4500], 500, 500, 1, 0, [True, False])
        s, c = Solver(cfg, precision=EXACT), O.OracleSolver(cfg)
        s.iterate(30)
        c.iterate(30)
        rg, sg = s.storage()
        rc, sc = c.storage()
        assert same(rg, rc) and same(sg, sc)
        s.iterate(170)
        ev = s.query([]).hand_ev(0)
        assert abs(ev[0]) < 1e-3 and abs(ev[1] - expect) < 1.0


# ---------------------------------------------------------------- fast precision
@pytest.mark.parametrize("name,make", SPOTS, ids=[s[0] for s in SPOTS])
def test_fast_traversals_close_to_the_oracle(name, make):
    # one traversal from random storage: terminal sums in block-scan order
    # differ from the sequential ones by f64 rounding only
    cfg = make()
    o, t = oracle_tree(cfg)
    r, s = random_storage(o.storage_len, 12)
    o.set_storage(r, s)
    e = Engine(t, r, s, precision=FAST)
    for p in (0, 1):
        for mode in (1, 2):
            g = e.traverse(p, mode)
            c = o.traverse_root_f64(p, mode)
            assert np.allclose(g, c, rtol=1e-12, atol=1e-15), describe(g, c)
    d = O.discount(3)
    for p in (0, 1):
        g = e.traverse(p, 0, (d[0], d[1], d[2], 1))
        c = o.traverse_root(p, 0, 3)
        assert np.allclose(g, c, rtol=1e-5, atol=1e-9), describe(g, c)
        # almost every storage cell is bit-identical after one update
        rg, _ = e.download()
        rc, _ = o.storage()
        assert np.mean(rg == rc) > 0.999, describe(rg, rc)


def _raked_dump():
    from pts.dump import read_dump
    from test_dump import TURN_FACING
    return read_dump(TURN_FACING)[1]


def _raked_chop():
    # crates/solver/src/tests.rs:106-127 with rake: every showdown chops
    from pts import BettingState
    cfg = make_config("AsKsQsJsTs", "22,33,44", "55,66,AA", 1500, 5000, PostflopMenu.one(["100%"], ["3x"]),
                      dcfr=REF, rake_rate=0.05, rake_cap=300.0)
    cfg.root = BettingState(3, 1500, [0, 500], [5000, 4500], 500, 500, 1, 0, [True, False])
    return cfg


def _raked_river():
    cfg = spots.river_spot(0)
    cfg.rake_rate, cfg.rake_cap = 0.05, 300.0
    return cfg


RAKED = [("dump-turn", _raked_dump), ("chop-river", _raked_chop), ("river0", _raked_river)]


@pytest.mark.parametrize("name,make", RAKED, ids=[s[0] for s in RAKED])
def test_fast_traversals_with_rake(name, make):
    # the fast terminal kernel's raked paths: fold payoffs less rake, and the
    # tie sums (eval.rs:347-379) that only a raked showdown reaches
    cfg = make()
    o, t = oracle_tree(cfg)
    r, s = random_storage(o.storage_len, 13)
    o.set_storage(r, s)
    e = Engine(t, r, s, precision=FAST)
    for p in (0, 1):
        for mode in (1, 2):
            g = e.traverse(p, mode)
            c = o.traverse_root_f64(p, mode)
            assert np.allclose(g, c, rtol=1e-12, atol=1e-12), describe(g, c)
    d = O.discount(3)
    for p in (0, 1):
        g = e.traverse(p, 0, (d[0], d[1], d[2], 1))
        c = o.traverse_root(p, 0, 3)
        assert np.allclose(g, c, rtol=1e-5, atol=1e-6), describe(g, c)


def test_fast_solve_tracks_the_exact_solve():
    # crates/solver/src/gpu/engine/tests.rs:493-522 (fast vs exact after a
    # converged solve): the same convergence, close strategies
    cfg = spots.river_spot(1)
    f, x = Solver(cfg, precision=FAST), Solver(cfg, precision=EXACT)
    f.iterate(200)
    x.iterate(200)
    ef, ex = f.exploitability(), x.exploitability()
    assert abs(ef - ex) < 1e-4 * cfg.root.pot, (ef, ex)
    a, b = f.query([]), x.query([])
    ra, rb = a.range_strategy(), b.range_strategy()
    assert np.max(np.abs(ra - rb)) < 0.03


def test_fast_turn_converges_like_the_oracle():
    cfg = spots.turn_spot(0)
    f, c = Solver(cfg, precision=FAST), O.OracleSolver(cfg)
    f.iterate(30)
    c.iterate(30)
    ef, ec = f.exploitability(), c.exploitability_f64()
    # exploitability within 0.01% of the pot after the same iterations
    assert abs(ef - ec) < 1e-4 * cfg.root.pot, (ef, ec)


# ---------------------------------------------------------------- chunk groups
def test_grouped_turn_solve_bit_exact():
    # A 16 MiB scratch budget: the 720 river subtrees run in 5 (DCFR) and 17
    # (FP64 best response) groups that share the scratch rows.
    cfg = spots.turn_spot(0)
    g, c = Solver(cfg, precision=EXACT, scratch_budget=16 << 20), O.OracleSolver(cfg)
    assert g.num_groups() == 5 and g.num_groups(True) == 17
    g.iterate(6)
    c.iterate(6)
    rg, sg = g.storage()
    rc, sc = c.storage()
    assert same(rg, rc), describe(rg, rc)
    assert same(sg, sc), describe(sg, sc)
    assert g.exploitability() == c.exploitability_f64()
    path = [c.query([]).actions[0]]
    path = path + [c.query(path).actions[0]]
    card = next(x for x in range(52) if x not in cfg.board)
    for pth in ([], path + [("card", card)]):
        a, b = g.query(pth), c.query(pth)
        assert same(a.strategy, b.strategy) and same(a.ev, b.ev), describe(a.ev, b.ev)


@pytest.mark.parametrize("seed", [0, 1])
def test_small_flop_solve_bit_exact(seed):
    # flop -> 49 turns -> 48 rivers, in groups of turn subtrees
    cfg = spots.small_flop_spot(seed)
    g, c = Solver(cfg, precision=EXACT, scratch_budget=4 << 20), O.OracleSolver(cfg)
    assert g.num_groups() > 10
    g.iterate(4)
    c.iterate(4)
    rg, sg = g.storage()
    rc, sc = c.storage()
    assert same(rg, rc), describe(rg, rc)
    assert same(sg, sc), describe(sg, sc)
    assert g.exploitability() == c.exploitability_f64()


def test_grouped_compressed_flop_bit_exact():
    # 16-bit storage under chunk groups: every group runs both up walks (the
    # maxima, then the encoding) before its rows are reused, and the regret
    # scales are committed once per traversal
    cfg = spots.small_flop_spot(0)
    cfg.storage_mode = 1
    g, c = Solver(cfg, precision=EXACT, scratch_budget=4 << 20), O.OracleSolver(cfg)
    assert g.compressed and g.num_groups() > 10
    g.iterate(4)
    c.iterate(4)
    for x, y, what in zip(g.storage_q(), c.storage_q(), ("regrets", "sums", "regret scales", "sum scales")):
        assert same(x, y), f"{what}: {describe(x, y)}"
    assert g.exploitability() == c.exploitability_f64()


def test_grouping_does_not_change_the_fast_solve():
    cfg = spots.small_flop_spot(1)
    a, b = Solver(cfg, precision=FAST), Solver(cfg, precision=FAST, scratch_budget=4 << 20)
    assert a.num_groups() == 1 and b.num_groups() > 10
    a.iterate(12)
    b.iterate(12)
    ra, sa = a.storage()
    rb, sb = b.storage()
    assert same(ra, rb) and same(sa, sb)
    assert a.exploitability() == b.exploitability()


def test_queries_from_reader_threads():
    # study.rs:995-1051: queries from other threads after (or during) a
    # solve; each call takes the solver's lock, so concurrent queries return
    # what sequential ones do
    import threading
    cfg = spots.turn_spot(3)
    g = Solver(cfg, precision=EXACT)
    g.iterate(6)
    q0 = [g.query([]).actions[0]]
    paths = [[], q0] + [[("check"), ("check"), ("card", c)] for c in range(52) if c not in cfg.board][:6]
    want = [g.query(p) for p in paths]
    got = [None] * (len(paths) * 3)
    def worker(k):
        got[k] = g.query(paths[k % len(paths)])
    threads = [threading.Thread(target=worker, args=(k,)) for k in range(len(got))]
    for t in threads:
        t.start()
    solver_side = threading.Thread(target=lambda: g.exploitability())
    solver_side.start()
    for t in threads:
        t.join()
    solver_side.join()
    for k, r in enumerate(got):
        w = want[k % len(paths)]
        assert same(r.strategy, w.strategy) and same(r.ev, w.ev) and same(r.reach, w.reach)


def full_range(board):
    used = set(parse_cards(board))
    w = np.zeros(1326, np.float32)
    for c in range(1326):
        hi = 1
        while (hi + 1) * hi // 2 <= c:
            hi += 1
        lo = c - hi * (hi - 1) // 2
        if hi not in used and lo not in used:
            w[c] = 1.0
    return w


EXTREMES = [
    # every combo the board leaves (1081 river hands, 1128 turn hands) on
    # both sides, and one combo against all of them
    ("river-full", "Kc8d5h3s2c", "full", "full", 8),
    ("river-one-vs-full", "Kc8d5h3s2c", "AhAd", "full", 8),
    ("turn-full", "Kc8d5h3s", "full", "full", 2),
]


@pytest.mark.parametrize("name,board,r0,r1,iters", EXTREMES, ids=[e[0] for e in EXTREMES])
def test_extreme_range_sizes_bit_exact(name, board, r0, r1, iters):
    ranges = [full_range(board) if r == "full" else r for r in (r0, r1)]
    cfg = make_config(board, ranges[0], ranges[1], 1000, 4000, PostflopMenu.one(["50%", "allin"], ["3x"], max_raises=2))
    g, c = Solver(cfg, precision=EXACT), O.OracleSolver(cfg)
    n = [int(np.count_nonzero(x)) if not isinstance(x, str) else 1 for x in ranges]
    assert list(g.n) == n
    g.iterate(iters)
    c.iterate(iters)
    rg, sg = g.storage()
    rc, sc = c.storage()
    assert same(rg, rc), describe(rg, rc)
    assert same(sg, sc), describe(sg, sc)
    assert g.exploitability() == c.exploitability_f64()


WIDE = [("river-93", lambda: spots.batch_spot("river", 0)), ("river-75", lambda: spots.batch_spot("river", 7)),
        ("turn-69", lambda: spots.batch_spot("turn", 1))]


@pytest.mark.parametrize("storage", [0, 1], ids=["f32", "16bit"])
@pytest.mark.parametrize("name,make", WIDE, ids=[w[0] for w in WIDE])
def test_wide_segments_bit_exact(name, make, storage):
    # segments of 65 to 80 members: the register walk holding two member
    # records per lane (CFR, the 16-bit passes, the FP64 best response and
    # the root query); the 93-member segment takes the LDS walk
    cfg = make()
    cfg.storage_mode = storage
    g, c = Solver(cfg, precision=EXACT), O.OracleSolver(cfg)
    g.iterate(6)
    c.iterate(6)
    rg, sg = g.storage()
    rc, sc = c.storage()
    assert same(rg, rc) and same(sg, sc), describe(rg, rc)
    assert g.exploitability() == c.exploitability_f64()
    a, b = g.query([]), c.query([])
    assert same(a.strategy, b.strategy) and same(a.ev, b.ev), describe(a.ev, b.ev)
    # below the root: a walk that starts inside the segment (the queried
    # node's span), two actions deep
    path = [b.actions[1]]  # a bet, then the largest raise
    path = path + [c.query(path).actions[-1]]
    for pth in (path[:1], path):
        a, b = g.query(pth), c.query(pth)
        assert a.actions == b.actions
        assert same(a.strategy, b.strategy) and same(a.reach, b.reach)
        assert same(a.ev, b.ev), describe(a.ev, b.ev)


def test_query_facing_a_bet_on_the_river():
    # lib.rs:854-863 traverses each child of the queried node on its own: a
    # fold or showdown child is evaluated by itself (cfr.rs:277-281), not
    # with its river node's batch (river.rs:163-213)
    cfg = spots.river_spot(0)
    g, c = Solver(cfg, precision=EXACT), O.OracleSolver(cfg)
    g.iterate(10)
    c.iterate(10)
    root = c.query([])
    for bet in root.actions[1:]:
        a, b = g.query([bet]), c.query([bet])
        assert a.actions == b.actions and a.actions[0] == "fold"
        assert same(a.ev, b.ev), describe(a.ev, b.ev)
