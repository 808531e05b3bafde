"""Batched spots (pts_batch_*): many independent trees merged into one plan
and iterated by one hipGraph launch. A spot's arithmetic is the same as
alone, so in exact precision every spot's storage after N batched
iterations equals the single-spot GPU solve's and the oracle's bit for bit."""
import numpy as np
import pytest

import pyoracle as O
from pts import spots
from pts.native import EXACT, FAST
from pts.solver import BatchSolver, Solver, SolveOptions


def small_batch():
    # mixed streets and hand counts, one set of DCFR parameters; the last
    # spot's chance nodes merge suit-isomorphic cards
    from test_host_cpu import iso_turn
    return [spots.river_spot(0), spots.turn_spot(1), spots.river_spot(2), spots.batch_spot("river", 3),
            spots.batch_spot("turn", 0), spots.river_spot(5), iso_turn()]


def test_batch_plan_merges_and_checks_on_the_host():
    cfgs = small_batch()
    b = BatchSolver(cfgs, device=-1)
    nodes, cells, s32, s64 = b.host_info()
    single = [Solver(c, device=-1) for c in cfgs]
    assert cells == sum(s.storage_len for s in single)
    assert nodes == sum(s.stats().player_nodes + s.stats().chance_nodes + s.stats().terminal_nodes for s in single)
    assert 0 < s32 < s64


def test_batch_workload_spots_are_seeded_and_varied():
    a, b = spots.batch_spot("river", 7), spots.batch_spot("river", 7)
    assert a.board == b.board and a.ranges == b.ranges and a.root.stack == b.root.stack
    cs = [spots.batch_spot("river", i) for i in range(40)]
    n = [sum(1 for x in c.ranges[0] if x > 0) for c in cs]
    assert min(n) >= 80 and max(n) <= 1326 and len(set(n)) > 30
    spr = [c.root.stack[0] / c.root.pot for c in cs]
    assert min(spr) >= 0.5 - 1e-9 and max(spr) <= 10 + 1e-9
    sizes = {len(c.menu.river.bet) - 1 for c in cs}
    assert sizes == {1, 2, 3}


@pytest.mark.gpu
def test_batch_iterations_bit_exact_per_spot():
    cfgs = small_batch()
    b = BatchSolver(cfgs, precision=EXACT)
    b.iterate(8)
    xb = b.exploitability()
    for k, c in enumerate(cfgs):
        s = Solver(c, precision=EXACT)
        s.iterate(8)
        rb, sb = b.storage(k)
        rs, ss = s.storage()
        assert np.array_equal(rb, rs) and np.array_equal(sb, ss), k
        assert xb[k] == s.exploitability(), k
    # and the oracle for one river and one turn spot
    for k in (0, 1):
        o = O.OracleSolver(cfgs[k])
        o.iterate(8)
        rb, sb = b.storage(k)
        rc, sc = o.storage()
        assert np.array_equal(rb, rc) and np.array_equal(sb, sc), k


@pytest.mark.gpu
def test_batch_solve_freezes_each_spot_at_its_target():
    # every spot keeps the single solve's check schedule (lib.rs:155-172,
    # 594-642), so it stops at the same check with the same exploitability
    # as the same spot solved alone, and a frozen spot's storage is that of
    # the spot alone after the same iterations
    cfgs = small_batch()
    b = BatchSolver(cfgs, precision=FAST)
    pr = b.solve(400, 1.0, 10)
    for k, p in enumerate(pr):
        assert p.finished
        assert p.exploitability_pct <= 1.0 or p.iteration == 400
        s = Solver(cfgs[k], precision=FAST)
        alone = s.solve(SolveOptions(400, 1.0, 10))
        assert (p.iteration, p.exploitability) == (alone.iteration, alone.exploitability), k
        rb, sb = b.storage(k)
        rs, ss = s.storage()
        assert np.array_equal(rb, rs) and np.array_equal(sb, ss), k


@pytest.mark.gpu
def test_full_batch_workload_bit_exact():
    # the batched workload at its size: 256 river and 64 turn spots in one
    # plan, exact precision; every spot equals its own solve after the same
    # iterations, and a sample (every 16th river spot, every 8th turn spot)
    # equals the oracle bit for bit
    cfgs = spots.batch_spots()
    assert len(cfgs) == 320
    b = BatchSolver(cfgs, precision=EXACT)
    b.iterate(4)
    xb = b.exploitability()
    for k, c in enumerate(cfgs):
        s = Solver(c, precision=EXACT)
        s.iterate(4)
        rb, sb = b.storage(k)
        rs, ss = s.storage()
        assert np.array_equal(rb, rs) and np.array_equal(sb, ss), k
        assert xb[k] == s.exploitability(), k
        if (k < 256 and k % 16 == 0) or (k >= 256 and (k - 256) % 8 == 0):
            o = O.OracleSolver(c)
            o.iterate(4)
            rc, sc = o.storage()
            assert np.array_equal(rb, rc) and np.array_equal(sb, sc), k
            assert xb[k] == o.exploitability_f64(), k
        del s


@pytest.mark.gpu
def test_autotune_keeps_the_arithmetic():
    cfgs = small_batch()
    a, b = BatchSolver(cfgs, precision=FAST), BatchSolver(cfgs, precision=FAST)
    (rl, dt), ms = b.autotune(2)
    assert rl in (1, 2, 4, 8, 16, 32, 64, 256) and dt in (128, 256, 512, 1024) and ms > 0
    assert b.iterations == 0
    a.iterate(6)
    b.iterate(6)
    for k in range(len(cfgs)):
        ra, sa = a.storage(k)
        rb, sb = b.storage(k)
        assert np.array_equal(ra, rb) and np.array_equal(sa, sb), k


@pytest.mark.gpu
def test_down_workgroup_sizes_keep_the_exact_bits():
    # the down pass's workgroup size (128 threads for levels of many
    # segments, such as the flop's river level) changes only how the
    # (node, hand tile) units are dealt to waves
    cfgs = small_batch()
    ref = None
    for dt in (128, 256, 512, 1024):
        b = BatchSolver(cfgs, precision=EXACT)
        b.set_tiles(4, dt)
        b.iterate(5)
        got = [b.storage(k) for k in range(len(cfgs))]
        if ref is None:
            ref = got
        else:
            for k, ((ra, sa), (rb, sb)) in enumerate(zip(ref, got)):
                assert np.array_equal(ra, rb) and np.array_equal(sa, sb), (dt, k)
        del b


@pytest.mark.gpu
def test_terminal_run_lengths_keep_the_fast_bits():
    # a run only decides which terminals one workgroup sums in sequence (and
    # loads a board's tables for): the fast values are the same bits for runs
    # of 1, 16 and 256 terminals
    cfgs = small_batch()
    ref = None
    for rl in (1, 16, 256):
        b = BatchSolver(cfgs, precision=FAST)
        b.set_tiles(rl, 256)
        b.iterate(6)
        got = [b.storage(k) for k in range(len(cfgs))]
        if ref is None:
            ref = got
        else:
            for k, ((ra, sa), (rb, sb)) in enumerate(zip(ref, got)):
                assert np.array_equal(ra, rb) and np.array_equal(sa, sb), (rl, k)
        del b
