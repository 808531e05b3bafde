"""Shape-specialised DCFR up walks (HIP source generated per segment shape at
engine creation and compiled with hiprtc): the arithmetic of the generic
register walk, operation for operation, so a solve that runs them is
bit-identical to the oracle in exact precision and to the generic walk in
either precision. PTS_JIT_MIN_SEGMENTS=1 puts every eligible segment (a walk
of 2 to 80 members that is not a spot's root) on a specialised kernel."""
import numpy as np
import pytest

import pyoracle as O
from pts import spots
from pts.native import EXACT, FAST
from pts.solver import Solver

from test_host_cpu import iso_turn

pytestmark = pytest.mark.gpu


def same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and np.array_equal(a, b)


def describe(a, b):
    d = np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64))
    return f"{int((d > 0).sum())} of {d.size} differ, max {d.max(initial=0):.3e}"


def specialise(monkeypatch, on):
    if on:
        monkeypatch.delenv("PTS_JIT", raising=False)
        monkeypatch.setenv("PTS_JIT_MIN_SEGMENTS", "1")
    else:
        monkeypatch.setenv("PTS_JIT", "0")


# (name, spot, iterations, scratch budget): the turn's river segments of up to
# 69 members, a turn with 33- to 99-member segments (those above 80 keep the
# generic walk), suit isomorphism, chunk groups, and a flop in groups of turn
# subtrees
CASES = [("turn0", lambda: spots.turn_spot(0), 6, 0),
         ("turn-wide", lambda: spots.batch_spot("turn", 1), 3, 0),
         ("iso-turn", iso_turn, 8, 0),
         ("grouped-turn", lambda: spots.turn_spot(0), 6, 16 << 20),
         ("small-flop", lambda: spots.small_flop_spot(0), 4, 4 << 20)]


@pytest.mark.parametrize("name,make,iters,budget", CASES, ids=[c[0] for c in CASES])
def test_specialised_walks_bit_exact(monkeypatch, name, make, iters, budget):
    specialise(monkeypatch, True)
    cfg = make()
    g, c = Solver(cfg, precision=EXACT, scratch_budget=budget), O.OracleSolver(cfg)
    assert g.jit_segments() > 0
    g.iterate(iters)
    c.iterate(iters)
    rg, sg = g.storage()
    rc, sc = c.storage()
    assert same(rg, rc), describe(rg, rc)
    assert same(sg, sc), describe(sg, sc)
    assert g.exploitability() == c.exploitability_f64()


def test_specialised_walks_equal_the_generic_walk(monkeypatch):
    cfg = spots.turn_spot(0)
    specialise(monkeypatch, True)
    a = Solver(cfg, precision=FAST)
    specialise(monkeypatch, False)
    b = Solver(cfg, precision=FAST)
    assert a.jit_segments() > 0 and b.jit_segments() == 0
    a.iterate(30)
    b.iterate(30)
    ra, sa = a.storage()
    rb, sb = b.storage()
    assert same(ra, rb), describe(ra, rb)
    assert same(sa, sb), describe(sa, sb)
    assert a.exploitability() == b.exploitability()



def test_small_plans_keep_the_generic_walks(monkeypatch):
    # the turn spot has 721 segments with a walk, under the 8192 a plan needs
    # before its shapes are compiled; forcing a per-shape count compiles them
    monkeypatch.delenv("PTS_JIT", raising=False)
    monkeypatch.delenv("PTS_JIT_MIN_SEGMENTS", raising=False)
    assert Solver(spots.turn_spot(0), precision=FAST).jit_segments() == 0
    monkeypatch.setenv("PTS_JIT_MIN_SEGMENTS", "64")
    assert Solver(spots.turn_spot(0), precision=FAST).jit_segments() > 0


def compressed(cfg):
    # 16-bit storage (storage_mode 1: i16 regrets, u16 sums, per-node scales)
    cfg.storage_mode = 1
    return cfg


# 16-bit storage runs each DCFR up walk twice (the largest new regret of every
# own node, then the new regrets encoded with it); the specialised kernels of
# both passes and of the best response from 16-bit storage, against the
# oracle's 16-bit solve bit for bit, from zero storage and from random 16-bit
# storage (every path of the codec: integers over their whole range, some
# scales 0, frozen nodes none)
@pytest.mark.parametrize("name,make,iters", [("turn0", lambda: spots.turn_spot(0), 6),
                                             ("small-flop", lambda: spots.small_flop_spot(0), 3)],
                         ids=["turn0", "small-flop"])
def test_specialised_16bit_walks_bit_exact(monkeypatch, name, make, iters):
    specialise(monkeypatch, True)
    cfg = compressed(make())
    g, c = Solver(cfg, precision=EXACT), O.OracleSolver(cfg)
    assert g.compressed and c.compressed and g.jit_segments() > 0
    g.iterate(iters)
    c.iterate(iters)
    for x, y, what in zip(g.storage_q(), c.storage_q(), ("regrets", "sums", "regret scales", "sum scales")):
        assert same(x, y), f"{what}: {describe(x, y)}"
    assert g.exploitability() == c.exploitability_f64()


def test_specialised_16bit_walks_from_random_storage(monkeypatch):
    specialise(monkeypatch, True)
    cfg = compressed(spots.turn_spot(1))
    g, c = Solver(cfg, precision=EXACT), O.OracleSolver(cfg)
    assert g.jit_segments() > 0
    n, k = c.storage_len, c.stats().player_nodes
    rng = np.random.default_rng(11)
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
