"""Seeded synthetic spots for the five workloads.

The generator is SplitMix64-seeded xoshiro256** (one stream per
(workload, spot, purpose)); boards come from a Fisher-Yates shuffle of the
52 cards. The reference has no synthetic spots of its own; these follow the
workload definitions:

  river  : 5-card board, pot 10bb, stacks 40bb, ranges U(0,1) per unblocked
           combo with 40% zeroed; bets 50%/100%/all-in, raise 2.5x, 3 bets.
  turn   : 4-card board, pot 20bb, stacks 90bb, ranges U(0,1) with 30% zeroed;
           bets 33%/75%/all-in, raises 2.5x/all-in, one raise (cap 2).
  flop   : 3-card board, pot 6bb, stacks 97bb, ranges from 169-class
           weights sigmoid(score + N(0, 0.5)), bets 33%/75%/all-in,
           raises 2.5x/all-in, 3 bets per street.
Menus use no all-in add/force/merge thresholds (the explicit menu of the
reference's tests, crates/solver/src/tests.rs:6-17).
"""
from __future__ import annotations

import math
from typing import List

from .config import (RIVER, TURN, FLOP, BettingState, DcfrParams, PostflopMenu, SolveConfig, COMBOS)
from .ranges import HandClass

M64 = (1 << 64) - 1


def splitmix64(x: int):
    x = (x + 0x9E3779B97F4A7C15) & M64
    z = x
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return x, z ^ (z >> 31)


class Rng:
    """xoshiro256** seeded by SplitMix64."""

    def __init__(self, seed: int):
        s = []
        x = seed & M64
        for _ in range(4):
            x, z = splitmix64(x)
            s.append(z)
        self.s = s

    def next_u64(self) -> int:
        s = self.s
        r = (((s[1] * 5) & M64) << 7 | ((s[1] * 5) & M64) >> 57) & M64
        r = (r * 9) & M64
        t = (s[1] << 17) & M64
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = ((s[3] << 45) | (s[3] >> 19)) & M64
        return r

    def uniform(self) -> float:
        return (self.next_u64() >> 11) * (1.0 / (1 << 53))

    def below(self, n: int) -> int:
        return self.next_u64() % n

    def normal(self) -> float:
        u1 = max(self.uniform(), 1e-300)
        u2 = self.uniform()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)


PURPOSE = {"board": 1, "range0": 2, "range1": 3, "score": 4, "shape": 5}


def stream(workload: int, spot: int, purpose: str) -> Rng:
    return Rng((workload * 0x100000001B3 + spot * 0x9E3779B1 + PURPOSE[purpose] * 0x85EBCA77) & M64)


def board(rng: Rng, k: int) -> List[int]:
    deck = list(range(52))
    for i in range(51, 0, -1):
        j = rng.below(i + 1)
        deck[i], deck[j] = deck[j], deck[i]
    return deck[:k]


def uniform_range(rng: Rng, board_cards: List[int], p_zero: float) -> List[float]:
    used = 0
    for c in board_cards:
        used |= 1 << c
    w = [0.0] * 1326
    for i, (a, b) in enumerate(COMBOS):
        if used >> a & 1 or used >> b & 1:
            continue
        x = rng.uniform()
        if rng.uniform() < p_zero:
            x = 0.0
        w[i] = x
    return w


def class_range(score: List[float], rng: Rng) -> List[float]:
    w = [0.0] * 1326
    for g in range(169):
        v = 1.0 / (1.0 + math.exp(-(score[g] + 0.5 * rng.normal())))
        for c in HandClass.from_grid_index(g).combos():
            w[c] = v
    return w


def river_menu() -> PostflopMenu:
    return PostflopMenu.one(["50%", "100%", "allin"], ["2.5x"], max_raises=3)


def turn_menu() -> PostflopMenu:
    return PostflopMenu.one(["33%", "75%", "allin"], ["2.5x", "allin"], max_raises=2)


def flop_menu() -> PostflopMenu:
    return PostflopMenu.one(["33%", "75%", "allin"], ["2.5x", "allin"], max_raises=3)


def river_spot(seed: int = 0) -> SolveConfig:
    b = board(stream(0, seed, "board"), 5)
    r = [uniform_range(stream(0, seed, f"range{p}"), b, 0.4) for p in range(2)]
    return SolveConfig(b, r, BettingState.street_start(RIVER, 1000, 4000, 4000), river_menu(),
                       max_memory_bytes=1 << 40, dcfr=DcfrParams())


def turn_spot(seed: int = 0) -> SolveConfig:
    b = board(stream(1, seed, "board"), 4)
    r = [uniform_range(stream(1, seed, f"range{p}"), b, 0.3) for p in range(2)]
    return SolveConfig(b, r, BettingState.street_start(TURN, 2000, 9000, 9000), turn_menu(),
                       max_memory_bytes=1 << 40, dcfr=DcfrParams())


def flop_spot(seed: int = 0) -> SolveConfig:
    b = board(stream(2, seed, "board"), 3)
    sr = stream(2, seed, "score")
    score = [sr.normal() for _ in range(169)]
    r = [class_range(score, stream(2, seed, f"range{p}")) for p in range(2)]
    return SolveConfig(b, r, BettingState.street_start(FLOP, 600, 9700, 9700), flop_menu(),
                       max_memory_bytes=1 << 40, dcfr=DcfrParams())


ITERATIONS = {"river": 1000, "turn": 500, "flop": 300}


def small_flop_spot(seed: int = 0, classes: int = 14) -> SolveConfig:
    """A flop spot small enough for the oracle within seconds: the flop
    board of `flop_spot(seed)`, pot 6bb, stacks 97bb, bet 75% only, no raises
    beyond one, and ranges of the `classes` best-scoring 169-grid classes of
    each player (weights as in `flop_spot`)."""
    b = board(stream(2, seed, "board"), 3)
    sr = stream(2, seed, "score")
    score = [sr.normal() for _ in range(169)]
    keep = sorted(range(169), key=lambda c: -score[c])[:classes]
    r = []
    for p in range(2):
        w = class_range(score, stream(2, seed, f"range{p}"))
        mask = [0.0] * 1326
        for c in keep:
            for k in HandClass.from_grid_index(c).combos():
                mask[k] = 1.0
        r.append([x * m if not _blocked(k, b) else 0.0 for k, (x, m) in enumerate(zip(w, mask))])
    return SolveConfig(b, r, BettingState.street_start(FLOP, 600, 9700, 9700),
                       PostflopMenu.one(["75%"], [], max_raises=1), max_memory_bytes=1 << 40, dcfr=DcfrParams())


def _blocked(k: int, b: List[int]) -> bool:
    a, c = COMBOS[k]
    return a in b or c in b


BATCH_SIZES = ["25%", "33%", "50%", "75%", "100%", "150%"]


def batch_spot(kind: str, index: int) -> SolveConfig:
    """One spot of the batched workload (workload id 3): a river or turn spot
    with a seeded board; per player a range of n ~ U[80, 1326] nonzero combos
    (n capped at the unblocked combos; weights U(0, 1]); stack = SPR x pot
    with SPR log-uniform in [0.5, 10] (pot 10bb on the river, 20bb on the
    turn); 1-3 bet sizes drawn from {25,33,50,75,100,150}% plus all-in,
    raises 2.5x / all-in, 3 bets per street."""
    river = kind == "river"
    spot = index if river else 10000 + index
    b = board(stream(3, spot, "board"), 5 if river else 4)
    used = 0
    for c in b:
        used |= 1 << c
    free = [i for i, (x, y) in enumerate(COMBOS) if not (used >> x & 1 or used >> y & 1)]
    ranges = []
    for p in range(2):
        rng = stream(3, spot, f"range{p}")
        n = min(80 + rng.below(1326 - 80 + 1), len(free))
        pick = list(free)
        for i in range(len(pick) - 1, 0, -1):  # Fisher-Yates, the first n kept
            j = rng.below(i + 1)
            pick[i], pick[j] = pick[j], pick[i]
        w = [0.0] * 1326
        for k in pick[:n]:
            w[k] = 1.0 - rng.uniform()  # (0, 1]
        ranges.append(w)
    sh = stream(3, spot, "shape")
    pot = 1000 if river else 2000
    spr = math.exp(math.log(0.5) + sh.uniform() * (math.log(10.0) - math.log(0.5)))
    stack = max(100, int(round(spr * pot)))
    nsz = 1 + sh.below(3)
    sizes = list(BATCH_SIZES)
    for i in range(len(sizes) - 1, 0, -1):
        j = sh.below(i + 1)
        sizes[i], sizes[j] = sizes[j], sizes[i]
    bets = sorted(sizes[:nsz], key=lambda x: float(x[:-1])) + ["allin"]
    menu = PostflopMenu.one(bets, ["2.5x", "allin"], max_raises=3)
    return SolveConfig(b, ranges, BettingState.street_start(RIVER if river else TURN, pot, stack, stack), menu,
                       max_memory_bytes=1 << 40, dcfr=DcfrParams())


def batch_spots(rivers: int = 256, turns: int = 64) -> List[SolveConfig]:
    """The batched workload: `rivers` river spots then `turns` turn spots."""
    return [batch_spot("river", i) for i in range(rivers)] + [batch_spot("turn", i) for i in range(turns)]
