"""The value-network path: turn subgames solved by the batched GPU DCFR, the
counterfactual-value network trained on them by its own GPU kernels, the
depth-limited flop solve that uses it at the turn, and the exact judge.

A training row is one solved turn subgame: input = the 4-card board mask,
pot / (pot + stack), stack / (pot + stack) and both players' root ranges
normalized to sum 1 (combo order); target = each player's expected value per
combo in fractions of the pot,

    EV_p(h) = v_p(h) / (norm * compat_p(h) * pot),

with v_p the subgame solver's root value of hand h under the average
strategy (FP64), norm its 1 / (compatible weight) and compat_p(h) the
opponent's root weight over hands avoiding h; the mask keeps the combos in
p's range with compat > 0. The depth-limited solver applies the inverse
at its leaves (norm * pot * EV * compatible reach, summed over the dealt
card).

The judge (`judge`): the depth-limited flop regrets are copied into the full
flop tree, the exact DCFR continues from there, and the solve is done when
the exact FP64 best response puts exploitability at or below the target;
the same exact solve from scratch is the comparison. `resolve` is the other
judge: the full tree's flop decision nodes pinned to the depth-limited
average strategy and only the later streets solved by exact DCFR under it.
"""
from __future__ import annotations

import math
import time
from typing import List, Optional

import numpy as np

from . import spots
from .config import COMBOS, FLOP, TURN, BettingState, DcfrParams, SolveConfig
from .net import ValueNet
from .ranges import HandClass
from .solver import BatchSolver, Solver, copy_root_street, depth_limited, pin_root_street

WORKLOAD = 4
CARD_MASK = np.array([[1 << a | 1 << b for (a, b) in COMBOS]], np.uint64)[0]


def class_weights(rng: spots.Rng, classes: int) -> List[float]:
    """Weights of the `classes` best of 169 classes under a seeded score
    (sigmoid(score + N(0, 0.5)), the flop workload's shape), others 0."""
    score = [rng.normal() for _ in range(169)]
    noise = [rng.normal() for _ in range(169)]
    keep = sorted(range(169), key=lambda c: -score[c])[:classes]
    w = [0.0] * 1326
    for g in keep:
        v = 1.0 / (1.0 + math.exp(-(score[g] + 0.5 * noise[g])))
        for c in HandClass.from_grid_index(g).combos():
            w[c] = v
    return w


def _unblocked(w, board):
    used = 0
    for c in board:
        used |= 1 << c
    return [x if not (used >> a & 1 or used >> b & 1) else 0.0 for x, (a, b) in zip(w, COMBOS)]


def eval_flop(seed: int, classes: int = 25) -> SolveConfig:
    """A flop of the value-network workload: the flop workload's tree (pot
    6bb, stacks 97bb, bets 33%/75%/all-in, raises 2.5x/all-in, 3 bets per
    street) with ranges of `classes` hand classes per player."""
    b = spots.board(spots.stream(WORKLOAD, seed, "board"), 3)
    r = [_unblocked(class_weights(spots.stream(WORKLOAD, seed, f"range{p}"), classes), b) for p in range(2)]
    return SolveConfig(b, r, BettingState.street_start(FLOP, 600, 9700, 9700), spots.flop_menu(),
                       max_memory_bytes=1 << 40, dcfr=DcfrParams())


def leaf_pots(cfg: SolveConfig):
    """(pot, stack behind) at the depth-limited flop tree's leaves."""
    s = depth_limited(cfg, None, device=-1)
    n = s.nodes()
    out = set()
    for i in np.nonzero(n["kind"] == 6)[0]:
        inv = n["invest"][i]
        stack = min(cfg.root.stack[0] - int(inv[0]), cfg.root.stack[1] - int(inv[1]))
        out.add((int(n["offset"][i]), int(stack)))
    return sorted(out)


def turn_subgame(seed: int, classes: int, pots, base=None) -> SolveConfig:
    """A training subgame: a seeded flop plus turn card, ranges of `classes`
    hand classes (or the flop ranges `base`, the ranges the depth-limited
    solve's leaves start from) each scaled by a random per-class factor in
    [0.05, 1] (the narrowing of flop betting), a pot and stack from the flop
    tree's leaves, the flop workload's menu on turn and river."""
    rng = spots.stream(WORKLOAD, 100000 + seed, "board")
    b = spots.board(rng, 4)
    r = []
    for p in range(2):
        rr = spots.stream(WORKLOAD, 100000 + seed, f"range{p}")
        w = list(base[p]) if base is not None else class_weights(rr, classes)
        f = [0.05 + 0.95 * rr.uniform() for _ in range(169)]
        for g in range(169):
            for c in HandClass.from_grid_index(g).combos():
                w[c] *= f[g]
        r.append(_unblocked(w, b))
    sh = spots.stream(WORKLOAD, 100000 + seed, "shape")
    pot, stack = pots[sh.below(len(pots))]
    return SolveConfig(b, r, BettingState.street_start(TURN, pot, stack, stack), spots.flop_menu(),
                       max_memory_bytes=1 << 40, dcfr=DcfrParams())


def _compat(w_opp):
    """compat(h) = sum of the opponent's weights over combos disjoint from h, per combo h."""
    w = np.asarray(w_opp, np.float64)
    card = np.zeros(52)
    for i, (a, b) in enumerate(COMBOS):
        card[a] += w[i]
        card[b] += w[i]
    a = np.array([x for x, _ in COMBOS])
    b = np.array([y for _, y in COMBOS])
    return w.sum() - card[a] - card[b] + w


def rows_of(cfg: SolveConfig, values, ncomb):
    """(input, target, mask) of one solved subgame; values: (2, 1326) root
    rows in combo order."""
    x = np.zeros(2706, np.float32)
    for c in cfg.board:
        x[c] = 1.0
    pot, stack = cfg.root.pot, min(cfg.root.stack)
    x[52] = pot / (pot + stack)
    x[53] = stack / (pot + stack)
    y = np.zeros(2652, np.float32)
    m = np.zeros(2652, np.float32)
    for p in range(2):
        w = np.asarray(cfg.ranges[p], np.float64)
        x[54 + p * 1326:54 + (p + 1) * 1326] = w / w.sum()
        comp = _compat(cfg.ranges[p ^ 1])
        ok = (w > 0) & (comp > 0)
        ev = np.zeros(1326)
        ev[ok] = values[p][ok] * ncomb / (comp[ok] * pot)
        y[p * 1326:(p + 1) * 1326] = ev
        m[p * 1326:(p + 1) * 1326] = ok
    return x, y, m


def generate(count: int, classes: int, pots, batch: int = 64, target_pct: float = 0.2, max_iterations: int = 1000,
             seed0: int = 0, log=None, base=None):
    """Solves `count` turn subgames in batches on the GPU; returns (X, Y, M)
    and the subgames' solve statistics."""
    X, Y, M, its = [], [], [], []
    t0 = time.perf_counter()
    for b0 in range(0, count, batch):
        cfgs = [turn_subgame(seed0 + i, classes, pots, base) for i in range(b0, min(count, b0 + batch))]
        bs = BatchSolver(cfgs)
        pr = bs.solve(max_iterations, target_pct, 10)
        rows = bs.root_rows(2)
        for k, c in enumerate(cfgs):
            x, y, m = rows_of(c, rows[k], bs.num_combinations(k))
            X.append(x)
            Y.append(y)
            M.append(m)
            its.append(pr[k].iteration)
        if log:
            log(f"[valuenet] solved {min(count, b0 + batch)}/{count} subgames ({time.perf_counter() - t0:.1f}s)")
    return np.array(X), np.array(Y), np.array(M), {"iterations": its, "seconds": time.perf_counter() - t0}


def _suit_perms():
    import itertools
    out = []
    for perm in itertools.permutations(range(4)):
        card = np.array([(c >> 2) * 4 + perm[c & 3] for c in range(52)])
        index = {}
        for i, (a, b) in enumerate(COMBOS):
            index[(max(a, b), min(a, b))] = i
        combo = np.zeros(1326, np.int64)
        for i, (a, b) in enumerate(COMBOS):
            x, y = card[a], card[b]
            combo[i] = index[(max(x, y), min(x, y))]
        out.append((card, combo))
    return out


PERMS = None


def permute(X, Y, M, k):
    """Rows with the suits relabelled by permutation k (an equivalent game)."""
    global PERMS
    if PERMS is None:
        PERMS = _suit_perms()
    card, combo = PERMS[k]
    X2 = np.zeros_like(X)
    X2[:, card] = X[:, :52]
    X2[:, 52:54] = X[:, 52:54]
    Y2, M2 = np.zeros_like(Y), np.zeros_like(M)
    for p in range(2):
        X2[:, 54 + p * 1326 + combo] = X[:, 54 + p * 1326:54 + (p + 1) * 1326]
        Y2[:, p * 1326 + combo] = Y[:, p * 1326:(p + 1) * 1326]
        M2[:, p * 1326 + combo] = M[:, p * 1326:(p + 1) * 1326]
    return X2, Y2, M2


def train(net: ValueNet, X, Y, M, steps: int, batch: int = 1024, lr: float = 1e-3, seed: int = 0, log=None):
    """Adam on minibatches (with replacement, suits relabelled at random) on
    the GPU; cosine learning-rate decay. Returns the loss per step."""
    rng = np.random.default_rng(seed)
    losses = []
    n = X.shape[0]
    for s in range(steps):
        idx = rng.integers(0, n, batch)
        x, y, m = permute(X[idx], Y[idx], M[idx], int(rng.integers(0, 24)))
        a = lr * 0.5 * (1.0 + math.cos(math.pi * s / max(steps, 1)))
        losses.append(net.train_step(x, y, m, max(a, lr * 0.02)))
        if log and (s % 200 == 0 or s == steps - 1):
            log(f"[valuenet] step {s}: loss {losses[-1]:.5f}")
    return losses


def heldout_loss(net: ValueNet, X, Y, M):
    """Masked Huber loss on rows the network was not trained on (rows padded
    to a multiple of 64 with empty masks)."""
    n = X.shape[0]
    pad = (-n) % 64
    if pad:
        X = np.concatenate([X, np.zeros((pad, X.shape[1]), X.dtype)])
        Y = np.concatenate([Y, np.zeros((pad, Y.shape[1]), Y.dtype)])
        M = np.concatenate([M, np.zeros((pad, M.shape[1]), M.dtype)])
    num, den = 0.0, 0.0
    for i in range(0, X.shape[0], 64):
        w = float(M[i:i + 64].sum())
        if w > 0:
            num += net.gradients(X[i:i + 64], Y[i:i + 64], M[i:i + 64]) * w
            den += w
    return num / max(den, 1e-30)


def projects_beyond(hist, target_pct: float, budget: int) -> bool:
    """True when the last two checks (iteration, % pot) fall too slowly to
    reach `target_pct` within `budget` iterations at their rate, or do not
    fall at all."""
    if len(hist) < 2:
        return False
    (t1, e1), (t2, e2) = hist[-2], hist[-1]
    rate = (e1 - e2) / (t2 - t1)
    return rate <= 0.0 or t2 + (e2 - target_pct) / rate > budget


def resolve(cfg: SolveConfig, dl: Solver, target_pct: float = 0.5, max_iterations: int = 300, check_every: int = 10,
            log=None):
    """The full tree with its flop decision nodes pinned (frozen) to the
    depth-limited solver's average strategy and the later streets solved by
    exact DCFR, until the exact FP64 best response of that profile is at or
    below the target. The re-solve gives up early when the last two checks'
    rate of fall projects the target beyond `max_iterations` (the pinned flop
    strategy bounds how far the later streets can bring the profile down).
    Returns (progress, seconds, solver, reached)."""
    full = Solver(cfg)
    t0 = time.perf_counter()
    pin_root_street(full, dl)
    hist = []
    t, e = 0, float("inf")
    nxt = check_every
    while t < max_iterations:
        full.iterate(1)
        t += 1
        if t >= nxt or t == max_iterations:
            e = full.exploitability()
            hist.append((t, 100.0 * e / cfg.root.pot))
            if log:
                log(f"[valuenet] resolve: iteration {t}, {hist[-1][1]:.4f}% pot")
            if hist[-1][1] <= target_pct:
                break
            if projects_beyond(hist, target_pct, max_iterations):
                if log:
                    log(f"[valuenet] resolve: the target projects beyond {max_iterations} iterations; giving up")
                break
            nxt = t + check_every
    reached = bool(hist) and hist[-1][1] <= target_pct
    return hist, time.perf_counter() - t0, full, reached


def judge(cfg: SolveConfig, warm: Optional[Solver] = None, target_pct: float = 0.5, max_iterations: int = 600,
          check_every: int = 10, regret_scale: float = 1.0, warm_t: int = 0):
    """Exact full-tree solve to the target, optionally warm-started from a
    depth-limited solver's flop strategy. Returns (progress, seconds)."""
    from .solver import SolveOptions
    full = Solver(cfg)
    t0 = time.perf_counter()
    if warm is not None:
        # the flop's current strategy (regrets) starts where the depth-limited
        # solve ended; the average strategy is rebuilt by the exact solve
        copy_root_street(full, warm, regret_scale, 0.0)
        full.iterations = warm_t
    pr = full.solve(SolveOptions(max_iterations, target_pct, check_every))
    return pr, time.perf_counter() - t0, full
