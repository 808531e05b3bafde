"""Strategy comparison between two solves of the same tree.

`node_diff` is the reference's per-node comparison (crates/solver/examples/
compare.rs:55-94): per-hand strategy distance, EV difference and the EV a
hand loses by following the other solve's best action, weighted by the first
solve's reach at the node.

`infoset_l1` compares every information set (decision node x hand) of two
storages: the average strategy of each (the normalized strategy sums,
crates/solver/src/cfr.rs:124-154) and the L1 distance sum_a |p_a - q_a|
between them.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class NodeDiff:
    range_freq_pp: float    # largest range-frequency difference of an action, percentage points
    strat_l1_pct: float     # reach-weighted mean of sum_a |s_a - f_a| / 2, percent
    ev_abs_pct: float       # reach-weighted mean |EV difference| per action, % of the pot
    loss_pct: float         # reach-weighted mean EV lost following the second solve's best action, % pot
    disagree_pct: float     # reach share of hands whose best action differs and loses more than 0.5% pot
    worst_loss_pct: float   # the largest such loss of one hand, % pot


def node_diff(exact, fast) -> NodeDiff:
    """compare.rs:55-94 on two NodeInfo of the same decision node."""
    assert exact.actions == fast.actions, (exact.actions, fast.actions)
    na = len(exact.actions)
    pot = float(exact.pot)
    rf_e, rf_f = exact.range_strategy(), fast.range_strategy()
    range_freq = float(np.max(np.abs(rf_e - rf_f))) if na else 0.0
    wsum = l1 = evd = loss = dis = worst = 0.0
    for h in range(len(exact.combos)):
        w = float(exact.reach[h])
        if w <= 0.0:
            continue
        hf = fast.hand_index(int(exact.combos[h]))
        se, sf = exact.hand_strategy(h).astype(np.float64), fast.hand_strategy(hf).astype(np.float64)
        ee, ef = exact.hand_ev(h).astype(np.float64), fast.hand_ev(hf).astype(np.float64)
        best_e, best_f = int(np.argmax(ee)), int(np.argmax(ef))
        lh = (ee[best_e] - ee[best_f]) / pot * 100.0
        wsum += w
        l1 += w * float(np.sum(np.abs(se - sf))) / 2.0
        evd += w * float(np.sum(np.abs(ee - ef))) / na / pot * 100.0
        loss += w * lh
        dis += w * (1.0 if best_e != best_f and lh > 0.5 else 0.0)
        worst = max(worst, lh)
    wsum = max(wsum, 1e-300)
    return NodeDiff(range_freq * 100.0, l1 / wsum * 100.0, evd / wsum, loss / wsum, dis / wsum * 100.0, worst)


def average_strategy(sums: np.ndarray, na: int, n: int) -> np.ndarray:
    """cfr.rs:124-154: one node's strategy sums (na x n, action-major)
    normalized per hand in f32; hands with no positive total play uniform."""
    s = np.asarray(sums, np.float32).reshape(na, n)
    tot = s[0].copy()
    for a in range(1, na):
        tot = tot + s[a]
    with np.errstate(divide="ignore"):
        inv = np.float32(1.0) / tot
    inv = np.where((tot > 0) & np.isfinite(inv), inv, np.float32(0)).astype(np.float32)
    out = s * inv
    out[:, inv == 0] = np.float32(1.0) / np.float32(na)
    return out


@dataclass
class InfosetL1:
    infosets: int           # decision node x hand pairs compared
    max_l1: float           # the largest sum_a |p_a - q_a|
    mean_l1: float
    over: int               # infosets with an L1 above `tol`
    tol: float
    worst_node: int         # the decision node holding the largest L1
    worst_hand: int


def infoset_l1(nodes, hands, sums_a, sums_b, tol: float = 1e-3) -> InfosetL1:
    """Per-information-set L1 distance between the average strategies of two
    storages of one tree. `nodes`: the arena (kind, num_children, offset);
    `hands`: hands per player."""
    kind, nch, off = nodes["kind"], nodes["num_children"], nodes["offset"]
    count, mx, total, over, wn, wh = 0, 0.0, 0.0, 0, -1, -1
    for i in np.nonzero(kind <= 1)[0]:
        na, n, o = int(nch[i]), int(hands[int(kind[i])]), int(off[i])
        pa = average_strategy(sums_a[o:o + na * n], na, n)
        pb = average_strategy(sums_b[o:o + na * n], na, n)
        l1 = np.sum(np.abs(pa.astype(np.float64) - pb.astype(np.float64)), axis=0)
        count += n
        total += float(l1.sum())
        over += int(np.sum(l1 > tol))
        j = int(np.argmax(l1))
        if l1[j] > mx:
            mx, wn, wh = float(l1[j]), int(i), j
    return InfosetL1(count, mx, total / max(count, 1), over, tol, wn, wh)
