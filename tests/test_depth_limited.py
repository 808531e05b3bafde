"""The depth-limited solve: trees cut where the root street's betting ends,
leaves valued per dealt card by the value network inside the GPU iteration.

The leaf arithmetic is pinned on a network whose output does not depend on
its input (zero weights, random output biases b): a leaf then pays
pot * b[p, combo] against every compatible opponent hand, because the
cut chance node's factor times the cards that avoid both hands is 1. An
independent numpy recursion over the same tree gives the root values under
uniform strategies."""
import numpy as np
import pytest

from pts import PostflopMenu, make_config, spots
from pts.config import COMBOS
from pts.net import ValueNet
from pts.solver import Solver, depth_limited


def turn_cfg():
    return make_config("2c7d9hTs", "AA, KK, QQ, AKs, AQs, KQs, JTs, 98s, 76s", "TT, 99, 88, AJs, KJs, QJs, T9s, 87s, 65s",
                       1000, 5000, PostflopMenu.one(["50%"], [], max_raises=1))


def test_depth_limited_tree_has_leaves_and_no_later_street():
    cfg = turn_cfg()
    dl = Solver(cfg, device=-1)
    full_nodes = dl.nodes()
    from pts.native import lib
    import ctypes
    c = cfg.to_c()
    h = ctypes.c_void_p()
    # host-only build of the depth-limited tree is exercised on the GPU tests;
    # here: the full tree has river betting the cut tree will not
    assert (full_nodes["street"] == 3).any()


def hands_of(cfg, p):
    b = 0
    for x in cfg.board:
        b |= 1 << x
    idx = [c for c in range(1326) if cfg.ranges[p][c] > 0 and not (b >> COMBOS[c][0] & 1 or b >> COMBOS[c][1] & 1)]
    return idx, np.array([cfg.ranges[p][c] for c in idx], np.float64)


def compat_matrix(h0, h1):
    m = np.zeros((len(h0), len(h1)))
    for i, a in enumerate(h0):
        sa = set(COMBOS[a])
        for j, b in enumerate(h1):
            m[i, j] = 0.0 if sa & set(COMBOS[b]) else 1.0
    return m


def uniform_values(nodes, cfg, p, bias):
    """Per-hand values of player p in the depth-limited tree, both players
    uniform, leaves paying pot * bias[p, combo]."""
    hs = [hands_of(cfg, q) for q in range(2)]
    M = compat_matrix(hs[p][0], hs[p ^ 1][0])
    norm = 1.0 / float(np.sum(compat_matrix(hs[0][0], hs[1][0]) * np.outer(hs[0][1], hs[1][1])))
    hp = cfg.root.pot / 2.0
    ev = np.array([bias[p * 1326 + c] for c in hs[p][0]], np.float64)
    kind, fc, nc, inv, off = nodes["kind"], nodes["first_child"], nodes["num_children"], nodes["invest"], nodes["offset"]
    fac = nodes["chance_factor"]
    avoid = 52 - len(cfg.board) - 4  # cards dealt at a leaf that avoid both hands

    def rec(i, r):
        k = int(kind[i])
        if k in (0, 1):
            kids = [int(fc[i]) + a for a in range(int(nc[i]))]
            if k == p:
                return sum(rec(c, r) for c in kids) / len(kids)
            return sum(rec(c, r / len(kids)) for c in kids)
        compat = M @ r
        if k in (3, 4):
            f = k - 3
            pay = -(hp + inv[i][p]) if f == p else hp + inv[i][f]
            return pay * norm * compat
        assert k == 6, k
        # sum over the dealt cards of the compatible reach times the stored
        # (f32) chance factor, as the leaf kernel forms it
        return float(off[i]) * ev * norm * compat * (avoid * float(fac[i]))

    return rec(0, hs[p ^ 1][1]), hs[p][1]


@pytest.mark.gpu
def test_leaf_values_with_an_input_independent_network():
    cfg = turn_cfg()
    net = ValueNet(64 * 4, seed=2)
    p = np.zeros(ValueNet.num_params(), np.float32)
    rng = np.random.default_rng(3)
    bias = (rng.standard_normal(2652) * 0.3).astype(np.float32)
    p[-2652:] = bias
    net.set_params(p)
    dl = depth_limited(cfg, net)
    assert dl.num_groups() >= 1
    nodes = dl.nodes()
    assert (nodes["kind"] == 6).any() and not (nodes["street"] == 3).any()
    got = dl.root_values(2)  # average strategy: uniform from zero sums
    for q in range(2):
        v, w = uniform_values(nodes, cfg, q, bias.astype(np.float64))
        want = float(np.dot(v, w))
        assert abs(got[q] - want) <= 1e-9 * max(1.0, abs(want)), (q, got[q], want)


@pytest.mark.gpu
def test_depth_limited_solve_runs_and_copies_into_the_full_tree():
    cfg = turn_cfg()
    net = ValueNet(64 * 4, seed=4)
    dl = depth_limited(cfg, net)
    dl.iterate(20)
    full = Solver(cfg)
    n = full.storage_len
    from pts.solver import copy_root_street
    copied = copy_root_street(full, dl)
    assert copied == int(np.sum((dl.nodes()["kind"] <= 1)))
    r, s = full.storage()
    assert np.count_nonzero(s) > 0 and len(r) == n


@pytest.mark.gpu
def test_pinned_root_street_plays_the_depth_limited_average():
    cfg = turn_cfg()
    dl = depth_limited(cfg, ValueNet(64 * 4, seed=4))
    dl.iterate(20)
    full = Solver(cfg)
    from pts.solver import pin_root_street
    assert pin_root_street(full, dl) == int(np.sum(dl.nodes()["kind"] <= 1))
    nd, nf = dl.nodes(), full.nodes()
    rd, sd = dl.storage()
    rf, sf = full.storage()
    # the root: sums copied, regrets the average strategy (uniform where a
    # hand's sums are zero), so regret matching in the full tree plays it
    k, na = int(nd["kind"][0]), int(nd["num_children"][0])
    nh = dl.n[k]
    od, of = int(nd["offset"][0]), int(nf["offset"][0])
    sdb = sd[od:od + na * nh].reshape(na, nh)
    assert np.array_equal(sf[of:of + na * nh].reshape(na, nh), sdb)
    tot = sdb.astype(np.float64).sum(axis=0)
    want = np.where(tot > 0, sdb / np.where(tot > 0, tot, 1.0), 1.0 / na).astype(np.float32)
    assert np.allclose(rf[of:of + na * nh].reshape(na, nh), want, rtol=1e-6, atol=0)
    assert np.any(tot > 0)
    # the pinned nodes are frozen: iterations of the full tree (both
    # traversals) leave every pinned node's regrets as written, so the root
    # keeps playing the pinned strategy and its average stays that strategy
    pinned = [i for i in np.nonzero(nf["kind"] <= 1)[0] if nf["street"][i] == nf["street"][0]]
    before = {i: rf[int(nf["offset"][i]):int(nf["offset"][i]) + int(nf["num_children"][i]) * full.n[int(nf["kind"][i])]].copy()
              for i in pinned}
    full.iterate(6)
    r2, s2 = full.storage()
    for i, blk in before.items():
        o = int(nf["offset"][i])
        assert np.array_equal(r2[o:o + len(blk)], blk), i
    s2b = s2[of:of + na * nh].reshape(na, nh).astype(np.float64)
    tot2 = s2b.sum(axis=0)
    live = tot2 > 0
    assert np.allclose((s2b / np.where(live, tot2, 1.0))[:, live], want[:, live], rtol=0, atol=1e-5)
    # the later streets are not frozen: their regrets moved
    later = [i for i in np.nonzero(nf["kind"] <= 1)[0] if nf["street"][i] != nf["street"][0]]
    assert any(not np.array_equal(r2[int(nf["offset"][i]):int(nf["offset"][i]) + 4], rf[int(nf["offset"][i]):int(nf["offset"][i]) + 4])
               for i in later)
