"""The chance-sampled mode (crates/solver/src/sampling.rs, tree.rs:128-136,
314-334). Its cards come from rand 0.9.5's StdRng (rand_chacha 0.9.0's
ChaCha12Rng, seeded by rand_core 0.9.5's seed_from_u64), a crate absent from
the reference tree; the oracle restates its published algorithm, and the
product's is checked against the oracle's. Pinned here by published known
answers: the ChaCha block function (RFC 8439 section 2.3.2), the stream's
layout (key, 64-bit block counter from 0, block and word order, next_u64 low
word first) by the keystream of RFC 8439 appendix A.1 test vectors 1 and 2
run through the oracle's StdRng with 20 rounds, and the seed expansion's
PCG32 step and XSH RR output by the PCG reference generator's first outputs
(seed 42, stream 54). The 12-round count and the increment
11634580027462260723 are the crates' constants; the sample is pinned by the
reference's own test and by the stratification it promises."""
import numpy as np
import pytest

import pyoracle as O
from pts import PostflopMenu, make_config, parse_cards, DcfrParams
from pts.solver import Solver


def test_chacha20_block_rfc8439():
    # RFC 8439 section 2.3.2: key 00..1f, block count 1, nonce 00:00:00:09:00:00:00:4a:00:00:00:00
    state = [0x61707865, 0x3320646e, 0x79622d32, 0x6b206574,
             0x03020100, 0x07060504, 0x0b0a0908, 0x0f0e0d0c,
             0x13121110, 0x17161514, 0x1b1a1918, 0x1f1e1d1c,
             0x00000001, 0x09000000, 0x4a000000, 0x00000000]
    want = [0xe4e7f110, 0x15593bd1, 0x1fdd0f50, 0xc47120a3,
            0xc7f4d1c7, 0x0368c033, 0x9aaa2204, 0x4e6cd4c3,
            0x466482d2, 0x09aa9f07, 0x05d7c214, 0xa2028bd9,
            0xd19c12b5, 0xb94e16de, 0xe883d0cb, 0x4e3c50a2]
    assert O.chacha_block(state, 20).tolist() == want


def test_stream_layout_by_the_rfc8439_keystream():
    # RFC 8439 appendix A.1, test vectors 1 and 2: key and nonce all zero,
    # block counters 0 and 1; StdRng's stream with 20 rounds and a zero key
    # must give these 128 bytes in order, and next_u64 the words in pairs,
    # low word first
    tv1 = ("76b8e0ada0f13d90405d6ae55386bd28bdd219b8a08ded1aa836efcc8b770dc7"
           "da41597c5157488d7724e03fb8d84a376a43b8f41518a11cc387b669b2ee6586")
    tv2 = ("9f07e7be5551387a98ba977c732d080dcb0f29a048e3656912c6533e32ee7aed"
           "29b721769ce64e43d57133b074d839d531ed1f28510afb45ace10a1f4b794d6f")
    w, w64 = O.stream_words(np.zeros(8, np.uint32), 20, 32)
    assert w.astype("<u4").tobytes() == bytes.fromhex(tv1 + tv2)
    assert w64.tolist() == [int(w[2 * i]) | int(w[2 * i + 1]) << 32 for i in range(16)]


def test_seed_expansion_by_the_pcg32_reference_outputs():
    # the PCG reference generator (M. O'Neill, pcg32_srandom_r(42, 54)):
    # increment 54 * 2 + 1, its state before the first output (0 -> step ->
    # + 42) is 151, and its first outputs are these six words; seed_from_u64
    # draws its eight key words with the same step and output
    inc = 54 * 2 + 1
    assert O.pcg32_words(151, inc, 6).tolist() == [0xa15c02b7, 0x7b47f409, 0xba1d3330,
                                                   0x83d2f293, 0xbfa4784b, 0xcbed606e]


def test_stream_is_deterministic_and_uniform():
    a, b = O.stdrng_u64(7, 5000), O.stdrng_u64(7, 5000)
    assert np.array_equal(a, b) and not np.array_equal(a, O.stdrng_u64(8, 5000))
    f = (a >> np.uint64(11)).astype(np.float64) / 2.0 ** 53
    assert abs(f.mean() - 0.5) < 0.02 and f.min() >= 0.0 and f.max() < 1.0
    bits = np.unpackbits(a.view(np.uint8))
    assert abs(bits.mean() - 0.5) < 0.005


def available(board):
    used = set(parse_cards(board))
    return [c for c in range(52) if c not in used]


def test_stratified_by_suit():
    # sampling.rs:58-76
    av = available("AsKs7d")
    picks = O.sample_cards(av, 12, 1)
    assert len(picks) == 12 and all(p in av for p in picks)
    assert 2 <= sum(1 for p in picks if p & 3 == 3) <= 3


@pytest.mark.parametrize("board", ["AsKs7d", "Kh7d2c", "2c3c4c", "Kh7d2cTs", "AhAdAc"])
def test_sample_properties(board):
    av = available(board)
    per_suit = [[c for c in av if c & 3 == s] for s in range(4)]
    for k in (1, 3, 8, 13, len(av) - 1):
        for seed in range(6):
            picks = O.sample_cards(av, k, seed)
            assert len(picks) == k and picks == sorted(set(picks)) and set(picks) <= set(av)
            for s in range(4):
                q = sum(1 for p in picks if p & 3 == s)
                share = k * len(per_suit[s]) / len(av)
                assert np.floor(share) <= q <= np.ceil(share)
                # even spread within the suit: gaps between picked positions
                # are the step rounded down or up
                if q >= 2:
                    pos = [per_suit[s].index(p) for p in picks if p & 3 == s]
                    step = len(per_suit[s]) / q
                    assert all(np.floor(step) <= g <= np.ceil(step) for g in np.diff(pos))
    assert O.sample_cards(av, len(av), 3) == av  # nothing to sample


def sampled_flop(seed=7, iso=True):
    # gpu/testutil.rs:120-131
    c = make_config("Ks7s2d", "AK, KQ, 77, 22, A5s, QJs, T9s", "AA, KK, AQ, QQ, JJ, 98s, 65s, A2s", 500, 3000,
                    PostflopMenu.one(["50%"], ["3x"]), isomorphism=iso, dcfr=DcfrParams.reference_default())
    c.sampled = (3, 2, seed)
    return c


def test_sampled_tree_shape_and_seeds():
    # without isomorphism the seed changes which cards are dealt, not the
    # tree's shape: 3 turn cards and 2 river cards per branch, all-in
    # runouts dealt in full
    a, b = O.OracleSolver(sampled_flop(7, iso=False)), O.OracleSolver(sampled_flop(8, iso=False))
    na, nb = a.nodes(), b.nodes()
    assert np.array_equal(na["kind"], nb["kind"]) and np.array_equal(na["num_children"], nb["num_children"])
    assert not np.array_equal(na["edge_card"], nb["edge_card"])
    chance = np.nonzero(na["kind"] == 2)[0]
    street = na["street"][chance]
    kids = na["num_children"][chance]
    # each chance node deals the sample unless the runout is all-in (every card left)
    assert set(kids[street == 1].tolist()) <= {3, 49} and set(kids[street == 2].tolist()) <= {2, 48}
    assert (kids[street == 1] == 3).any() and (kids[street == 2] == 2).any()
    # the estimate counts the same tree
    st = O.OracleSolver.estimate(sampled_flop(7, iso=False))
    assert st.player_nodes == a.stats().player_nodes and st.storage_cells == a.stats().storage_cells


def test_product_builds_the_oracle_tree():
    for seed in (7, 11):
        cfg = sampled_flop(seed)
        o = O.OracleSolver(cfg)
        s = Solver(cfg, device=-1)
        a, b = o.nodes(), s.nodes()
        for k in ("kind", "num_children", "first_child", "offset"):
            assert np.array_equal(a[k].astype(np.int64), b[k].astype(np.int64)), k
        assert np.array_equal(np.where(a["edge_card"] >= 0, a["edge_card"], -1), b["card"].astype(np.int64))
        assert np.array_equal(a["chance_factor"], b["chance_factor"])
        sa, sb = o.stats(), s.stats()
        assert (sa.player_nodes, sa.chance_nodes, sa.terminal_nodes, sa.storage_cells) == \
            (sb.player_nodes, sb.chance_nodes, sb.terminal_nodes, sb.storage_cells)


def test_sampled_mode_rejects_empty_samples():
    cfg = sampled_flop()
    cfg.sampled = (0, 2, 1)
    with pytest.raises(Exception):
        Solver(cfg, device=-1)
