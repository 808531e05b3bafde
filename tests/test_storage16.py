"""Pins of the oracle's 16-bit storage (crates/solver/src/storage.rs:87-330):
i16 regrets and u16 strategy sums, one f32 scale per player node. The codec
is checked against closed forms and the reference's round-trip test; the
solver against the f32 solver it approximates (tests.rs:273-302) and, for
the first update from zero storage, against the f32 update encoded."""
import numpy as np
import pytest

import pyoracle as O
from pts import DcfrParams, PostflopMenu, make_config

REF_DCFR = DcfrParams.reference_default()
VALS = np.array([1.0, -2.0, 0.5, 0.0, 3.0, -0.25, 7.0, -7.0], np.float32)


def test_codec_roundtrip():
    # storage.rs:333-357
    q, scale = O.encode_block(VALS, sums=False)
    assert scale == 7.0
    dec = q.astype(np.float32) * np.float32(scale / np.float32(32767.0))
    assert np.all(np.abs(dec - VALS) <= 7.0 / 32767.0 + 1e-6)
    q, scale = O.encode_block(np.abs(VALS), sums=True)
    assert scale == 7.0
    dec = q.astype(np.float32) * np.float32(scale / np.float32(65535.0))
    assert np.all(np.abs(dec - np.abs(VALS)) <= 7.0 / 65535.0 + 1e-6)


def test_codec_closed_forms():
    # 32767 / 7 = 4681 exactly: v * 4681 rounded half away from zero
    q, _ = O.encode_block(VALS, sums=False)
    assert q.tolist() == [4681, -9362, 2341, 0, 14043, -1170, 32767, -32767]
    # sums: v * (65535 / 7) + 0.5, truncated; the largest maps to 65535
    q, _ = O.encode_block(np.abs(VALS), sums=True)
    inv = np.float32(65535.0) / np.float32(7.0)
    want = np.floor(np.abs(VALS) * inv + np.float32(0.5)).astype(np.uint16)
    assert q.tolist() == want.tolist()
    assert q[6] == q[7] == 65535 and q[3] == 0
    # ties: 0.5 steps round away from zero in both directions
    q, _ = O.encode_block(np.array([32767.0, 0.5, -0.5, 1.5, -1.5, -2.5], np.float32), sums=False)
    assert q.tolist() == [32767, 1, -1, 2, -2, -3]
    # an all-zero block: scale 0, zeros
    q, s = O.encode_block(np.zeros(5, np.float32), sums=False)
    assert s == 0.0 and not q.any()
    # sums below zero (rounding noise) store as 0
    q, s = O.encode_block(np.array([-1e-7, 2.0], np.float32), sums=True)
    assert q.tolist() == [0, 65535] and s == 2.0


def turn_config(storage_mode):
    # tests.rs:273-302 (and gpu/testutil.rs:80-89)
    c = make_config("Qh9d4c2s", "AQ, KQ, QJs, JTs, 99, 44, T9s, 87s, A5s, KJs",
                    "AA, KK, AQ, QT, 99, JJ, TT, 98s, KJs, 65s, AJs", 600, 4000, PostflopMenu(), dcfr=REF_DCFR)
    c.storage_mode = storage_mode
    return c


def test_auto_compresses_above_one_gigabyte():
    # lib.rs:346-352
    a, b = O.OracleSolver(turn_config(1)), O.OracleSolver(turn_config(2))
    assert a.compressed and not b.compressed  # this turn tree is far below 1 GB


def test_first_update_is_the_f32_update_encoded():
    # From zero storage the decoded blocks are exact zeros, so player 0's
    # regret update and player 1's strategy-sum update (both in player 0's
    # traversal) are the f32 solver's blocks, stored through the codec.
    a, b = O.OracleSolver(turn_config(1)), O.OracleSolver(turn_config(0))
    a.traverse_root(0, 0, 1)
    b.traverse_root(0, 0, 1)
    rq, sq, rs, ss = a.storage_q()
    rf, sf = b.storage()
    nodes = a.nodes()
    n = [a.n[0], a.n[1]]
    checked = [0, 0]
    for i in np.nonzero(nodes["kind"] <= 1)[0]:
        k = int(nodes["kind"][i])
        off, na, slot = int(nodes["offset"][i]), int(nodes["num_children"][i]), int(nodes["slot"][i])
        blk = slice(off, off + na * n[k])
        if k == 0:
            want, scale = O.encode_block(rf[blk], sums=False)
            assert np.array_equal(rq[blk], want) and rs[slot] == scale == np.abs(rf[blk]).max()
        else:
            want, scale = O.encode_block(sf[blk], sums=True)
            assert np.array_equal(sq[blk], want) and ss[slot] == scale == np.abs(sf[blk]).max()
        checked[k] += 1
    assert checked[0] > 10 and checked[1] > 10


def test_compressed_storage_is_close():
    # tests.rs:273-302: half the memory, both converge, root strategies agree
    a, b = O.OracleSolver(turn_config(1)), O.OracleSolver(turn_config(0))
    assert a.compressed and not b.compressed
    cells, slots = a.storage_len, a.stats().player_nodes
    assert cells * 4 + slots * 8 < 0.55 * cells * 8  # Storage::bytes (storage.rs:197-205)
    pa, _ = a.solve(300, 0.2, 20)
    pb, _ = b.solve(300, 0.2, 20)
    assert pa.exploitability_pct < 0.5 and pb.exploitability_pct < 0.5
    ra, rb = a.query([]).range_strategy(), b.query([]).range_strategy()
    assert np.max(np.abs(ra - rb)) < 0.03
