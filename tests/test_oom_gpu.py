"""Out-of-memory errors carry the bytes needed and available
(GpuError::OutOfMemory { needed, available }, crates/solver/src/gpu/
mod.rs:86-136), counted with the storage's own cell size: 16-bit storage
needs 2 + 2 bytes per cell plus 12 bytes per decision node, so a tree too big
for f32 storage but not for 16-bit storage is created."""
import ctypes

import numpy as np
import pytest

from pts import spots
from pts.native import PtsError, lib
from pts.solver import Engine

from test_host_cpu import oracle_tree

pytestmark = pytest.mark.gpu

HBM = 288 << 30  # 288 GiB of HBM3E


def tree_with_storage(cells):
    o, t = oracle_tree(spots.river_spot(0))
    t.desc.storage_len = cells  # cells past the tree's own are valid, unused storage
    return o, t


def slots(o):
    nodes = o.nodes()
    return np.where(nodes["kind"] <= 1, nodes["slot"], 0).astype(np.uint32), o.stats().player_nodes


def test_f32_storage_out_of_memory():
    cells = 50 * 10 ** 9  # 400 GB of f32 regrets and sums
    _, t = tree_with_storage(cells)
    with pytest.raises(PtsError) as ex:
        Engine(t, None, None)
    e = ex.value
    assert e.status == 6 and e.hip_error == 0
    assert 8 * cells <= e.needed < 8 * cells + 10 ** 7
    assert 0 < e.available < HBM


def test_16bit_storage_out_of_memory():
    cells = 100 * 10 ** 9  # 400 GB of i16 regrets and u16 sums
    o, t = tree_with_storage(cells)
    slot, k = slots(o)
    with pytest.raises(PtsError) as ex:
        Engine(t, None, None, q16=(slot, None, None, None, None))
    e = ex.value
    assert e.status == 6 and e.hip_error == 0
    assert 4 * cells + 12 * k <= e.needed < 4 * cells + 10 ** 7
    assert 0 < e.available < HBM


def test_16bit_storage_that_f32_could_not_hold():
    cells = 45 * 10 ** 9  # f32: 360 GB (more than the HBM); 16-bit: 180 GB
    o, t = tree_with_storage(cells)
    slot, k = slots(o)
    e = Engine(t, None, None, q16=(slot, None, None, None, None))
    assert e.device_bytes() >= 4 * cells
    assert np.isfinite(e.exploitability())
    del e
