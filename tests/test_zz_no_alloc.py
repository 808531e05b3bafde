"""The allocation-free network forward (pts_net_forward_noalloc) against
pts_net_forward on the same rows: device buffers in and out (torch is only
the allocator of the caller's rows here), two calls in a row, the second
with fewer rows, and the size limit."""
import ctypes

import numpy as np
import pytest

from pts.native import PtsError, check, lib
from pts.net import IN, OUT, ValueNet

pytestmark = pytest.mark.gpu


def test_noalloc_forward_matches_forward():
    import torch
    net = ValueNet(1024, seed=5)
    rng = np.random.default_rng(7)
    for rows in (256, 64):
        x = rng.standard_normal((rows, IN)).astype(np.float32) * 0.1
        want = net.forward(x)
        xd = torch.from_numpy(x).to("cuda:0")
        yd = torch.full((rows, OUT), float("nan"), dtype=torch.float32, device="cuda:0")
        torch.cuda.synchronize()
        check(lib().pts_net_forward_noalloc(net.h, xd.data_ptr(), rows, yd.data_ptr(), None))
        torch.cuda.synchronize()
        got = yd.cpu().numpy()
        assert np.array_equal(got, want), f"{rows} rows: {int(np.sum(got != want))} outputs differ"


def test_noalloc_forward_rejects_sizes_beyond_its_scratch():
    net = ValueNet(64, seed=1)
    for rows in (0, 100, 1088):
        with pytest.raises(PtsError, match="at most 1024"):
            check(lib().pts_net_forward_noalloc(net.h, ctypes.c_void_p(16), rows, ctypes.c_void_p(16), None))
