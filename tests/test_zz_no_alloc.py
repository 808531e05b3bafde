"""The allocation-free network forward (pts_net_forward_noalloc) against
pts_net_forward on the same rows: device buffers in and out (allocated here
by the HIP runtime, as a caller would), two calls in a row, the second with
fewer rows, and the size limit."""
import ctypes

import numpy as np
import pytest

from pts.native import PtsError, check, lib
from pts.net import IN, OUT, ValueNet

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def hip():
    h = ctypes.CDLL("libamdhip64.so.7")  # the HIP runtime libpts.so links
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipSetDevice.argtypes = [ctypes.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


def test_noalloc_forward_matches_forward():
    H = hip()
    net = ValueNet(1024, seed=5)
    assert H.hipSetDevice(0) == 0
    rng = np.random.default_rng(7)
    xd, yd = ctypes.c_void_p(), ctypes.c_void_p()
    assert H.hipMalloc(ctypes.byref(xd), 256 * IN * 4) == 0
    assert H.hipMalloc(ctypes.byref(yd), 256 * OUT * 4) == 0
    try:
        for rows in (256, 64):
            x = (rng.standard_normal((rows, IN)) * 0.1).astype(np.float32)
            want = net.forward(x)
            nan = np.full((rows, OUT), np.nan, np.float32)
            assert H.hipMemcpy(xd, x.ctypes.data, x.nbytes, H2D) == 0
            assert H.hipMemcpy(yd, nan.ctypes.data, nan.nbytes, H2D) == 0
            check(lib().pts_net_forward_noalloc(net.h, xd, rows, yd, None))
            got = np.zeros((rows, OUT), np.float32)
            assert H.hipDeviceSynchronize() == 0  # the forward ran on the network's own stream
            assert H.hipMemcpy(got.ctypes.data, yd, got.nbytes, D2H) == 0
            assert np.array_equal(got, want), f"{rows} rows: {int(np.sum(got != want))} outputs differ"
    finally:
        H.hipFree(xd)
        H.hipFree(yd)


def test_noalloc_forward_rejects_sizes_beyond_its_scratch():
    net = ValueNet(64, seed=1)
    for rows in (0, 100, 1088):
        with pytest.raises(PtsError, match="at most 1024"):
            check(lib().pts_net_forward_noalloc(net.h, ctypes.c_void_p(16), rows, ctypes.c_void_p(16), None))
