"""`ValueNet`: the counterfactual-value MLP of the depth-limited flop solve
over libpts (hand-written BF16 MFMA kernels on gfx950)."""
from __future__ import annotations

import ctypes

import numpy as np

from .native import check, lib

IN, OUT = 2706, 2652


class ValueNet:
    def __init__(self, batch_cap: int = 1024, seed: int = 1, device: int = 0):
        h = ctypes.c_void_p()
        check(lib().pts_net_create(device, batch_cap, seed, ctypes.byref(h)))
        self.h = h
        self.cap = batch_cap

    def __del__(self):
        if getattr(self, "h", None):
            lib().pts_net_destroy(self.h)
            self.h = None

    @staticmethod
    def num_params() -> int:
        return lib().pts_net_num_params()

    @staticmethod
    def forward_flops_per_row() -> float:
        return lib().pts_net_forward_flops_per_row()

    def params(self):
        p = np.zeros(self.num_params(), np.float32)
        check(lib().pts_net_get_params(self.h, p))
        return p

    def set_params(self, p):
        check(lib().pts_net_set_params(self.h, np.ascontiguousarray(p, np.float32)))

    def grads(self):
        g = np.zeros(self.num_params(), np.float32)
        check(lib().pts_net_get_grads(self.h, g))
        return g

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros((x.shape[0], OUT), np.float32)
        check(lib().pts_net_forward(self.h, x, x.shape[0], y))
        return y

    def gradients(self, x, target, mask):
        loss = ctypes.c_double()
        check(lib().pts_net_gradients(self.h, np.ascontiguousarray(x, np.float32), np.ascontiguousarray(target, np.float32),
                                      np.ascontiguousarray(mask, np.float32), x.shape[0], ctypes.byref(loss)))
        return loss.value

    def train_step(self, x, target, mask, lr):
        loss = ctypes.c_double()
        check(lib().pts_net_train_step(self.h, np.ascontiguousarray(x, np.float32), np.ascontiguousarray(target, np.float32),
                                       np.ascontiguousarray(mask, np.float32), x.shape[0], lr, ctypes.byref(loss)))
        return loss.value

    def profile(self, rows: int):
        f, t = ctypes.c_double(), ctypes.c_double()
        check(lib().pts_net_profile(self.h, rows, ctypes.byref(f), ctypes.byref(t)))
        return f.value, t.value
