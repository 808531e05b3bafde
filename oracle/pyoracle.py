"""ctypes wrapper of the CPU oracle (oracle/build/liboracle.so).

Test infrastructure only: tests/, the smoke entry point and the benchmark's
cpu_baseline leg load it as the checker; the product never does.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liboracle.so")
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "poker-trainer_amd"))

from pts.config import ConfigC, MenuC, SolveConfig  # noqa: E402

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


class StatsC(ctypes.Structure):
    _fields_ = [("player_nodes", ctypes.c_uint64), ("chance_nodes", ctypes.c_uint64), ("terminal_nodes", ctypes.c_uint64),
                ("storage_cells", ctypes.c_uint64), ("memory_bytes", ctypes.c_uint64)]


class ProgressC(ctypes.Structure):
    _fields_ = [("iteration", ctypes.c_uint32), ("max_iterations", ctypes.c_uint32), ("exploitability", ctypes.c_double),
                ("exploitability_pct", ctypes.c_double), ("finished", ctypes.c_int32), ("_pad", ctypes.c_int32)]


class StepC(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("act_kind", ctypes.c_int32), ("to", ctypes.c_int64), ("card", ctypes.c_int32),
                ("_pad", ctypes.c_int32)]


P = ctypes.c_void_p
fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def _declare(L):
    L.orc_evaluate.restype = ctypes.c_uint32
    L.orc_evaluate.argtypes = [ctypes.POINTER(ctypes.c_int32), ctypes.c_int32]
    L.orc_combo_index.restype = ctypes.c_int32
    L.orc_evaluate_batch.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_uint64, ctypes.c_void_p]
    L.orc_enumerate_hands.restype = ctypes.c_uint64
    L.orc_enumerate_hands.argtypes = [ctypes.c_int32, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_void_p]
    L.orc_parse_range.argtypes = [ctypes.c_char_p, fp]
    L.orc_count_categories.argtypes = [ctypes.c_int32, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)]
    L.orc_postflop_actions.restype = ctypes.c_int32
    L.orc_postflop_actions.argtypes = [ctypes.c_int32, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                       ctypes.c_int64, ctypes.c_uint32, ctypes.POINTER(MenuC),
                                       ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int64), ctypes.c_int32]
    L.orc_discount.argtypes = [ctypes.c_uint32, ctypes.c_double, ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                               ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(ctypes.c_float)]
    L.orc_next_check_gap.restype = ctypes.c_uint32
    L.orc_next_check_gap.argtypes = [ctypes.c_int32, ctypes.c_uint32, ctypes.c_double, ctypes.c_uint32, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_uint32]
    L.orc_solver_new.restype = P
    L.orc_solver_new.argtypes = [ctypes.POINTER(ConfigC), ctypes.c_char_p, ctypes.c_int32]
    L.orc_estimate.argtypes = [ctypes.POINTER(ConfigC), ctypes.POINTER(StatsC), ctypes.c_char_p, ctypes.c_int32]
    L.orc_solver_free.argtypes = [P]
    L.orc_solver_set_parallel.argtypes = [P, ctypes.c_int32, ctypes.c_int32]
    L.orc_solver_set_batch_river.argtypes = [P, ctypes.c_int32]
    L.orc_solver_stats.argtypes = [P, ctypes.POINTER(StatsC)]
    L.orc_solver_storage_len.restype = ctypes.c_uint64
    L.orc_solver_storage_len.argtypes = [P]
    L.orc_solver_num_nodes.restype = ctypes.c_uint64
    L.orc_solver_num_nodes.argtypes = [P]
    L.orc_solver_num_hands.restype = ctypes.c_int32
    L.orc_solver_num_hands.argtypes = [P, ctypes.c_int32]
    L.orc_solver_num_combinations.restype = ctypes.c_double
    L.orc_solver_num_combinations.argtypes = [P]
    L.orc_solver_hands.argtypes = [P, ctypes.c_int32, np.ctypeslib.ndpointer(dtype=np.uint16), fp]
    L.orc_solver_get_storage.argtypes = [P, fp, fp]
    L.orc_solver_set_storage.restype = ctypes.c_int32
    L.orc_solver_set_storage.argtypes = [P, fp, fp]
    L.orc_solver_is_compressed.restype = ctypes.c_int32
    L.orc_solver_is_compressed.argtypes = [P]
    i16p, u16p = np.ctypeslib.ndpointer(dtype=np.int16), np.ctypeslib.ndpointer(dtype=np.uint16)
    L.orc_solver_get_storage_q.restype = ctypes.c_int32
    L.orc_solver_get_storage_q.argtypes = [P, i16p, u16p, fp, fp]
    L.orc_solver_set_storage_q.restype = ctypes.c_int32
    L.orc_solver_set_storage_q.argtypes = [P, i16p, u16p, fp, fp]
    L.orc_encode_block.argtypes = [ctypes.c_int32, fp, ctypes.c_uint64, i16p, u16p, fp]
    u32p, u64p, i32p = (np.ctypeslib.ndpointer(dtype=t) for t in (np.uint32, np.uint64, np.int32))
    L.orc_chacha_block.argtypes = [ctypes.c_int32, u32p, u32p]
    L.orc_stdrng_u64.argtypes = [ctypes.c_uint64, ctypes.c_uint64, u64p]
    L.orc_pcg32_words.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, u32p]
    L.orc_stream_words.argtypes = [u32p, ctypes.c_int32, ctypes.c_uint64, u32p, u64p]
    L.orc_sample_cards.restype = ctypes.c_int32
    L.orc_sample_cards.argtypes = [i32p, ctypes.c_int32, ctypes.c_int32, ctypes.c_uint64, i32p]
    L.orc_solver_iterations.restype = ctypes.c_uint32
    L.orc_solver_iterations.argtypes = [P]
    L.orc_solver_set_iterations.argtypes = [P, ctypes.c_uint32]
    L.orc_solver_iterate.argtypes = [P, ctypes.c_uint32]
    L.orc_solver_exploitability.restype = ctypes.c_double
    L.orc_solver_exploitability.argtypes = [P]
    L.orc_solver_exploitability_f64.restype = ctypes.c_double
    L.orc_solver_exploitability_f64.argtypes = [P]
    L.orc_solver_root_values.argtypes = [P, ctypes.c_int32, ctypes.c_int32, dp]
    L.orc_solver_traverse_root.argtypes = [P, ctypes.c_int32, ctypes.c_int32, ctypes.c_uint32, fp]
    L.orc_solver_traverse_root_f64.argtypes = [P, ctypes.c_int32, ctypes.c_int32, dp]
    L.orc_solver_solve.argtypes = [P, ctypes.c_uint32, ctypes.c_double, ctypes.c_uint32, ctypes.POINTER(ProgressC),
                                   ctypes.POINTER(ProgressC), ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]
    L.orc_solver_nodes.argtypes = [P] + [ctypes.c_void_p] * 12
    L.orc_solver_num_boards.restype = ctypes.c_int32
    L.orc_solver_num_boards.argtypes = [P]
    L.orc_solver_board.argtypes = [P, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]
    L.orc_solver_query.restype = ctypes.c_int32
    L.orc_solver_query.argtypes = [P, ctypes.POINTER(StepC), ctypes.c_int32] + [ctypes.c_void_p] * 12 + \
        [ctypes.c_int32, ctypes.c_int32, ctypes.c_char_p, ctypes.c_int32]
    L.orc_solver_ranges_after.restype = ctypes.c_int32
    L.orc_solver_ranges_after.argtypes = [P, ctypes.POINTER(StepC), ctypes.c_int32, fp, fp, ctypes.c_char_p, ctypes.c_int32]
    L.orc_handset_new.restype = P
    L.orc_handset_new.argtypes = [fp, fp, ctypes.POINTER(ctypes.c_int32), ctypes.c_int32]
    L.orc_handset_free.argtypes = [P]
    L.orc_handset_len.restype = ctypes.c_int32
    L.orc_handset_len.argtypes = [P, ctypes.c_int32]
    L.orc_handset_cards.argtypes = [P, ctypes.c_int32, np.ctypeslib.ndpointer(dtype=np.uint8)]
    L.orc_handset_num_combinations.restype = ctypes.c_double
    L.orc_handset_num_combinations.argtypes = [P]
    L.orc_handset_compatible_reach.argtypes = [P, ctypes.c_int32, fp, dp]
    L.orc_eval_fold.argtypes = [P, ctypes.c_int32, fp, ctypes.c_double, fp]
    L.orc_eval_showdown.argtypes = [P, ctypes.POINTER(ctypes.c_int32), ctypes.c_int32, fp, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_double, fp]
    L.orc_fold_batch1.argtypes = [P, ctypes.c_int32, fp, ctypes.c_double, fp]
    L.orc_showdown_batch1.argtypes = [P, ctypes.POINTER(ctypes.c_int32), ctypes.c_int32, fp, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_double, ctypes.c_int32, ctypes.c_double, fp, fp]


def ints(xs):
    return (ctypes.c_int32 * len(xs))(*xs)


def steps_c(path):
    """path: list of ("fold"|"check"|"call"|("raise", to)|("card", c))."""
    arr = (StepC * max(1, len(path)))()
    for i, s in enumerate(path):
        if isinstance(s, tuple) and s[0] == "card":
            arr[i].kind, arr[i].card = 1, s[1]
        elif isinstance(s, tuple) and s[0] == "raise":
            arr[i].kind, arr[i].act_kind, arr[i].to = 0, 3, s[1]
        else:
            arr[i].kind, arr[i].act_kind = 0, {"fold": 0, "check": 1, "call": 2}[s]
    return arr


ACT_NAMES = {0: "fold", 1: "check", 2: "call"}


def action_of(kind, to):
    return ("raise", int(to)) if kind == 3 else ACT_NAMES[kind]


class NodeInfo:
    def __init__(self, **kw):
        self.__dict__.update(kw)

    def hand_index(self, combo):
        idx = np.nonzero(self.combos == combo)[0]
        return int(idx[0]) if len(idx) else None

    def hand_strategy(self, i):
        return self.strategy[:, i]

    def hand_ev(self, i):
        return self.ev[:, i]

    def range_strategy(self):
        tot = float(np.sum(self.reach.astype(np.float64)))
        if tot <= 0:
            return np.zeros(len(self.actions))
        return (self.strategy.astype(np.float64) * self.reach.astype(np.float64)).sum(axis=1) / tot


class OracleSolver:
    """The reference CPU solve (Solver, crates/solver/src/lib.rs:288-913)."""

    def __init__(self, cfg: SolveConfig, parallel=True, threads=0):
        L = lib()
        self.cfg = cfg
        self._c = cfg.to_c()
        err = ctypes.create_string_buffer(256)
        self.h = L.orc_solver_new(ctypes.byref(self._c), err, 256)
        if not self.h:
            raise ValueError(err.value.decode())
        L.orc_solver_set_parallel(self.h, int(parallel), threads)
        self.n = [L.orc_solver_num_hands(self.h, p) for p in range(2)]

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_solver_free(self.h)
            self.h = None

    @staticmethod
    def estimate(cfg):
        L = lib()
        c = cfg.to_c()
        st = StatsC()
        err = ctypes.create_string_buffer(256)
        if L.orc_estimate(ctypes.byref(c), ctypes.byref(st), err, 256):
            raise ValueError(err.value.decode())
        return st

    def stats(self):
        st = StatsC()
        lib().orc_solver_stats(self.h, ctypes.byref(st))
        return st

    def set_batch_river(self, on):
        lib().orc_solver_set_batch_river(self.h, int(on))

    @property
    def storage_len(self):
        return lib().orc_solver_storage_len(self.h)

    @property
    def num_combinations(self):
        return lib().orc_solver_num_combinations(self.h)

    def hands(self, p):
        c = np.zeros(self.n[p], np.uint16)
        w = np.zeros(self.n[p], np.float32)
        lib().orc_solver_hands(self.h, p, c, w)
        return c, w

    def storage(self):
        n = self.storage_len
        r = np.zeros(n, np.float32)
        s = np.zeros(n, np.float32)
        lib().orc_solver_get_storage(self.h, r, s)
        return r, s

    def set_storage(self, r, s):
        if lib().orc_solver_set_storage(self.h, np.ascontiguousarray(r, np.float32), np.ascontiguousarray(s, np.float32)):
            raise ValueError("16-bit storage: use set_storage_q")

    @property
    def compressed(self):
        return bool(lib().orc_solver_is_compressed(self.h))

    def storage_q(self):
        """16-bit storage (storage.rs:87-106): i16 regrets, u16 sums, and the
        regret and strategy-sum scales of every player node."""
        n, k = self.storage_len, self.stats().player_nodes
        r, s = np.zeros(n, np.int16), np.zeros(n, np.uint16)
        rs, ss = np.zeros(k, np.float32), np.zeros(k, np.float32)
        if lib().orc_solver_get_storage_q(self.h, r, s, rs, ss):
            raise ValueError("f32 storage")
        return r, s, rs, ss

    def set_storage_q(self, r, s, rs, ss):
        if lib().orc_solver_set_storage_q(self.h, np.ascontiguousarray(r, np.int16), np.ascontiguousarray(s, np.uint16),
                                          np.ascontiguousarray(rs, np.float32), np.ascontiguousarray(ss, np.float32)):
            raise ValueError("f32 storage")

    @property
    def iterations(self):
        return lib().orc_solver_iterations(self.h)

    @iterations.setter
    def iterations(self, t):
        lib().orc_solver_set_iterations(self.h, t)

    def iterate(self, count=1):
        lib().orc_solver_iterate(self.h, count)

    def exploitability(self):
        return lib().orc_solver_exploitability(self.h)

    def exploitability_f64(self):
        return lib().orc_solver_exploitability_f64(self.h)

    def root_values(self, mode=2, f64=False):
        out = np.zeros(2, np.float64)
        lib().orc_solver_root_values(self.h, mode, int(f64), out)
        return out

    def traverse_root(self, p, mode, t=1):
        out = np.zeros(self.n[p], np.float32)
        lib().orc_solver_traverse_root(self.h, p, mode, t, out)
        return out

    def traverse_root_f64(self, p, mode):
        out = np.zeros(self.n[p], np.float64)
        lib().orc_solver_traverse_root_f64(self.h, p, mode, out)
        return out

    def solve(self, max_iterations, target_pct, check_every):
        out = ProgressC()
        hist = (ProgressC * 4096)()
        hl = ctypes.c_int32(0)
        lib().orc_solver_solve(self.h, max_iterations, target_pct, check_every, ctypes.byref(out), hist, 4096,
                               ctypes.byref(hl))
        return out, [hist[i] for i in range(min(hl.value, 4096))]

    def nodes(self):
        L = lib()
        m = L.orc_solver_num_nodes(self.h)
        a = dict(kind=np.zeros(m, np.uint8), street=np.zeros(m, np.uint8), num_children=np.zeros(m, np.uint16),
                 first_child=np.zeros(m, np.uint32), invest=np.zeros((m, 2), np.int32), offset=np.zeros(m, np.uint64),
                 slot=np.zeros(m, np.uint32), rake=np.zeros(m, np.float32), chance_factor=np.zeros(m, np.float32),
                 edge_card=np.zeros(m, np.int32), edge_act=np.zeros(m, np.int32), edge_to=np.zeros(m, np.int64))
        L.orc_solver_nodes(self.h, *[x.ctypes.data for x in a.values()])
        return a

    def boards(self):
        L = lib()
        out = []
        for i in range(L.orc_solver_num_boards(self.h)):
            b = (ctypes.c_int32 * 5)()
            L.orc_solver_board(self.h, i, b)
            out.append([x for x in b if x >= 0])
        return out

    def query(self, path):
        L = lib()
        st = steps_c(path)
        cap_a, cap_h = 16, 1326
        player, na, n = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        kinds = np.zeros(cap_a, np.int32)
        tos = np.zeros(cap_a, np.int64)
        combos = np.zeros(cap_h, np.uint16)
        strat = np.zeros(cap_a * cap_h, np.float32)
        ev = np.zeros(cap_a * cap_h, np.float32)
        reach = np.zeros(cap_h, np.float32)
        rf = np.zeros(2, np.float64)
        pot = ctypes.c_int64()
        inv = np.zeros(2, np.int64)
        err = ctypes.create_string_buffer(256)
        rc = L.orc_solver_query(self.h, st, len(path), ctypes.byref(player), ctypes.byref(na), ctypes.byref(n),
                                kinds.ctypes.data, tos.ctypes.data, combos.ctypes.data, strat.ctypes.data,
                                ev.ctypes.data, reach.ctypes.data, rf.ctypes.data, ctypes.addressof(pot),
                                inv.ctypes.data, cap_a, cap_h, err, 256)
        if rc:
            raise ValueError(err.value.decode())
        na, n = na.value, n.value
        return NodeInfo(player=player.value, actions=[action_of(kinds[a], tos[a]) for a in range(na)],
                        combos=combos[:n].copy(), strategy=strat[:na * n].reshape(na, n).copy(),
                        ev=ev[:na * n].reshape(na, n).copy(), reach=reach[:n].copy(), reach_fraction=rf.copy(),
                        pot=pot.value, invest=inv.copy())

    def ranges_after(self, path):
        o0 = np.zeros(1326, np.float32)
        o1 = np.zeros(1326, np.float32)
        err = ctypes.create_string_buffer(256)
        if lib().orc_solver_ranges_after(self.h, steps_c(path), len(path), o0, o1, err, 256):
            raise ValueError(err.value.decode())
        return o0, o1


def encode_block(values, sums):
    """One block through the 16-bit codec (storage.rs:253-325) from zero
    storage: (stored integers, scale)."""
    v = np.ascontiguousarray(values, np.float32)
    rq, sq, sc = np.zeros(len(v), np.int16), np.zeros(len(v), np.uint16), np.zeros(1, np.float32)
    lib().orc_encode_block(int(sums), v, len(v), rq, sq, sc)
    return (sq if sums else rq), float(sc[0])


def chacha_block(words, rounds):
    """The ChaCha block function on 16 input words (RFC 8439 section 2.3)."""
    out = np.zeros(16, np.uint32)
    lib().orc_chacha_block(rounds, np.ascontiguousarray(words, np.uint32), out)
    return out


def stdrng_u64(seed, n):
    """The first n next_u64 draws of rand 0.9's StdRng::seed_from_u64(seed)."""
    out = np.zeros(n, np.uint64)
    lib().orc_stdrng_u64(seed, n, out)
    return out


def pcg32_words(state, inc, n):
    """n PCG32 words as seed_from_u64 draws them (state advanced, then the
    XSH RR output of the new state), from the given state and increment."""
    out = np.zeros(n, np.uint32)
    lib().orc_pcg32_words(state & (2**64 - 1), inc & (2**64 - 1), n, out)
    return out


def stream_words(key, rounds, n):
    """StdRng's stream from a given 8-word key and round count: the first n
    next_u32 words, and the first n // 2 next_u64 draws of a fresh copy."""
    w = np.zeros(n, np.uint32)
    w64 = np.zeros(max(n // 2, 1), np.uint64)
    lib().orc_stream_words(np.ascontiguousarray(key, np.uint32), rounds, n, w, w64)
    return w, w64[:n // 2]


def sample_cards(available, k, seed):
    """sampling.rs's stratified sample with a fresh StdRng::seed_from_u64(seed)."""
    av = np.ascontiguousarray(available, np.int32)
    out = np.zeros(max(len(av), 1), np.int32)
    n = lib().orc_sample_cards(av, len(av), k, seed, out)
    return out[:n].tolist()


class OracleHandSet:
    def __init__(self, r0, r1, board):
        L = lib()
        self.h = L.orc_handset_new(np.asarray(r0, np.float32), np.asarray(r1, np.float32), ints(board), len(board))
        self.n = [L.orc_handset_len(self.h, p) for p in range(2)]

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_handset_free(self.h)
            self.h = None

    def cards(self, p):
        c = np.zeros(2 * self.n[p], np.uint8)
        lib().orc_handset_cards(self.h, p, c)
        return c.reshape(-1, 2)

    @property
    def num_combinations(self):
        return lib().orc_handset_num_combinations(self.h)

    def compatible_reach(self, p, opp):
        out = np.zeros(self.n[p], np.float64)
        lib().orc_handset_compatible_reach(self.h, p, np.asarray(opp, np.float32), out)
        return out

    def eval_fold(self, p, opp, pay):
        out = np.zeros(self.n[p], np.float32)
        lib().orc_eval_fold(self.h, p, np.asarray(opp, np.float32), pay, out)
        return out

    def eval_showdown(self, board5, p, opp, w, l, t):
        out = np.zeros(self.n[p], np.float32)
        lib().orc_eval_showdown(self.h, ints(board5), p, np.asarray(opp, np.float32), w, l, t, out)
        return out

    def fold_batch(self, p, opp, pay):
        out = np.zeros(self.n[p], np.float32)
        lib().orc_fold_batch1(self.h, p, np.asarray(opp, np.float32), pay, out)
        return out

    def showdown_batch(self, board5, p, opp, w, l, t, fold_pay=None):
        out = np.zeros(self.n[p], np.float32)
        fo = np.zeros(self.n[p], np.float32)
        lib().orc_showdown_batch1(self.h, ints(board5), p, np.asarray(opp, np.float32), w, l, t,
                                  int(fold_pay is not None), fold_pay or 0.0, out, fo)
        return out, (fo if fold_pay is not None else None)


def evaluate(cards):
    return lib().orc_evaluate(ints(cards), len(cards))


def evaluate_batch(cards):
    """evaluate_cards over the rows of a (count, ncards) uint8 array."""
    cards = np.ascontiguousarray(cards, np.uint8)
    out = np.zeros(cards.shape[0], np.uint32)
    lib().orc_evaluate_batch(cards.ctypes.data, cards.shape[1], cards.shape[0], out.ctypes.data)
    return out


def enumerate_hands(ncards, first, count):
    """Up to `count` ncards-card hands (ascending cards) from lexicographic
    rank `first` on: a (k, ncards) uint8 array."""
    out = np.zeros((count, ncards), np.uint8)
    k = lib().orc_enumerate_hands(ncards, first, count, out.ctypes.data)
    return out[:k]


def count_categories(ncards):
    counts = (ctypes.c_uint64 * 9)()
    distinct = ctypes.c_uint64()
    lib().orc_count_categories(ncards, counts, ctypes.byref(distinct))
    return list(counts), distinct.value


def postflop_actions(street, pot, to_call, current_bet, max_to, min_raise_to, num_raises, menu):
    from pts.config import MenuC, _menu_to_c
    m = MenuC()
    _menu_to_c(menu, m)
    kinds = (ctypes.c_int32 * 16)()
    tos = (ctypes.c_int64 * 16)()
    n = lib().orc_postflop_actions(street, pot, to_call, current_bet, max_to, min_raise_to, num_raises, ctypes.byref(m),
                                   kinds, tos, 16)
    return [action_of(kinds[i], tos[i]) for i in range(n)]


def parse_range(s):
    out = np.zeros(1326, np.float32)
    if lib().orc_parse_range(s.encode(), out):
        raise ValueError(s)
    return out


def discount(t, alpha=1.5, beta=0.0, gamma=2.0, restart=False, unweighted=False):
    out = (ctypes.c_float * 3)()
    lib().orc_discount(t, alpha, int(beta is not None), beta or 0.0, gamma, int(restart), int(unweighted), out)
    return tuple(out)


def next_check_gap(prev, cur, target, max_gap):
    has = prev is not None
    t0, e0 = prev if has else (0, 0.0)
    return lib().orc_next_check_gap(int(has), t0, e0, cur[0], cur[1], target, max_gap)
