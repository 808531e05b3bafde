"""`Solver` and `Engine`: the solver crate's API over libpts (HIP, gfx950).

`Solver` mirrors crates/solver/src/lib.rs:288-913 (new / estimate / solve /
query / ranges_after / exploitability); `Engine` is the lower boundary that
replaces crates/solver/src/gpu/engine.rs: it takes a caller-built tree and
storage (pts_tree_desc) and runs iterations, best responses and subtree
evaluations on the GPU.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from .config import ConfigC, SolveConfig
from .native import PROGRESS_FN, ProgressC, StatsC, TreeDescC, check, lib, steps_c

ACT_NAMES = {0: "fold", 1: "check", 2: "call"}


def _action(kind, to):
    return ("raise", int(to)) if kind == 3 else ACT_NAMES[int(kind)]


@dataclass
class SolveOptions:
    """lib.rs:137-153"""
    max_iterations: int = 1000
    target_exploitability_pct: float = 0.5
    check_every: int = 10


class NodeInfo:
    """lib.rs:219-275"""

    def __init__(self, **kw):
        self.__dict__.update(kw)

    def num_hands(self):
        return len(self.combos)

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


class Solver:
    """A built game whose storage lives in HBM; solves run in HIP kernels."""

    def __init__(self, cfg: SolveConfig, device: int = 0, precision: int = 0, scratch_budget: int = 0, _handle=None):
        """precision: 0 fast (default, as the reference's GpuPrecision), 1 exact.
        scratch_budget: bytes of reach / value rows (0: the HBM left after the
        storage); the turn or river subtrees run in groups that fit."""
        self.cfg = cfg
        self._c = cfg.to_c()
        h = ctypes.c_void_p()
        if _handle is not None:
            h = _handle
        else:
            check(lib().pts_solver_new(ctypes.byref(self._c), device, precision, scratch_budget, ctypes.byref(h)))
        self.h = h
        self.n = [lib().pts_solver_num_hands(self.h, p) for p in range(2)]

    def __del__(self):
        if getattr(self, "h", None):
            lib().pts_solver_free(self.h)
            self.h = None

    def close(self):
        self.__del__()

    @staticmethod
    def estimate(cfg: SolveConfig) -> StatsC:
        c = cfg.to_c()
        st = StatsC()
        check(lib().pts_solver_estimate(ctypes.byref(c), ctypes.byref(st)))
        return st

    def stats(self) -> StatsC:
        st = StatsC()
        lib().pts_solver_stats(self.h, ctypes.byref(st))
        return st

    @property
    def storage_len(self):
        return lib().pts_solver_storage_len(self.h)

    def num_groups(self, f64: bool = False) -> int:
        """Chunk groups a traversal runs in (DCFR rows, or FP64 best-response rows)."""
        return lib().pts_solver_num_groups(self.h, int(f64))

    def jit_segments(self) -> int:
        """Segments whose DCFR up walk runs in a shape-specialised kernel
        (compiled at creation, in a plan of at least 8192 walk segments, for
        shapes of at least 64 of them; PTS_JIT_MIN_SEGMENTS=n sets the
        per-shape count for any plan; PTS_JIT=0 turns them off)."""
        return int(lib().pts_solver_jit_segments(self.h))

    @property
    def iterations(self):
        return lib().pts_solver_iterations(self.h)

    @iterations.setter
    def iterations(self, t):
        lib().pts_solver_set_iterations(self.h, t)

    def solve(self, opts: SolveOptions, history: Optional[list] = None):
        out = ProgressC()
        hist = (ProgressC * 4096)()
        hl = ctypes.c_int32(0)
        check(lib().pts_solver_solve(self.h, opts.max_iterations, opts.target_exploitability_pct, opts.check_every,
                                     ctypes.byref(out), hist, 4096, ctypes.byref(hl)))
        if history is not None:
            history.extend(hist[i] for i in range(min(hl.value, 4096)))
        return out

    def solve_cancellable(self, opts: SolveOptions, cancel: Optional[ctypes.c_uint8] = None, on_progress=None):
        """Solver::solve_cancellable (lib.rs:480-494): `cancel` is a ctypes
        c_uint8 another thread may set to 1 (the solve stops within one
        iteration); `on_progress(progress)` runs after every check and stops
        the solve by returning False. Returns (progress, cancelled)."""
        out = ProgressC()

        def cb(p, _user):
            return 1 if on_progress is None or on_progress(p.contents) is not False else 0

        fn = PROGRESS_FN(cb)
        st = lib().pts_solver_solve_cancellable(self.h, opts.max_iterations, opts.target_exploitability_pct,
                                                opts.check_every, ctypes.byref(cancel) if cancel is not None else None,
                                                fn, None, ctypes.byref(out))
        if st == 7:  # PTS_CANCELLED
            return out, True
        check(st)
        return out, False

    def iterate(self, count: int = 1):
        check(lib().pts_solver_iterate(self.h, count))

    def exploitability(self) -> float:
        x = ctypes.c_double()
        check(lib().pts_solver_exploitability(self.h, ctypes.byref(x)))
        return x.value

    def root_values(self, mode: int = 2):
        out = np.zeros(2, np.float64)
        check(lib().pts_solver_root_values(self.h, mode, out))
        return out

    def profile_iteration(self):
        """The next DCFR iteration with HIP events around every kernel launch:
        ({class: ms}, {class: launches}, wall_ms) for the classes begin, down,
        terminal, up."""
        ms = np.zeros(4, np.float64)
        ln = (ctypes.c_uint32 * 4)()
        wall = ctypes.c_double()
        check(lib().pts_solver_profile_iteration(self.h, ms, ln, ctypes.byref(wall)))
        names = ("begin", "down", "terminal", "up")
        return dict(zip(names, ms.tolist())), dict(zip(names, list(ln))), wall.value

    def storage(self):
        n = self.storage_len
        r = np.zeros(n, np.float32)
        s = np.zeros(n, np.float32)
        check(lib().pts_solver_get_storage(self.h, r, s))
        return r, s

    def set_storage(self, r, s):
        check(lib().pts_solver_set_storage(self.h, np.ascontiguousarray(r, np.float32), np.ascontiguousarray(s, np.float32)))

    @property
    def compressed(self) -> bool:
        """16-bit storage (storage.rs:74-81; lib.rs:346-352 for Auto)."""
        return bool(lib().pts_solver_is_compressed(self.h))

    def storage_q(self):
        """16-bit storage: (i16 regrets, u16 sums, regret scales, strategy-sum
        scales), one scale of each per player node."""
        n, k = self.storage_len, self.stats().player_nodes
        r, s = np.zeros(n, np.int16), np.zeros(n, np.uint16)
        rs, ss = np.zeros(k, np.float32), np.zeros(k, np.float32)
        check(lib().pts_solver_get_storage_q(self.h, r, s, rs, ss))
        return r, s, rs, ss

    def set_storage_q(self, r, s, rs, ss):
        check(lib().pts_solver_set_storage_q(self.h, np.ascontiguousarray(r, np.int16), np.ascontiguousarray(s, np.uint16),
                                             np.ascontiguousarray(rs, np.float32), np.ascontiguousarray(ss, np.float32)))

    def nodes(self):
        m = lib().pts_solver_num_nodes(self.h)
        a = dict(kind=np.zeros(m, np.uint8), street=np.zeros(m, np.uint8), num_children=np.zeros(m, np.uint16),
                 first_child=np.zeros(m, np.uint32), invest=np.zeros((m, 2), np.int32), offset=np.zeros(m, np.uint64),
                 rake=np.zeros(m, np.float32), chance_factor=np.zeros(m, np.float32), card=np.zeros(m, np.int8),
                 edge_act=np.zeros(m, np.int32), edge_to=np.zeros(m, np.int64))
        lib().pts_solver_nodes(self.h, *[x.ctypes.data for x in a.values()])
        return a

    def query(self, path) -> NodeInfo:
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
        check(lib().pts_solver_query(self.h, steps_c(path), len(path), ctypes.byref(player), ctypes.byref(na),
                                     ctypes.byref(n), kinds.ctypes.data, tos.ctypes.data, combos.ctypes.data,
                                     strat.ctypes.data, ev.ctypes.data, reach.ctypes.data, rf.ctypes.data,
                                     ctypes.addressof(pot), inv.ctypes.data, cap_a, cap_h))
        na, n = na.value, n.value
        return NodeInfo(player=player.value, actions=[_action(kinds[a], tos[a]) for a in range(na)],
                        combos=combos[:n].copy(), strategy=strat[:na * n].reshape(na, n).copy(),
                        ev=ev[:na * n].reshape(na, n).copy(), reach=reach[:n].copy(), reach_fraction=rf.copy(),
                        pot=pot.value, invest=inv.copy())

    def ranges_after(self, path):
        o0 = np.zeros(1326, np.float32)
        o1 = np.zeros(1326, np.float32)
        check(lib().pts_solver_ranges_after(self.h, steps_c(path), len(path), o0, o1))
        return o0, o1


def run_solve(iterate, measure, pot: float, opts: SolveOptions, iterations: int = 0,
              exploitability: float = float("inf"), cancel: Optional[ctypes.c_uint8] = None, on_progress=None):
    """pts_run_solve (Solver::run_solve, lib.rs:576-642) over Python
    callables: iterate(t) runs iteration t, measure() returns the
    exploitability in chips. Returns (progress, cancelled, iterations,
    exploitability)."""
    from .native import ITERATE_FN, MEASURE_FN
    errors = []

    def it(_u, t):
        try:
            iterate(int(t))
            return 0
        except Exception as e:  # noqa: BLE001 - handed back after the C loop returns
            errors.append(e)
            return 5

    def me(_u, x):
        try:
            x[0] = float(measure())
            return 0
        except Exception as e:  # noqa: BLE001
            errors.append(e)
            return 5

    def pr(p, _u):
        return 1 if on_progress is None or on_progress(p.contents) is not False else 0

    fi, fm, fp = ITERATE_FN(it), MEASURE_FN(me), PROGRESS_FN(pr)
    t = ctypes.c_uint32(iterations)
    e = ctypes.c_double(exploitability)
    out = ProgressC()
    st = lib().pts_run_solve(ctypes.byref(t), ctypes.byref(e), pot, opts.max_iterations, opts.target_exploitability_pct,
                             opts.check_every, ctypes.byref(cancel) if cancel is not None else None, fi, fm, fp, None,
                             ctypes.byref(out))
    if errors:
        raise errors[0]
    if st == 7:
        return out, True, t.value, e.value
    check(st)
    return out, False, t.value, e.value


def depth_limited(cfg: SolveConfig, net=None, device: int = 0, precision: int = 0) -> Solver:
    """A Solver over the depth-limited tree of `cfg` (the root street's betting,
    cut points valued by the value network `net`, a pts.net.ValueNet; None:
    zeros)."""
    c = cfg.to_c()
    h = ctypes.c_void_p()
    check(lib().pts_dl_new(ctypes.byref(c), device, precision, net.h if net is not None else None, ctypes.byref(h)))
    s = Solver(cfg, _handle=h)
    s._net = net  # keep the network alive while the solver uses it
    return s


def copy_root_street(dst: Solver, src: Solver, regret_scale: float = 1.0, sum_scale: float = 1.0) -> int:
    n = ctypes.c_uint32()
    check(lib().pts_solver_copy_root_street(dst.h, src.h, regret_scale, sum_scale, ctypes.byref(n)))
    return n.value


def pin_root_street(dst: Solver, src: Solver) -> int:
    """Pins dst's root-street decision nodes to src's average strategy
    (pts_solver_pin_root_street); returns the number of nodes pinned."""
    n = ctypes.c_uint32()
    check(lib().pts_solver_pin_root_street(dst.h, src.h, ctypes.byref(n)))
    return n.value


class TreeArrays:
    """The arrays of a pts_tree_desc, kept alive while an Engine uses them."""

    def __init__(self, kind, street, num_children, first_child, invest, offset, rake, chance_factor, card, boards,
                 hand_cards, hand_weights, hand_same, half_pot, num_combinations, storage_len, batch_river=True):
        self.a = dict(kind=np.ascontiguousarray(kind, np.uint8), street=np.ascontiguousarray(street, np.uint8),
                      num_children=np.ascontiguousarray(num_children, np.uint16),
                      first_child=np.ascontiguousarray(first_child, np.uint32),
                      invest=np.ascontiguousarray(invest, np.int32).reshape(-1),
                      offset=np.ascontiguousarray(offset, np.uint64), rake=np.ascontiguousarray(rake, np.float32),
                      chance_factor=np.ascontiguousarray(chance_factor, np.float32),
                      card=np.ascontiguousarray(card, np.int8),
                      boards=np.ascontiguousarray(np.asarray(boards, np.int32).reshape(-1)))
        self.hc = [np.ascontiguousarray(x, np.uint8).reshape(-1) for x in hand_cards]
        self.hw = [np.ascontiguousarray(x, np.float32) for x in hand_weights]
        self.hs = [np.ascontiguousarray(x, np.uint32) for x in hand_same]
        d = TreeDescC()
        d.num_nodes = len(self.a["kind"])
        for k in ("kind", "street", "num_children", "first_child", "invest", "offset", "rake", "chance_factor", "card",
                  "boards"):
            setattr(d, k, self.a[k].ctypes.data)
        d.num_boards = len(self.a["boards"]) // 5
        for q in range(2):
            d.num_hands[q] = len(self.hw[q])
            d.hand_cards[q] = self.hc[q].ctypes.data
            d.hand_weights[q] = self.hw[q].ctypes.data
            d.hand_same[q] = self.hs[q].ctypes.data
        d.half_pot, d.num_combinations, d.storage_len = half_pot, num_combinations, storage_len
        d.batch_river = int(batch_river)
        self.desc = d


class Engine:
    """pts_engine_*: the GPU backend boundary (crates/solver/src/gpu/engine.rs)."""

    def __init__(self, tree: TreeArrays, regrets, sums, device: int = 0, precision: int = 0, scratch_budget: int = 0,
                 q16=None):
        """f32 storage (regrets, sums), or with q16 = (slot per node, i16
        regrets, u16 sums, regret scales, strategy-sum scales) 16-bit storage
        (HostStorage::Compressed)."""
        self.tree = tree
        h = ctypes.c_void_p()
        if q16 is None:
            # regrets / sums None: zeros
            arrs = [None if x is None else np.ascontiguousarray(x, np.float32) for x in (regrets, sums)]
            check(lib().pts_engine_create(device, ctypes.byref(tree.desc), *[None if a is None else a.ctypes.data
                                                                             for a in arrs],
                                          precision, scratch_budget, ctypes.byref(h)))
            self.slots = 0
        else:
            slot, r, s, rs, ss = q16
            self.slots = len(slot) if rs is None else len(rs)
            arrs = None if r is None else [np.ascontiguousarray(r, np.int16), np.ascontiguousarray(s, np.uint16),
                                           np.ascontiguousarray(rs, np.float32), np.ascontiguousarray(ss, np.float32)]
            self._q = arrs  # kept alive across the call
            ptrs = [None] * 4 if arrs is None else [a.ctypes.data for a in arrs]
            check(lib().pts_engine_create_q(device, ctypes.byref(tree.desc), np.ascontiguousarray(slot, np.uint32),
                                            len(slot) if rs is None else len(rs), *ptrs, precision, scratch_budget,
                                            ctypes.byref(h)))
        self.h = h
        self.storage_len = tree.desc.storage_len
        self.n = [tree.desc.num_hands[0], tree.desc.num_hands[1]]

    def __del__(self):
        if getattr(self, "h", None):
            lib().pts_engine_destroy(self.h)
            self.h = None

    def iterate(self, d):
        check(lib().pts_engine_iterate(self.h, d[0], d[1], d[2], int(d[3]) if len(d) > 3 else 1))

    def traverse(self, p, mode, d=(0.0, 0.0, 0.0, 1)):
        out = np.zeros(self.n[p], np.float64)
        check(lib().pts_engine_traverse(self.h, p, mode, d[0], d[1], d[2], int(d[3]), out, self.n[p]))
        return out

    def root_values(self, mode):
        out = np.zeros(2, np.float64)
        check(lib().pts_engine_root_values(self.h, mode, out))
        return out

    def exploitability(self, raked=False):
        x = ctypes.c_double()
        check(lib().pts_engine_exploitability(self.h, int(raked), ctypes.byref(x)))
        return x.value

    def download(self):
        r = np.zeros(self.storage_len, np.float32)
        s = np.zeros(self.storage_len, np.float32)
        check(lib().pts_engine_download(self.h, r, s, self.storage_len))
        return r, s

    def download_q(self):
        r, s = np.zeros(self.storage_len, np.int16), np.zeros(self.storage_len, np.uint16)
        rs, ss = np.zeros(self.slots, np.float32), np.zeros(self.slots, np.float32)
        check(lib().pts_engine_download_q(self.h, r, s, rs, ss))
        return r, s, rs, ss

    def upload(self, r, s):
        check(lib().pts_engine_upload(self.h, np.ascontiguousarray(r, np.float32), np.ascontiguousarray(s, np.float32),
                                      self.storage_len))

    def subtree_values(self, node, player, opp_reach, na):
        out = np.zeros(na * self.n[player], np.float32)
        opp = np.ascontiguousarray(opp_reach, np.float32)
        check(lib().pts_engine_subtree_values(self.h, node, player, opp, len(opp), out, len(out)))
        return out.reshape(na, self.n[player])

    def device_bytes(self):
        return lib().pts_engine_device_bytes(self.h)

    def num_groups(self, f64: bool = False) -> int:
        return lib().pts_engine_num_groups(self.h, int(f64))

    def jit_segments(self) -> int:
        return int(lib().pts_engine_jit_segments(self.h))


class BatchSolver:
    """pts_batch_*: many independent spots solved together on one GPU (one
    hipGraph launch per DCFR iteration for all of them)."""

    def __init__(self, cfgs: List[SolveConfig], device: int = 0, precision: int = 0):
        self.cfgs = list(cfgs)
        arr = (ConfigC * len(self.cfgs))()
        for i, c in enumerate(self.cfgs):
            arr[i] = c.to_c()
        self._arr = arr
        h = ctypes.c_void_p()
        check(lib().pts_batch_new(arr, len(self.cfgs), device, precision, ctypes.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            lib().pts_batch_free(self.h)
            self.h = None

    def __len__(self):
        return len(self.cfgs)

    def host_info(self):
        """(merged nodes, storage cells, scratch bytes of DCFR rows, of FP64 rows)."""
        nn, sl = ctypes.c_uint64(), ctypes.c_uint64()
        sc = (ctypes.c_uint64 * 2)()
        check(lib().pts_batch_host_info(self.h, ctypes.byref(nn), ctypes.byref(sl), sc))
        return nn.value, sl.value, sc[0], sc[1]

    @property
    def iterations(self):
        return lib().pts_batch_iterations(self.h)

    def stats(self, k) -> StatsC:
        st = StatsC()
        check(lib().pts_batch_stats(self.h, k, ctypes.byref(st)))
        return st

    def iterate(self, count: int = 1):
        check(lib().pts_batch_iterate(self.h, count))

    def exploitability(self):
        out = np.zeros(len(self.cfgs), np.float64)
        check(lib().pts_batch_exploitability(self.h, out))
        return out

    def solve(self, max_iterations: int, target_pct: float, check_every: int = 10):
        out = (ProgressC * len(self.cfgs))()
        check(lib().pts_batch_solve(self.h, max_iterations, target_pct, check_every, out))
        return [out[i] for i in range(len(self.cfgs))]

    def storage(self, k):
        n = lib().pts_batch_storage_len(self.h, k)
        r = np.zeros(n, np.float32)
        s = np.zeros(n, np.float32)
        check(lib().pts_batch_get_storage(self.h, k, r, s))
        return r, s

    def autotune(self, iterations: int = 3):
        best = (ctypes.c_uint32 * 2)()
        ms = ctypes.c_double()
        check(lib().pts_batch_autotune(self.h, iterations, best, ctypes.byref(ms)))
        return (best[0], best[1]), ms.value

    def root_rows(self, mode: int = 2):
        """Per-hand root values of every spot, combo order: (spots, 2, 1326)."""
        out = np.zeros(len(self.cfgs) * 2 * 1326, np.float64)
        check(lib().pts_batch_root_rows(self.h, mode, out))
        return out.reshape(len(self.cfgs), 2, 1326)

    def num_combinations(self, k):
        return lib().pts_batch_num_combinations(self.h, k)

    def set_tiles(self, run_length: int, down_threads: int):
        check(lib().pts_batch_set_tiles(self.h, run_length, down_threads))

    def profile_iteration(self):
        ms = np.zeros(4, np.float64)
        ln = (ctypes.c_uint32 * 4)()
        wall = ctypes.c_double()
        check(lib().pts_batch_profile_iteration(self.h, ms, ln, ctypes.byref(wall)))
        names = ("begin", "down", "terminal", "up")
        return dict(zip(names, ms.tolist())), dict(zip(names, list(ln))), wall.value
