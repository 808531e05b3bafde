"""Loads libpts.so (the HIP build for gfx950) and declares the C ABI of
include/pts.h. There is no CPU fallback: without the library, or without a
GPU, the calls fail loudly."""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .config import ConfigC

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG, "libpts.so")

_lib = None


class PtsError(RuntimeError):
    """A pts_status other than PTS_OK (crates/solver/src/gpu/mod.rs:86-136)."""

    NAMES = {1: "DriverMissing", 2: "DriverTooOld", 3: "NoDevice", 4: "Driver", 5: "InvalidInput",
             6: "OutOfMemory", 7: "Cancelled"}

    def __init__(self, status, msg, hip_error=0, needed=0, available=0):
        super().__init__(f"{self.NAMES.get(status, status)}: {msg}")
        self.status = status
        self.hip_error = hip_error  # GpuError::Driver's error code
        self.needed = needed        # GpuError::OutOfMemory { needed, available }
        self.available = available


class StatsC(ctypes.Structure):
    _fields_ = [("player_nodes", ctypes.c_uint64), ("chance_nodes", ctypes.c_uint64), ("terminal_nodes", ctypes.c_uint64),
                ("storage_cells", ctypes.c_uint64), ("memory_bytes", ctypes.c_uint64)]


class ProgressC(ctypes.Structure):
    _fields_ = [("iteration", ctypes.c_uint32), ("max_iterations", ctypes.c_uint32), ("exploitability", ctypes.c_double),
                ("exploitability_pct", ctypes.c_double), ("finished", ctypes.c_int32), ("_pad", ctypes.c_int32)]


PROGRESS_FN = ctypes.CFUNCTYPE(ctypes.c_int32, ctypes.POINTER(ProgressC), ctypes.c_void_p)
ITERATE_FN = ctypes.CFUNCTYPE(ctypes.c_int32, ctypes.c_void_p, ctypes.c_uint32)
MEASURE_FN = ctypes.CFUNCTYPE(ctypes.c_int32, ctypes.c_void_p, ctypes.POINTER(ctypes.c_double))


class StepC(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("act_kind", ctypes.c_int32), ("to", ctypes.c_int64), ("card", ctypes.c_int32),
                ("_pad", ctypes.c_int32)]


class TreeDescC(ctypes.Structure):
    _fields_ = [("num_nodes", ctypes.c_uint64), ("kind", ctypes.c_void_p), ("street", ctypes.c_void_p),
                ("num_children", ctypes.c_void_p), ("first_child", ctypes.c_void_p), ("invest", ctypes.c_void_p),
                ("offset", ctypes.c_void_p), ("rake", ctypes.c_void_p), ("chance_factor", ctypes.c_void_p),
                ("card", ctypes.c_void_p), ("num_boards", ctypes.c_uint32), ("boards", ctypes.c_void_p),
                ("num_hands", ctypes.c_uint32 * 2), ("hand_cards", ctypes.c_void_p * 2),
                ("hand_weights", ctypes.c_void_p * 2), ("hand_same", ctypes.c_void_p * 2),
                ("half_pot", ctypes.c_double), ("num_combinations", ctypes.c_double), ("storage_len", ctypes.c_uint64),
                ("batch_river", ctypes.c_int32), ("_pad", ctypes.c_int32)]


EXPORTS = [
    "pts_abi_sizeof", "pts_last_error", "pts_device_count", "pts_backend_wants_gpu", "pts_engine_create", "pts_engine_iterate",
    "pts_engine_traverse", "pts_engine_root_values", "pts_engine_exploitability", "pts_engine_download",
    "pts_engine_upload", "pts_engine_subtree_values", "pts_engine_device_bytes", "pts_engine_destroy",
    "pts_solver_estimate", "pts_solver_new", "pts_solver_free", "pts_solver_stats", "pts_solver_solve",
    "pts_solver_iterate", "pts_solver_exploitability", "pts_solver_root_values", "pts_solver_iterations",
    "pts_solver_set_iterations", "pts_solver_storage_len", "pts_solver_num_hands", "pts_solver_get_storage",
    "pts_solver_set_storage", "pts_solver_profile_iteration", "pts_engine_profile_iteration", "pts_solver_num_nodes", "pts_solver_nodes", "pts_solver_query",
    "pts_solver_ranges_after", "pts_engine_num_groups", "pts_solver_num_groups", "pts_solver_plan_groups",
    "pts_batch_new", "pts_batch_free", "pts_batch_num_spots", "pts_batch_iterations", "pts_batch_stats",
    "pts_batch_storage_len", "pts_batch_get_storage", "pts_batch_iterate", "pts_batch_exploitability", "pts_batch_solve",
    "pts_batch_autotune", "pts_batch_set_tiles", "pts_batch_profile_iteration", "pts_batch_host_info",
    "pts_net_num_params", "pts_net_forward_flops_per_row", "pts_net_create", "pts_net_destroy", "pts_net_set_params",
    "pts_net_get_params", "pts_net_get_grads", "pts_net_forward", "pts_net_gradients", "pts_net_train_step", "pts_net_profile",
    "pts_dl_new", "pts_solver_num_leaves", "pts_solver_copy_root_street", "pts_solver_pin_root_street", "pts_batch_root_rows",
    "pts_batch_num_combinations", "pts_solver_is_compressed", "pts_solver_get_storage_q", "pts_solver_set_storage_q",
    "pts_engine_create_q", "pts_engine_download_q", "pts_last_error_detail", "pts_evaluate",
    "pts_solver_solve_cancellable", "pts_run_solve", "pts_net_forward_noalloc", "pts_engine_jit_segments",
    "pts_solver_jit_segments",
]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it (python -c 'import __graft_entry__ as g; g.build()')")
        _lib = ctypes.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


P = ctypes.c_void_p
f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def _declare(L):
    L.pts_abi_sizeof.restype = ctypes.c_size_t
    L.pts_abi_sizeof.argtypes = [ctypes.c_int]
    L.pts_last_error.restype = ctypes.c_size_t
    L.pts_last_error.argtypes = [ctypes.c_char_p, ctypes.c_size_t]
    L.pts_last_error_detail.argtypes = [ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_uint64),
                                        ctypes.POINTER(ctypes.c_uint64)]
    L.pts_device_count.argtypes = [ctypes.POINTER(ctypes.c_int)]
    L.pts_evaluate.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_void_p]
    L.pts_backend_wants_gpu.argtypes = [ctypes.c_int, ctypes.c_int]
    L.pts_engine_create.argtypes = [ctypes.c_int, ctypes.POINTER(TreeDescC), ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_int, ctypes.c_uint64, ctypes.POINTER(P)]
    L.pts_engine_iterate.argtypes = [P, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_int]
    L.pts_engine_traverse.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                      ctypes.c_int, f64p, ctypes.c_size_t]
    L.pts_engine_root_values.argtypes = [P, ctypes.c_int, f64p]
    L.pts_engine_exploitability.argtypes = [P, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    L.pts_engine_download.argtypes = [P, f32p, f32p, ctypes.c_uint64]
    L.pts_engine_upload.argtypes = [P, f32p, f32p, ctypes.c_uint64]
    L.pts_engine_subtree_values.argtypes = [P, ctypes.c_uint32, ctypes.c_int, f32p, ctypes.c_size_t, f32p, ctypes.c_size_t]
    L.pts_engine_device_bytes.restype = ctypes.c_uint64
    L.pts_engine_device_bytes.argtypes = [P]
    L.pts_engine_destroy.argtypes = [P]
    L.pts_solver_estimate.argtypes = [ctypes.POINTER(ConfigC), ctypes.POINTER(StatsC)]
    L.pts_solver_new.argtypes = [ctypes.POINTER(ConfigC), ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.POINTER(P)]
    L.pts_solver_num_groups.restype = ctypes.c_uint32
    L.pts_solver_num_groups.argtypes = [P, ctypes.c_int]
    L.pts_solver_plan_groups.argtypes = [P, ctypes.c_uint64, ctypes.c_int, ctypes.POINTER(ctypes.c_uint32),
                                         ctypes.POINTER(ctypes.c_uint64)]
    L.pts_batch_new.argtypes = [ctypes.POINTER(ConfigC), ctypes.c_uint32, ctypes.c_int, ctypes.c_int, ctypes.POINTER(P)]
    L.pts_batch_free.argtypes = [P]
    L.pts_batch_num_spots.restype = ctypes.c_uint32
    L.pts_batch_num_spots.argtypes = [P]
    L.pts_batch_iterations.restype = ctypes.c_uint32
    L.pts_batch_iterations.argtypes = [P]
    L.pts_batch_stats.argtypes = [P, ctypes.c_uint32, ctypes.POINTER(StatsC)]
    L.pts_batch_storage_len.restype = ctypes.c_uint64
    L.pts_batch_storage_len.argtypes = [P, ctypes.c_uint32]
    L.pts_batch_get_storage.argtypes = [P, ctypes.c_uint32, f32p, f32p]
    L.pts_batch_iterate.argtypes = [P, ctypes.c_uint32]
    L.pts_batch_exploitability.argtypes = [P, f64p]
    L.pts_batch_solve.argtypes = [P, ctypes.c_uint32, ctypes.c_double, ctypes.c_uint32, ctypes.POINTER(ProgressC)]
    L.pts_batch_autotune.argtypes = [P, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_double)]
    L.pts_batch_set_tiles.argtypes = [P, ctypes.c_uint32, ctypes.c_uint32]
    L.pts_batch_profile_iteration.argtypes = [P, f64p, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_double)]
    L.pts_batch_host_info.argtypes = [P, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64),
                                      ctypes.POINTER(ctypes.c_uint64)]
    L.pts_net_num_params.restype = ctypes.c_uint64
    L.pts_net_num_params.argtypes = []
    L.pts_net_forward_flops_per_row.restype = ctypes.c_double
    L.pts_net_forward_flops_per_row.argtypes = []
    L.pts_net_create.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint64, ctypes.POINTER(P)]
    L.pts_net_destroy.argtypes = [P]
    L.pts_net_set_params.argtypes = [P, f32p]
    L.pts_net_get_params.argtypes = [P, f32p]
    L.pts_net_get_grads.argtypes = [P, f32p]
    L.pts_net_forward.argtypes = [P, f32p, ctypes.c_uint32, f32p]
    L.pts_net_forward_noalloc.argtypes = [P, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p]
    L.pts_net_gradients.argtypes = [P, f32p, f32p, f32p, ctypes.c_uint32, ctypes.POINTER(ctypes.c_double)]
    L.pts_net_train_step.argtypes = [P, f32p, f32p, f32p, ctypes.c_uint32, ctypes.c_float, ctypes.POINTER(ctypes.c_double)]
    L.pts_net_profile.argtypes = [P, ctypes.c_uint32, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    L.pts_dl_new.argtypes = [ctypes.POINTER(ConfigC), ctypes.c_int, ctypes.c_int, P, ctypes.POINTER(P)]
    L.pts_solver_num_leaves.restype = ctypes.c_uint32
    L.pts_solver_num_leaves.argtypes = [P]
    L.pts_solver_copy_root_street.argtypes = [P, P, ctypes.c_float, ctypes.c_float, ctypes.POINTER(ctypes.c_uint32)]
    L.pts_solver_pin_root_street.argtypes = [P, P, ctypes.POINTER(ctypes.c_uint32)]
    L.pts_batch_root_rows.argtypes = [P, ctypes.c_int, f64p]
    L.pts_batch_num_combinations.restype = ctypes.c_double
    L.pts_batch_num_combinations.argtypes = [P, ctypes.c_uint32]
    L.pts_engine_num_groups.restype = ctypes.c_uint32
    L.pts_engine_num_groups.argtypes = [P, ctypes.c_int]
    L.pts_engine_jit_segments.argtypes = [P]
    L.pts_engine_jit_segments.restype = ctypes.c_uint64
    L.pts_solver_jit_segments.argtypes = [P]
    L.pts_solver_jit_segments.restype = ctypes.c_uint64
    L.pts_solver_free.argtypes = [P]
    L.pts_solver_stats.argtypes = [P, ctypes.POINTER(StatsC)]
    L.pts_solver_solve.argtypes = [P, ctypes.c_uint32, ctypes.c_double, ctypes.c_uint32, ctypes.POINTER(ProgressC),
                                   ctypes.POINTER(ProgressC), ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]
    L.pts_solver_solve_cancellable.argtypes = [P, ctypes.c_uint32, ctypes.c_double, ctypes.c_uint32, ctypes.c_void_p,
                                               PROGRESS_FN, ctypes.c_void_p, ctypes.POINTER(ProgressC)]
    L.pts_run_solve.argtypes = [ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_double), ctypes.c_double,
                                ctypes.c_uint32, ctypes.c_double, ctypes.c_uint32, ctypes.c_void_p, ITERATE_FN,
                                MEASURE_FN, PROGRESS_FN, ctypes.c_void_p, ctypes.POINTER(ProgressC)]
    L.pts_solver_iterate.argtypes = [P, ctypes.c_uint32]
    L.pts_solver_exploitability.argtypes = [P, ctypes.POINTER(ctypes.c_double)]
    L.pts_solver_root_values.argtypes = [P, ctypes.c_int, f64p]
    L.pts_solver_iterations.restype = ctypes.c_uint32
    L.pts_solver_iterations.argtypes = [P]
    L.pts_solver_set_iterations.argtypes = [P, ctypes.c_uint32]
    L.pts_solver_storage_len.restype = ctypes.c_uint64
    L.pts_solver_storage_len.argtypes = [P]
    L.pts_solver_num_hands.restype = ctypes.c_int32
    L.pts_solver_num_hands.argtypes = [P, ctypes.c_int]
    L.pts_solver_get_storage.argtypes = [P, f32p, f32p]
    L.pts_solver_set_storage.argtypes = [P, f32p, f32p]
    i16p, u16p = np.ctypeslib.ndpointer(dtype=np.int16), np.ctypeslib.ndpointer(dtype=np.uint16)
    L.pts_solver_is_compressed.restype = ctypes.c_int32
    L.pts_solver_is_compressed.argtypes = [P]
    L.pts_solver_get_storage_q.argtypes = [P, i16p, u16p, f32p, f32p]
    L.pts_solver_set_storage_q.argtypes = [P, i16p, u16p, f32p, f32p]
    L.pts_engine_create_q.argtypes = [ctypes.c_int, ctypes.c_void_p, np.ctypeslib.ndpointer(dtype=np.uint32),
                                      ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64, ctypes.POINTER(ctypes.c_void_p)]
    L.pts_engine_download_q.argtypes = [P, i16p, u16p, f32p, f32p]
    L.pts_solver_profile_iteration.argtypes = [P, f64p, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_double)]
    L.pts_engine_profile_iteration.argtypes = [P, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_int, f64p,
                                               ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_double)]
    L.pts_solver_num_nodes.restype = ctypes.c_uint64
    L.pts_solver_num_nodes.argtypes = [P]
    L.pts_solver_nodes.argtypes = [P] + [ctypes.c_void_p] * 11
    L.pts_solver_query.argtypes = [P, ctypes.POINTER(StepC), ctypes.c_int32] + [ctypes.c_void_p] * 12 + \
        [ctypes.c_int32, ctypes.c_int32]
    L.pts_solver_ranges_after.argtypes = [P, ctypes.POINTER(StepC), ctypes.c_int32, f32p, f32p]


FAST, EXACT = 0, 1


def last_error() -> str:
    buf = ctypes.create_string_buffer(1024)
    lib().pts_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def last_error_detail():
    """(hip_error, needed, available) of the last failure on this thread."""
    e, n, a = ctypes.c_int32(), ctypes.c_uint64(), ctypes.c_uint64()
    lib().pts_last_error_detail(ctypes.byref(e), ctypes.byref(n), ctypes.byref(a))
    return e.value, n.value, a.value


def check(status: int):
    if status != 0:
        raise PtsError(status, last_error(), *last_error_detail())


def evaluate(cards) -> np.ndarray:
    """pts_evaluate over the rows of a (count, ncards) array of cards
    (evaluate_cards, crates/holdem/src/eval.rs:142-198)."""
    c = np.ascontiguousarray(cards, np.uint8)
    out = np.zeros(c.shape[0], np.uint32)
    check(lib().pts_evaluate(c.ctypes.data, c.shape[1], c.shape[0], out.ctypes.data))
    return out


def device_count() -> int:
    n = ctypes.c_int(0)
    st = lib().pts_device_count(ctypes.byref(n))
    return n.value if st == 0 else 0


def steps_c(path):
    """path: list of "fold" | "check" | "call" | ("raise", to) | ("card", c)."""
    arr = (StepC * max(1, len(path)))()
    for i, s in enumerate(path):
        if isinstance(s, tuple) and s[0] == "card":
            arr[i].kind, arr[i].card = 1, s[1]
        elif isinstance(s, tuple) and s[0] == "raise":
            arr[i].kind, arr[i].act_kind, arr[i].to = 0, 3, s[1]
        else:
            arr[i].kind, arr[i].act_kind = 0, {"fold": 0, "check": 1, "call": 2}[s]
    return arr
