"""Cancelling GPU solves (Solver::solve_cancellable, crates/solver/src/
lib.rs:480-494,614-630; the reference's gpu/engine/tests.rs:298-316): a flag
set before the solve runs nothing, the progress callback stops at the check
it returns False from, a flag set by another thread stops the solve within
one iteration, and the storage of a stopped solve is the storage of that many
iterations, which a later solve continues."""
import ctypes
import threading
import time

import numpy as np
import pytest

import pyoracle as O
from pts import spots
from pts.native import EXACT
from pts.solver import Solver, SolveOptions

pytestmark = pytest.mark.gpu


def test_cancel_flag_set_before_the_solve():
    s = Solver(spots.river_spot(0))
    pr, cancelled = s.solve_cancellable(SolveOptions(100, 0.5, 10), ctypes.c_uint8(1))
    assert cancelled and s.iterations == 0 and pr.iteration == 0 and pr.finished == 1


def test_progress_callback_stops_and_a_later_solve_continues():
    cfg = spots.river_spot(4)
    s = Solver(cfg, precision=EXACT)
    checks = []
    pr, cancelled = s.solve_cancellable(SolveOptions(1000, 0.0, 10),
                                        on_progress=lambda p: checks.append(p.iteration) or len(checks) < 2)
    assert cancelled and checks == [10, 20] and s.iterations == 20 and pr.iteration == 20
    o = O.OracleSolver(cfg)
    o.iterate(20)
    rg, sg = s.storage()
    rc, sc = o.storage()
    assert np.array_equal(rg, rc) and np.array_equal(sg, sc)
    pr, cancelled = s.solve_cancellable(SolveOptions(35, 0.0, 10))
    assert not cancelled and s.iterations == 35 and pr.iteration == 35
    o.iterate(15)
    rg, sg = s.storage()
    rc, sc = o.storage()
    assert np.array_equal(rg, rc) and np.array_equal(sg, sc)
    assert s.exploitability() == o.exploitability_f64()


def test_cancel_flag_from_another_thread():
    s = Solver(spots.turn_spot(0))
    flag = ctypes.c_uint8(0)
    timer = threading.Timer(0.3, lambda: setattr(flag, "value", 1))
    timer.start()
    t0 = time.perf_counter()
    pr, cancelled = s.solve_cancellable(SolveOptions(1_000_000, 0.0, 10), flag)
    dt = time.perf_counter() - t0
    timer.join()
    assert cancelled and 0 < s.iterations < 1_000_000 and dt < 5.0
    # the solver stays usable: the same iteration count and exploitability readable
    n = s.iterations
    s.iterate(1)
    assert s.iterations == n + 1 and np.isfinite(s.exploitability())
