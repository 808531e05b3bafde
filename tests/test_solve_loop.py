"""The solve loop (pts_run_solve: Solver::run_solve, crates/solver/src/
lib.rs:576-642) driven by host callbacks, no GPU: the check schedule against
the oracle's pinned next_check_gap (lib.rs:155-172), cancellation before and
during a solve and by the progress callback (lib.rs:482-494,614-630; the
reference's engine/tests.rs:298-316 cancels a GPU solve the same way), and
error propagation."""
import ctypes

import pytest

import pyoracle as O
from pts.native import PtsError
from pts.solver import SolveOptions, run_solve

POT = 1000.0


def expected_checks(e_of, opts, start=0):
    """The reference loop restated with the oracle's next_check_gap."""
    t, nxt, last, out = start, start + max(opts.check_every, 1), None, []
    target = opts.target_exploitability_pct * POT / 100.0
    while t < opts.max_iterations:
        t += 1
        end = t >= opts.max_iterations
        if t >= nxt or end:
            e = e_of(t)
            out.append(t)
            if 100.0 * e / POT <= opts.target_exploitability_pct or end:
                break
            nxt = t + O.next_check_gap(last, (t, e), target, max(opts.check_every, 1))
            last = (t, e)
    return out


@pytest.mark.parametrize("k,every,max_it", [(0.5, 10, 1000), (1.0, 10, 1000), (1.3, 25, 5000), (0.2, 10, 60)])
def test_check_schedule_follows_the_reference(k, every, max_it):
    def e_of(t):
        return 400.0 * t ** -k
    seen, done = [], []
    opts = SolveOptions(max_it, 0.5, every)
    pr, cancelled, t, e = run_solve(lambda t: done.append(t), lambda: e_of(len(done)), POT, opts,
                                    on_progress=lambda p: seen.append(p.iteration) is None)
    assert not cancelled
    assert done == list(range(1, t + 1))
    assert seen == expected_checks(e_of, opts)
    assert pr.iteration == seen[-1] == t and pr.finished == 1
    assert e == e_of(t) and pr.exploitability_pct == pytest.approx(100.0 * e / POT)


def test_cancel_before_the_first_iteration():
    flag = ctypes.c_uint8(1)
    done = []
    pr, cancelled, t, e = run_solve(done.append, lambda: 1.0, POT, SolveOptions(100, 0.5, 10), iterations=7,
                                    exploitability=3.0, cancel=flag)
    assert cancelled and done == [] and t == 7 and e == 3.0
    assert pr.iteration == 7 and pr.finished == 1 and pr.exploitability == 3.0


def test_cancel_during_a_solve_stops_within_one_iteration():
    flag = ctypes.c_uint8(0)
    done = []

    def it(t):
        done.append(t)
        if t == 23:
            flag.value = 1  # another thread would set it; the loop reads it before the next iteration

    pr, cancelled, t, e = run_solve(it, lambda: 100.0, POT, SolveOptions(1000, 0.5, 10), cancel=flag)
    assert cancelled and done == list(range(1, 24)) and t == 23
    assert pr.iteration == 20 and pr.finished == 1  # the last check's progress


def test_progress_callback_stops_the_solve():
    calls = []
    pr, cancelled, t, _ = run_solve(lambda t: None, lambda: 100.0, POT, SolveOptions(1000, 0.5, 10),
                                    on_progress=lambda p: calls.append(p.iteration) or len(calls) < 3)
    assert cancelled and calls == [10, 20, 30] and t == 30 and pr.iteration == 30 and pr.finished == 1


def test_target_and_cap():
    pr, cancelled, t, _ = run_solve(lambda t: None, lambda: 4.0, POT, SolveOptions(1000, 0.5, 10))
    assert not cancelled and t == 10 and pr.exploitability_pct == pytest.approx(0.4)
    pr, cancelled, t, _ = run_solve(lambda t: None, lambda: 400.0, POT, SolveOptions(15, 0.5, 10))
    assert not cancelled and t == 15 and pr.iteration == 15  # checked at the cap
    pr, cancelled, t, _ = run_solve(lambda t: None, lambda: 400.0, POT, SolveOptions(5, 0.5, 10), iterations=5)
    assert not cancelled and t == 5  # nothing left to run


def test_errors_end_the_loop():
    def bad(t):
        if t == 4:
            raise RuntimeError("iteration failed")

    with pytest.raises(RuntimeError, match="iteration failed"):
        run_solve(bad, lambda: 1.0, POT, SolveOptions(100, 0.5, 10))
    with pytest.raises(PtsError, match="pot"):
        run_solve(lambda t: None, lambda: 1.0, 0.0, SolveOptions(100, 0.5, 10))
