"""Benchmark: time to 0.5%-of-pot exploitability of the exact HU DCFR solve on
one MI355X, with the reference CPU solve (the oracle) timed on the host.

Default workload: the flop spot (3-card board, 16.19 G storage cells, 1176
hands per player), the configuration the headline metric is quoted on.

A step is one DCFR iteration (both players' traversals) over the whole tree.
`value` is the wall time of a fresh solve from zero storage until the first
FP64 best-response check at or below the target (checks included; tree build
and upload excluded and reported as build_s). `ms_per_step` is the mean of K
timed iterations after W warmup iterations. `roofline` is for the kernel class
with the largest share of an iteration's event time: algorithmic bytes per
launch (down: 12 B per storage cell of the opponent's nodes; up: 8 B per cell
of the traverser's nodes; terminal: 4 B per opponent hand per terminal) over
its mean launch duration, both measured with HIP events on the engine's
stream; `traffic` is the class's HBM bytes per launch as rocprofv3 counted
them (profiles/traffic.json). `roofline.iteration` and `roofline.solve` give
the whole path's algorithmic bytes (20 B per storage cell plus 4 B per
opponent hand per terminal per iteration; 4 B per cell plus the terminal term
per FP64 check) over the timed iteration and over the solve.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
                       [--workload flop|turn|river|batch|valuenet|river-auto] [--precision fast|exact]
                       [--storage f32|16bit]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "poker-trainer_amd"))

METRIC = "time to 0.5%-pot exploitability, exact vs value-net DCFR, HU flop, 1 MI355X"
STREET_NAME = {"river": "HU river", "turn": "HU turn", "flop": "HU flop"}
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    "turn": ("turn HU exact DCFR: seeded 4-card board, pot 20bb, stacks 90bb, bets 33%/75%/all-in, raises "
             "2.5x/all-in, one raise (cap 2), 48 river cards; ranges U(0,1) per combo with 30% zeroed; f32 regrets "
             "and sums, alpha 1.5 beta 0 gamma 2", "turn_spot", 500),
    "river": ("river HU exact DCFR: seeded 5-card board, pot 10bb, stacks 40bb, bets 50%/100%/all-in, raise 2.5x, "
              "3 bets; ranges U(0,1) per combo with 40% zeroed; f32, alpha 1.5 beta 0 gamma 2", "river_spot", 1000),
    "flop": ("flop HU exact DCFR: seeded 3-card board, pot 6bb, stacks 97bb, bets 33%/75%/all-in, raises 2.5x/all-in, "
             "3 bets per street, 49 turn x 48 river cards; ranges from 169-class weights sigmoid(score + N(0,0.5)); f32 "
             "regrets and sums (16.19 G cells), alpha 1.5 beta 0 gamma 2", "flop_spot", 1000),
}
BATCH_DESC = ("batched HU exact DCFR: 256 river + 64 turn spots, seeded boards, per-range nonzero combos U[80,1326], "
              "SPR log-uniform 0.5-10, 1-3 bet sizes from {25,33,50,75,100,150}% + all-in, raises 2.5x/all-in, 3 bets; "
              "one hipGraph launch per iteration for all spots, autotuned tiles; f32")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def solve_logged(s, opts, hist=None):
    """Solver.solve with a line per exploitability check on stderr (the
    progress callback of solve_cancellable; it runs only at the checks)."""
    seen = hist if hist is not None else []

    def progress(p):
        seen.append(type(p).from_buffer_copy(p))
        log(f"[bench] check: iteration {p.iteration}, {p.exploitability_pct:.4f}% pot")
        return True

    pr, cancelled = s.solve_cancellable(opts, on_progress=progress)
    assert not cancelled
    return pr


def measured_traffic(workload, precision, kernel):
    """HBM bytes per launch of a kernel class as rocprofv3 counted them
    (FETCH_SIZE and WRITE_SIZE, scaled by a plain copy's counts), recorded in
    profiles/traffic.json with the commands that measured them; None when
    that workload was not counted."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            e = json.load(f).get(f"{workload}/{precision}/{kernel}")
    except (OSError, ValueError):
        return None
    return None if e is None else e.get("hbm_bytes_per_launch")


def cpu_baseline(cfg, iters_to_target, checks, seconds):
    """The oracle (C++ restatement of the reference CPU solve, parallel over
    chance children like the reference's rayon split) on the host cores: a
    bounded sample of iterations plus one exploitability check, scaled to the
    GPU run's iteration and check counts."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    o = pyoracle.OracleSolver(cfg, parallel=True, threads=threads)
    n = 0
    t0 = time.perf_counter()
    while True:
        o.iterate(1)
        n += 1
        if time.perf_counter() - t0 > seconds or n >= iters_to_target:
            break
    t_iter = (time.perf_counter() - t0) / n
    t1 = time.perf_counter()
    o.exploitability()
    t_check = time.perf_counter() - t1
    est = iters_to_target * t_iter + checks * t_check
    return {"value": round(est, 4), "unit": "s", "cores": threads, "kind": "port",
            "sample": f"{n} DCFR iterations ({t_iter * 1e3:.1f} ms each) and one best-response check "
                      f"({t_check * 1e3:.1f} ms) of the same spot, scaled to the GPU run's {iters_to_target} "
                      f"iterations and {checks} checks"}


def cpu_baseline_flop(cfg, iters_to_target, checks, seconds, flop_cells):
    """The flop's oracle solve needs ~130 GB of storage and hours of host
    time, so it is extrapolated: the oracle (parallel over chance children)
    runs a bounded sample of DCFR iterations and one exploitability check on
    the flop's largest turn subtree (the flop checked through: the flop
    board plus its first free card, pot 600, stacks 9700, the same ranges
    and menu), and the times are scaled by the flop's storage cells over the
    subtree's, then to the GPU run's iteration and check counts."""
    import copy
    from pts.config import BettingState, TURN
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    sub = copy.deepcopy(cfg)
    turn_card = next(c for c in range(52) if c not in cfg.board)
    sub.board = list(cfg.board) + [turn_card]
    sub.root = BettingState.street_start(TURN, cfg.root.pot, cfg.root.stack[0], cfg.root.stack[1])
    o = pyoracle.OracleSolver(sub, parallel=True, threads=threads)
    sub_cells = o.storage_len
    n = 0
    t0 = time.perf_counter()
    while True:
        o.iterate(1)
        n += 1
        if time.perf_counter() - t0 > seconds:
            break
    t_iter = (time.perf_counter() - t0) / n
    t1 = time.perf_counter()
    o.exploitability()
    t_check = time.perf_counter() - t1
    scale = flop_cells / sub_cells
    est = scale * (iters_to_target * t_iter + checks * t_check)
    return {"value": round(est, 1), "unit": "s", "cores": threads, "kind": "port",
            "sample": f"extrapolation: the oracle ran {n} DCFR iterations ({t_iter * 1e3:.1f} ms each) and one "
                      f"best-response check ({t_check * 1e3:.1f} ms) of the flop's checked-through turn subtree "
                      f"({sub_cells} cells), scaled by the flop's {flop_cells} cells ({scale:.1f}x) to the GPU run's "
                      f"{iters_to_target} iterations and {checks} checks"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="flop", choices=sorted(WORKLOADS) + ["batch", "valuenet", "river-auto"])
    ap.add_argument("--subgames", type=int, default=1024, help="valuenet: turn subgames solved for training")
    ap.add_argument("--flops", type=int, default=2, help="valuenet: flops of the flop workload judged (seeds 0..)")
    ap.add_argument("--train-steps", type=int, default=600, help="valuenet: Adam steps of 1024 rows")
    ap.add_argument("--resolve-iterations", type=int, default=300,
                    help="valuenet: most iterations of the later streets' re-solve before the exact fallback")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--target", type=float, default=0.5)
    ap.add_argument("--check-every", type=int, default=0,
                    help="most iterations between exploitability checks (0: the reference trainer's rule, "
                         "crates/trainer/src/review.rs:1065-1071: 20 for trees above 200,000 nodes, else 10)")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--max-iterations", type=int, default=0,
                    help="cap the solve at this many iterations (0: the workload's own count; a capped run that "
                         "stops short of the target says so in target_reached)")
    ap.add_argument("--net-only", action="store_true",
                    help="valuenet: only time the network's forward and training kernels on a synthetic batch")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB (5 launches) and exit: the known-size "
                         "kernel that scales the FETCH_SIZE / WRITE_SIZE counters")
    ap.add_argument("--precision", default="fast", choices=["fast", "exact"],
                    help="fast: block-scan terminal sums (default, as the reference's GPU default); exact: the "
                         "reference's summation order, bit-identical to the CPU solve")
    ap.add_argument("--storage", default="f32", choices=["f32", "16bit"],
                    help="regret / strategy-sum storage: f32 (default), or 16-bit (i16 regrets, u16 sums, one f32 "
                         "scale per decision node: half the bytes per cell)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (this build has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)

    if args.copy_probe_mib:
        import torch
        n = args.copy_probe_mib * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0.record()
        for _ in range(4):
            b.copy_(a)
        t1.record()
        torch.cuda.synchronize()
        ms = t0.elapsed_time(t1) / 4
        moved = 2 * n * 4  # read and written once per copy
        log(f"[bench] copy probe: 5 copies of {args.copy_probe_mib} MiB; the last 4 at {ms:.3f} ms each, "
            f"{moved / ms / 1e9:.2f} TB/s read + written")
        return

    from pts import spots
    from pts.solver import Solver, SolveOptions

    if args.workload == "batch":
        return bench_batch(args)
    if args.workload == "river-auto":
        return bench_river_auto(args)
    if args.workload == "valuenet":
        return bench_valuenet(args)
    desc, maker, max_iters = WORKLOADS[args.workload]
    if args.max_iterations:
        max_iters = min(max_iters, args.max_iterations)
    cfg = getattr(spots, maker)(args.seed)
    q16 = args.storage == "16bit"
    if q16:
        cfg.storage_mode = 1
        desc = desc.replace("f32 regrets and sums", "16-bit regrets and sums").replace("; f32,", "; 16-bit storage,")

    # steps: W warmup iterations, then K timed ones
    t0 = time.perf_counter()
    prec = 1 if args.precision == "exact" else 0
    s = Solver(cfg, precision=prec)
    build_s = time.perf_counter() - t0
    st = s.stats()
    n0, n1 = s.n
    log(f"[bench] {args.workload}: hands {n0}/{n1}, {st.player_nodes} decision nodes, {st.terminal_nodes} terminals, "
        f"{st.storage_cells} cells; build+upload {build_s:.2f}s")
    s.iterate(args.warmup)
    t0 = time.perf_counter()
    s.iterate(args.steps)
    step_s = (time.perf_counter() - t0) / max(args.steps, 1)
    log(f"[bench] {args.steps} timed iterations: {step_s * 1e3:.3f} ms each")
    ms, launches, wall = s.profile_iteration()
    log(f"[bench] profiled iteration: {ms} launches {launches} wall {wall:.3f} ms")
    del s

    # the metric: a fresh solve to the target
    if not args.check_every:
        args.check_every = 20 if st.player_nodes + st.chance_nodes + st.terminal_nodes > 200_000 else 10
    s = Solver(cfg, precision=prec)
    hist = []
    t0 = time.perf_counter()
    pr = solve_logged(s, SolveOptions(max_iters, args.target, args.check_every), hist)
    solve_s = time.perf_counter() - t0
    log(f"[bench] solve: {pr.iteration} iterations, {pr.exploitability_pct:.4f}% pot, {solve_s:.3f}s, "
        f"{len(hist)} checks")

    cells = st.storage_cells
    # per cell: down reads regrets and reads and writes sums, up reads and
    # writes regrets (4 bytes each, 2 with 16-bit storage)
    cb = 2.0 if q16 else 4.0
    bytes_cls = {"down": 3 * cb * cells, "up": 2 * cb * cells, "terminal": 4.0 * st.terminal_nodes * (n0 + n1)}
    dom = max(("down", "terminal", "up"), key=lambda k: ms[k])
    achieved = bytes_cls[dom] / (ms[dom] * 1e-3) / 1e9 if ms[dom] > 0 else 0.0
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5),
                "traffic": None if q16 else measured_traffic(args.workload, args.precision, dom), "kernel": dom,
                "launches_per_iteration": launches[dom], "avg_launch_ms": round(ms[dom] / max(launches[dom], 1), 5),
                "algorithmic_bytes_per_launch": round(bytes_cls[dom] / max(launches[dom], 1)),
                "iteration_ms_by_class": {k: round(v, 4) for k, v in ms.items()}}
    # the whole path's algorithmic bytes: 20 B per cell (10 with 16-bit
    # storage) + 4 B per opponent hand per terminal per iteration; 4 B (2) per
    # cell + the same terminal term per FP64 check
    term_b = 4.0 * st.terminal_nodes * (n0 + n1)
    b_iter = 5 * cb * cells + term_b
    b_br = cb * cells + term_b
    solve_b = b_iter * pr.iteration + b_br * len(hist)
    roofline["iteration"] = {"algorithmic_bytes": round(b_iter), "ms": round(step_s * 1e3, 4),
                             "achieved": round(b_iter / step_s / 1e9, 2),
                             "frac": round(b_iter / step_s / 1e9 / HBM_PEAK_GBS, 5)}
    roofline["solve"] = {"algorithmic_bytes": round(solve_b), "s": round(solve_s, 4),
                         "achieved": round(solve_b / solve_s / 1e9, 2),
                         "frac": round(solve_b / solve_s / 1e9 / HBM_PEAK_GBS, 5)}

    reached = pr.exploitability_pct <= args.target
    cpu = None
    if not args.no_cpu:
        if args.workload == "flop":
            cpu = cpu_baseline_flop(cfg, pr.iteration, len(hist), args.cpu_seconds, cells)
        else:
            cpu = cpu_baseline(cfg, pr.iteration, len(hist), args.cpu_seconds)
        log(f"[bench] cpu baseline: {cpu}")

    metric = METRIC if args.workload == "flop" else \
        f"time to {args.target:g}%-pot exploitability, exact DCFR, {STREET_NAME[args.workload]}, 1 MI355X"
    line = {
        "metric": metric, "value": round(solve_s, 4), "unit": "s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(step_s * 1e3, 4), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": desc, "seed": args.seed, "hands": [n0, n1], "decision_nodes": st.player_nodes,
                   "terminals": st.terminal_nodes, "storage_cells": cells, "target_pct": args.target,
                   "check_every": args.check_every, "precision": args.precision, "storage": args.storage,
                   "iterations_to_target": pr.iteration,
                   "target_reached": reached, "final_exploitability_pct": round(pr.exploitability_pct, 5),
                   "iterations_per_s": round(1.0 / step_s, 2), "build_s": round(build_s, 3),
                   "checks": len(hist), "parallelism": "1 GPU"},
        "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def bench_batch(args):
    """spots/s to the target for the batched workload: autotuned tiles, then
    one batched solve (checks every 10 iterations, a spot frozen at its
    target) timed from the first iteration to the last spot's check."""
    from pts import spots
    from pts.solver import BatchSolver
    prec = 1 if args.precision == "exact" else 0
    args.check_every = args.check_every or 10
    cfgs = spots.batch_spots()
    t0 = time.perf_counter()
    b = BatchSolver(cfgs, precision=prec)
    build_s = time.perf_counter() - t0
    sts = [b.stats(k) for k in range(len(cfgs))]
    cells = sum(st.storage_cells for st in sts)
    hands = [[sum(1 for x in c.ranges[p] if x > 0) for p in range(2)] for c in cfgs]
    log(f"[bench] batch: {len(cfgs)} spots, {cells} cells, {sum(st.terminal_nodes for st in sts)} terminals; "
        f"build+upload {build_s:.2f}s")
    (rl, dt), tune_ms = b.autotune(3)
    log(f"[bench] autotuned tiles: {rl} terminals per terminal workgroup, {dt} hands per down workgroup "
        f"({tune_ms:.3f} ms/iteration)")
    b.iterate(args.warmup)
    t0 = time.perf_counter()
    b.iterate(args.steps)
    step_s = (time.perf_counter() - t0) / max(args.steps, 1)
    ms, launches, wall = b.profile_iteration()
    log(f"[bench] {args.steps} timed iterations: {step_s * 1e3:.3f} ms each; profile {ms} {launches}")
    del b
    b = BatchSolver(cfgs, precision=prec)
    b.set_tiles(rl, dt)
    t0 = time.perf_counter()
    pr = b.solve(1000, args.target, args.check_every)
    solve_s = time.perf_counter() - t0
    reached = sum(1 for p in pr if p.exploitability_pct <= args.target)
    iters = [p.iteration for p in pr]
    log(f"[bench] batched solve: {reached}/{len(cfgs)} spots at <= {args.target}% pot, {solve_s:.3f}s, "
        f"iterations {min(iters)}..{max(iters)}")
    # roofline of the dominant kernel class over one profiled iteration
    terms = sum(st.terminal_nodes * (h[0] + h[1]) for st, h in zip(sts, hands))
    bytes_cls = {"down": 12.0 * cells, "up": 8.0 * cells, "terminal": 4.0 * terms}
    dom = max(("down", "terminal", "up"), key=lambda k: ms[k])
    achieved = bytes_cls[dom] / (ms[dom] * 1e-3) / 1e9 if ms[dom] > 0 else 0.0
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": measured_traffic("batch", args.precision, dom),
                "kernel": dom,
                "launches_per_iteration": launches[dom], "avg_launch_ms": round(ms[dom] / max(launches[dom], 1), 5),
                "algorithmic_bytes_per_launch": round(bytes_cls[dom] / max(launches[dom], 1)),
                "iteration_ms_by_class": {k: round(v, 4) for k, v in ms.items()}}
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_batch(cfgs, iters, args.cpu_seconds)
        log(f"[bench] cpu baseline: {cpu}")
    line = {
        "metric": "spots/s to 0.5%-pot exploitability, batched HU DCFR (256 river + 64 turn spots), 1 MI355X",
        "value": round(len(cfgs) / solve_s, 3), "unit": "spots/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(step_s * 1e3, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": BATCH_DESC, "spots": len(cfgs), "storage_cells": cells, "target_pct": args.target,
                   "check_every": args.check_every, "spots_reaching_target": reached, "solve_s": round(solve_s, 4),
                   "iterations_min_max": [min(iters), max(iters)], "tiles": {"terminal_run": rl, "down_hands": dt},
                   "precision": args.precision, "build_s": round(build_s, 3), "parallelism": "1 GPU"},
        "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def bench_valuenet(args):
    """The value-network path against the exact solve on flops of the flop
    workload (its tree and 169-class ranges; seeds 0..flops-1).

    Training (not counted in the per-flop times, as it serves every flop):
    `subgames` turn subgames, each a random turn card under one of the
    evaluated flops' leaf pots and that flop's ranges scaled per class,
    solved by the batched GPU DCFR to 0.2% of their pot; the network trained
    on them by its own kernels (Adam, batches of 1024 rows, suits relabelled
    at random).

    Per flop, the value-net path: the depth-limited solve (300 iterations,
    leaves valued by the network), then the exact check: the full tree with the
    flop pinned to the depth-limited average strategy and the later streets
    re-solved by exact DCFR until the exact FP64 best response of that whole
    profile is at or below the target. The exact path: the same tree solved
    from scratch to the same target."""
    import numpy as np
    from pts import spots
    from pts import valuenet as V
    from pts.net import ValueNet
    from pts.solver import Solver, SolveOptions, depth_limited
    if args.net_only:
        net = ValueNet(1024, seed=1)
        net.profile(1024)
        fwd_ms, train_ms = min(net.profile(1024) for _ in range(5))
        flops = ValueNet.forward_flops_per_row() * 1024
        tf = flops / (fwd_ms * 1e-3) / 1e12
        line = {"metric": "value-network forward, BF16 MFMA with FP32 accumulation, 1024 rows", "value": round(tf, 2),
                "unit": "TFLOP/s", "n_gpus": 1, "steps": 5, "warmup": 1, "ms_per_step": round(fwd_ms, 4),
                "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16", "data": "synthetic",
                "config": {"workload": "network 2706-4x1024-2652, He-normal weights, one batch of 1024 rows",
                           "train_step_ms": round(train_ms, 4),
                           "train_step_tflops": round(3 * flops / (train_ms * 1e-3) / 1e12, 2)},
                "roofline": {"bound": "mfma", "achieved": round(tf, 2), "peak": 2500.0, "unit": "TFLOP/s",
                             "frac": round(tf / 2500.0, 5), "traffic": None, "kernel": "gemm_kernel (forward)"},
                "cpu_baseline": None}
        print(json.dumps(line), flush=True)
        return
    cfgs = [spots.flop_spot(k) for k in range(args.flops)]
    if not args.check_every:
        # the reference trainer's rule, as for the flop workload
        st = Solver.estimate(cfgs[0])
        args.check_every = 20 if st.player_nodes + st.chance_nodes + st.terminal_nodes > 200_000 else 10
    # training rows: subgames spread over the evaluated flops
    X, Y, M, its, gen_s = [], [], [], [], 0.0
    per = -(-args.subgames // len(cfgs))
    for k, cfg in enumerate(cfgs):
        x, y, m, info = V.generate(per, 0, V.leaf_pots(cfg), batch=64, base=cfg.ranges, seed0=k * 1_000_000,
                                   log=log)
        X.append(x), Y.append(y), M.append(m)
        its += info["iterations"]
        gen_s += info["seconds"]
    X, Y, M = np.concatenate(X), np.concatenate(Y), np.concatenate(M)
    log(f"[bench] {len(X)} turn subgames solved in {gen_s:.1f}s (mean {np.mean(its):.0f} iterations to 0.2% pot)")
    n = X.shape[0]
    ntr = n * 7 // 8
    net = ValueNet(1024, seed=1)
    t0 = time.perf_counter()
    V.train(net, X[:ntr], Y[:ntr], M[:ntr], args.train_steps, batch=1024, lr=1e-3, log=log)
    train_s = time.perf_counter() - t0
    held = V.heldout_loss(net, X[ntr:], Y[ntr:], M[ntr:])
    zero = float(np.sum(M[ntr:] * np.where(np.abs(Y[ntr:]) <= 1, 0.5 * Y[ntr:] ** 2, np.abs(Y[ntr:]) - 0.5)) /
                 np.sum(M[ntr:]))
    fwd_ms, train_ms = net.profile(1024)
    flops = ValueNet.forward_flops_per_row() * 1024
    mfma = {"bound": "mfma", "achieved": round(flops / (fwd_ms * 1e-3) / 1e12, 2), "peak": 2500.0, "unit": "TFLOP/s",
            "frac": round(flops / (fwd_ms * 1e-3) / 1e12 / 2500.0, 5), "traffic": None,
            "kernel": "gemm_kernel (forward, 1024 rows)", "forward_ms": round(fwd_ms, 4),
            "train_step_tflops": round(3 * flops / (train_ms * 1e-3) / 1e12, 2)}
    log(f"[bench] network: {args.train_steps} steps of 1024 rows in {train_s:.1f}s, held-out loss {held:.5f} "
        f"(zero output {zero:.5f}), forward {fwd_ms:.3f} ms / 1024 rows = {mfma['achieved']} TFLOP/s")
    per_flop = []
    for k, cfg in enumerate(cfgs):
        t0 = time.perf_counter()
        dl = depth_limited(cfg, net)
        dl.iterate(300)
        dl_s = time.perf_counter() - t0
        hist, rs_s, full, reached = V.resolve(cfg, dl, args.target, args.resolve_iterations, args.check_every, log=log)
        del full
        log(f"[bench] flop {k}: depth-limited {dl_s:.1f}s, re-solve {rs_s:.1f}s to {hist[-1][1]:.4f}% pot; exact solve next")
        t0 = time.perf_counter()
        ex = Solver(cfg)
        ex_hist = []
        pr = solve_logged(ex, SolveOptions(1000, args.target, args.check_every), ex_hist)
        ex_s = time.perf_counter() - t0
        cells = ex.stats().storage_cells
        del ex
        # a profile that misses the target falls back to the exact solve
        vn_s = dl_s + rs_s + (0.0 if reached else ex_s)
        r = {"flop": k, "value_net_s": round(vn_s, 2), "depth_limited_s": round(dl_s, 2), "resolve_s": round(rs_s, 2),
             "resolve_iterations": hist[-1][0], "value_net_exploitability_pct": round(hist[-1][1], 4),
             "reached": reached, "fell_back": not reached, "exact_s": round(ex_s, 2), "exact_iterations": pr.iteration,
             "exact_exploitability_pct": round(pr.exploitability_pct, 4), "exact_checks": len(ex_hist),
             "cells": cells, "speedup": round(ex_s / vn_s, 3)}
        per_flop.append(r)
        log(f"[bench] flop {k}: {r}")
    vn = sum(r["value_net_s"] for r in per_flop)
    ex = sum(r["exact_s"] for r in per_flop)
    cpu = None
    if not args.no_cpu:
        # the exact solve's CPU baseline, on the first flop
        r0 = per_flop[0]
        cpu = cpu_baseline_flop(cfgs[0], r0["exact_iterations"], r0["exact_checks"], args.cpu_seconds, r0["cells"])
        log(f"[bench] cpu baseline (exact solve of flop 0): {cpu}")
    line = {
        "metric": METRIC, "value": round(vn / len(per_flop), 3), "unit": "s", "n_gpus": 1, "steps": len(per_flop),
        "warmup": 0, "ms_per_step": None, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16+f32", "data": "synthetic",
        "config": {"workload": f"value-net depth-limited flop solve judged by the exact FP64 best response of the whole "
                               f"profile (flop pinned to the depth-limited average, later streets re-solved by exact "
                               f"DCFR): {len(per_flop)} flops of the flop workload (seeds 0-{len(per_flop) - 1}); "
                               f"network 2706-4x1024-2652 trained on {n} GPU-solved turn subgames, batches of 1024",
                   "per_flop": per_flop, "value_net_mean_s": round(vn / len(per_flop), 3),
                   "exact_mean_s": round(ex / len(per_flop), 3), "speedup_vs_exact": round(ex / vn, 3),
                   "target_pct": args.target, "check_every": args.check_every, "training_subgames": n,
                   "training_data_s": round(gen_s, 2), "training_s": round(train_s, 2),
                   "heldout_loss": round(held, 6), "zero_output_loss": round(zero, 6), "parallelism": "1 GPU"},
        "roofline": mfma, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def cpu_baseline_batch(cfgs, iters, seconds):
    """The oracle on a bounded sample of the batch (every 16th river spot and
    every 8th turn spot), each run for the iterations the GPU solve gave it
    plus one check per 10 iterations, scaled to the whole batch."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    sample = [k for k in range(len(cfgs)) if (k < 256 and k % 16 == 0) or (k >= 256 and (k - 256) % 8 == 0)]
    spent, covered, done = 0.0, 0, []
    for k in sample:
        t0 = time.perf_counter()
        o = pyoracle.OracleSolver(cfgs[k], parallel=True, threads=threads)
        n = 0
        while n < iters[k]:
            o.iterate(min(10, iters[k] - n))
            n += min(10, iters[k] - n)
            o.exploitability()
        spent += time.perf_counter() - t0
        covered += 1
        done.append(k)
        if spent > seconds:
            break
    # scale: river and turn samples separately
    riv = [k for k in done if k < 256]
    trn = [k for k in done if k >= 256]
    per = spent / max(covered, 1)
    est = per * len(cfgs)
    return {"value": round(len(cfgs) / est, 4), "unit": "spots/s", "cores": threads, "kind": "port",
            "sample": f"{len(riv)} river + {len(trn)} turn spots of the batch solved by the oracle for the GPU run's "
                      f"iteration counts with a check every 10 iterations ({spent:.1f}s), scaled to all {len(cfgs)} spots"}


if __name__ == "__main__":
    main()
