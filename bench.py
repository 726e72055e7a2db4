"""Benchmark: FastRRT planning on MI355X, 256x256 scenes.

Metric: plans/s at p99 <= 20 ms, with success %.  One "step" is one batch of
synthetic 256x256 scenes planned end to end (set_plan_data goal-connection
precompute, loop() until the goal is reached or the plan fails,
path_optimize rounds, path extraction).

Two forms are timed on the same scenes and the same per-plan budget:
  plain   - one plan per launch sequence through the FastRRT handle (the
            reference's loop protocol: one host sync per loop() call,
            timeout_ms = the budget);
  batched - every plan of a step submitted to one Batch and advanced together
            (one hipGraph replay per loop iteration of all plans), each with
            deadline_ms = the budget from submission.
`value` is the batched plans/s; the plain figure and the ratio are reported
beside it, with p50/p99 latency and success rates of both.

Prints one JSON line.  Single GPU only: --gpus N > 1 exits with a message.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--plans", type=int, default=24, help="scenes per step")
    ap.add_argument("--max-iter", type=int, default=200)
    ap.add_argument("--opt-rounds", type=int, default=20)
    ap.add_argument("--cpu-sample", type=int, default=2, help="scenes timed on the CPU oracle (0 = skip)")
    ap.add_argument("--budget-ms", type=float, default=20.0, help="per-plan timeout / deadline")
    ap.add_argument("--batch", type=int, default=128, help="plans per batched step")
    ap.add_argument("--plain-plans", type=int, default=24, help="plans timed in the plain form")
    ap.add_argument("--iters-per-poll", type=int, default=1)
    return ap.parse_args()


def plain_plan(P, rrt, frame, sc, args):
    """set_plan_data -> search_init -> loop() until goal/timeout -> path_optimize
    rounds -> get_planned_path, through the FastRRT handle."""
    budget = int(round(args.budget_ms)) if args.budget_ms > 0 else -1
    sp = P.SearchParams(frame, P.Waypoint(sc.start[0], sc.start[1], P.angle(sc.start[2])),
                        P.Waypoint(sc.goal[0], sc.goal[1], P.angle(sc.goal[2])), 1.0, sc.min_dist, budget, 40.0,
                        20.0, P.angle.new_deg(10))
    t0 = time.perf_counter()
    rrt.set_plan_data(sp)
    rrt.search_init()
    it = 0
    stats = [0, 0]
    while not rrt.goal_reached() and rrt.loop(False):
        it += 1
        e, n, _c = rrt.iteration_stats()
        stats[0] += e
        stats[1] += n
        if it >= args.max_iter:
            break
    ok = rrt.goal_reached()
    if ok:
        r = 0
        while r < args.opt_rounds and rrt.path_optimize():
            r += 1
    path = rrt.get_planned_path() if ok else None
    t1 = time.perf_counter()
    return (t1 - t0) * 1e3, bool(ok and path), it, stats


def cpu_baseline(scs, args):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from _parity import OraclePlanner, Params
    prm = Params()
    times = []
    for sc in scs:
        t0 = time.perf_counter()
        o = OraclePlanner(sc, prm, 1234, sc.seed)
        o.search_init()
        it = 0
        while not o.goal_reached() and o.loop():
            it += 1
            if it >= args.max_iter:
                break
        if o.goal_reached():
            r = 0
            while r < args.opt_rounds and o.path_optimize():
                r += 1
            o.path()
        times.append(time.perf_counter() - t0)
    return len(scs) / sum(times), times


def read_prof(L):
    import ctypes
    prof = {}
    for k, name in enumerate(("goal_connection", "expand", "connect", "iteration")):
        n, ms = ctypes.c_long(), ctypes.c_double()
        L.fastrrt_profile_read(k, ctypes.byref(n), ctypes.byref(ms))
        prof[name] = (n.value, ms.value)
    return prof


def run_plain(P, L, scs, args, frame_of):
    ego = P.EgoParams(256, 256, 34.641016151377535, 34.641016151377535, 40.0, 5.412658773,
                      scs[0].ego_lb, scs[0].ego_ub, -1.0, scs[0].class_costs)
    rrt = P.FastRRT(ego)
    for sc in scs[:2]:  # warm-up
        rrt.set_seed(1234, sc.seed)
        plain_plan(P, rrt, frame_of(sc), sc, args)
    lat, oks, iters = [], [], []
    t0 = time.perf_counter()
    for sc in scs:
        rrt.set_seed(1234, sc.seed)
        ms, ok, it, _st = plain_plan(P, rrt, frame_of(sc), sc, args)
        lat.append(ms)
        oks.append(ok)
        iters.append(it)
    t1 = time.perf_counter()
    return {"plans_per_s": len(scs) / (t1 - t0), "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)), "success": float(np.mean(oks)),
            "mean_iterations": float(np.mean(iters)), "plans": len(scs)}


def run_batched(B, scs, args, frame_of, batch=None):
    from fastrrt_amd.batch import Batch, PlanSpec
    n = len(scs)
    b = batch or Batch(max_plans=max(n, 1), max_cells=max(n, 1) * 256 * 256)
    t0 = time.perf_counter()
    c0 = b.clock_ms()
    for i, sc in enumerate(scs):
        b.submit(PlanSpec(frame_of(sc), sc.start, sc.goal, min_dist=sc.min_dist, ego_lb=sc.ego_lb, ego_ub=sc.ego_ub,
                          max_iterations=args.max_iter, optimize_rounds=args.opt_rounds,
                          deadline_ms=args.budget_ms, seed=1234, stream=sc.seed, tag=i))
    res = []
    while b.step(args.iters_per_poll) > 0:
        res += b.poll()
    res += b.poll()
    t1 = time.perf_counter()
    lat = np.array([r.latency_ms for r in res])
    return b, {"plans_per_s": n / (t1 - t0), "p50_ms": float(np.percentile(lat, 50)),
               "p99_ms": float(np.percentile(lat, 99)), "success": float(np.mean([r.ok for r in res])),
               "mean_iterations": float(np.mean([r.iterations for r in res])), "plans": n,
               "wall_ms": (t1 - t0) * 1e3}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(plans are independent: more GPUs shard requests, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    from fastrrt_amd import pyfastrrt as P, scenes
    L = P._lib.load()
    per_cov = max(1, args.batch // 3)
    pool = scenes.scenes_256_sweep(per_cov * (args.steps + args.warmup), base_seed=5000)
    by_cov = {}
    for sc in pool:
        by_cov.setdefault(sc.coverage, []).append(sc)
    steps = []
    for k in range(args.steps + args.warmup):
        step = []
        for cov in sorted(by_cov):
            step += by_cov[cov][k * per_cov:(k + 1) * per_cov]
        steps.append(step)
    frames = {}

    def frame_of(sc):
        f = frames.get(sc.seed)
        if f is None:
            f = P.SearchFrame(sc.frame, sc.class_costs)
            frames[sc.seed] = f
        return f

    for st in steps:
        for sc in st:
            frame_of(sc)
    # batched form
    batch = None
    for k in range(args.warmup):
        batch, _ = run_batched(P, steps[k], args, frame_of, batch)
    L.fastrrt_profile_reset()
    L.fastrrt_profile_enable(True)
    agg = []
    t0 = time.perf_counter()
    for k in range(args.warmup, args.warmup + args.steps):
        batch, r = run_batched(P, steps[k], args, frame_of, batch)
        agg.append(r)
    t1 = time.perf_counter()
    L.fastrrt_profile_enable(False)
    prof_b = read_prof(L)
    n_b = sum(r["plans"] for r in agg)
    batched = {"plans_per_s": n_b / (t1 - t0), "p50_ms": float(np.median([r["p50_ms"] for r in agg])),
               "p99_ms": float(max(r["p99_ms"] for r in agg)),
               "success": float(np.average([r["success"] for r in agg], weights=[r["plans"] for r in agg])),
               "mean_iterations": float(np.mean([r["mean_iterations"] for r in agg])), "plans": n_b}
    # plain form on a sample of the same scenes
    plain_scs = [sc for st in steps[args.warmup:] for sc in st][:args.plain_plans]
    L.fastrrt_profile_reset()
    L.fastrrt_profile_enable(True)
    plain = run_plain(P, L, plain_scs, args, frame_of) if args.plain_plans > 0 else None
    L.fastrrt_profile_enable(False)
    prof_p = read_prof(L)
    # roofline of the dominant kernel of the batched run: the per-plan
    # goal-connection precompute reads 1 B/cell (obstacle distance plane) and
    # writes 9 B/cell (flag + heading + cost) plus 3 B/cell of prep planes
    W = H = 256
    gc_n, gc_ms = prof_b["goal_connection"]
    it_n, it_ms = prof_b["iteration"]
    plans_per_gc = n_b / max(1, gc_n)
    cand = []
    if gc_n:
        cand.append(("k_setup_goal_connection", gc_ms / gc_n, 13.0 * W * H * plans_per_gc, gc_n))
    if it_n:
        cand.append(("loop_iteration(graph)", it_ms / it_n, 6.0 * W * H * n_b / max(1, args.steps) / 1.0, it_n))
    dom = max(cand, key=lambda c: c[1] * c[3])
    achieved = dom[2] / (dom[1] * 1e-3) / 1e9
    roof = {"bound": "hbm", "kernel": dom[0], "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": None, "avg_launch_ms": round(dom[1], 4),
            "launches": dom[3], "algorithmic_bytes_per_launch": dom[2]}
    cpu = None
    if args.cpu_sample > 0:
        sample = steps[args.warmup][:args.cpu_sample]
        v, times = cpu_baseline(sample, args)
        cpu = {"value": round(v, 4), "unit": "plans/s", "cores": 1, "kind": "port",
               "sample": f"{len(sample)} scenes of the 256x256 workload through the oracle, 1 thread, "
                         f"{sum(times):.1f} s"}
    out = {
        "metric": "plans/s at p99<=20 ms (served/batched vs plain loop), 256^2 grid, success %",
        "value": round(batched["plans_per_s"], 3), "unit": "plans/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round((t1 - t0) * 1e3 / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32+fp64+int32", "data": "synthetic",
        "config": {"workload": "256x256 frames, rectangles 4-30 px at 5/15/30% coverage, start (128,128,0), "
                               "goal x~U[64,192] on row 0, heading~U[-pi/4,pi/4]; min_dist 22x40, P=40, r=20, "
                               "steer 40 deg, L=5.41; loop to goal, then path_optimize rounds",
                   "mode": "batched", "plans_per_step": len(steps[0]), "budget_ms": args.budget_ms,
                   "max_iter": args.max_iter, "opt_rounds": args.opt_rounds},
        "p50_ms": round(batched["p50_ms"], 3), "p99_ms": round(batched["p99_ms"], 3),
        "success": round(batched["success"], 4),
        "batched": batched, "plain": plain,
        "batched_vs_plain": (round(batched["plans_per_s"] / plain["plans_per_s"], 3) if plain else None),
        "kernel_ms_batched": {k: {"launches": v[0], "avg_ms": round(v[1] / v[0], 4) if v[0] else None}
                              for k, v in prof_b.items()},
        "kernel_ms_plain": {k: {"launches": v[0], "avg_ms": round(v[1] / v[0], 4) if v[0] else None}
                            for k, v in prof_p.items()},
        "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
