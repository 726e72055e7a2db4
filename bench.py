"""Benchmark: FastRRT planning on MI355X, 256x256 scenes.

Metric: plans/s at p99 <= 20 ms, with success %.  One "step" is one batch of
synthetic 256x256 scenes planned end to end (set_plan_data goal-connection
precompute, loop() until the goal is reached or the plan fails,
path_optimize rounds, path extraction).

This version times the plain form: one plan per launch sequence through the
FastRRT handle (the reference's loop protocol, one host sync per loop()).

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
    return ap.parse_args()


def plain_plan(P, rrt, frame, sc, args):
    sp = P.SearchParams(frame, P.Waypoint(sc.start[0], sc.start[1], P.angle(sc.start[2])),
                        P.Waypoint(sc.goal[0], sc.goal[1], P.angle(sc.goal[2])), 1.0, sc.min_dist, -1, 40.0, 20.0,
                        P.angle.new_deg(10))
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
    return (t1 - t0) * 1e3, ok and path is not None, it, stats


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


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(plans are independent: more GPUs shard requests, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    from fastrrt_amd import pyfastrrt as P, scenes
    L = P._lib.load()
    n_per_cov = max(1, args.plans // 3)
    all_scenes = scenes.scenes_256_sweep(n_per_cov * (args.steps + args.warmup), base_seed=5000)
    # interleave coverages so every step holds each coverage level
    by_cov = {}
    for sc in all_scenes:
        by_cov.setdefault(sc.coverage, []).append(sc)
    steps = []
    for k in range(args.steps + args.warmup):
        batch = []
        for cov in sorted(by_cov):
            batch += by_cov[cov][k * n_per_cov:(k + 1) * n_per_cov]
        steps.append(batch)
    ego = P.EgoParams(256, 256, 34.641016151377535, 34.641016151377535, 40.0, 5.412658773, scenes.EGO_LOWER_BOUND,
                      scenes.EGO_UPPER_BOUND, -1.0, scenes.SEGMENTATION_CLASS_COST)
    rrt = P.FastRRT(ego)
    frames = {}

    def frame_of(sc):
        f = frames.get(sc.seed)
        if f is None:
            f = P.SearchFrame(sc.frame, sc.class_costs)
            frames[sc.seed] = f
        return f

    for k in range(args.warmup):
        for sc in steps[k]:
            rrt.set_seed(1234, sc.seed)
            plain_plan(P, rrt, frame_of(sc), sc, args)
    L.fastrrt_profile_reset()
    L.fastrrt_profile_enable(True)
    lat, oks, iters = [], [], []
    exp_total = [0, 0]
    t0 = time.perf_counter()
    for k in range(args.warmup, args.warmup + args.steps):
        for sc in steps[k]:
            rrt.set_seed(1234, sc.seed)
            ms, ok, it, st = plain_plan(P, rrt, frame_of(sc), sc, args)
            lat.append(ms)
            oks.append(ok)
            iters.append(it)
            exp_total[0] += st[0]
            exp_total[1] += st[1]
    t1 = time.perf_counter()
    L.fastrrt_profile_enable(False)
    import ctypes
    prof = {}
    for k, name in enumerate(("goal_connection", "expand", "connect", "iteration")):
        n, ms = ctypes.c_long(), ctypes.c_double()
        L.fastrrt_profile_read(k, ctypes.byref(n), ctypes.byref(ms))
        prof[name] = (n.value, ms.value)
    n_plans = len(lat)
    total_s = t1 - t0
    value = n_plans / total_s
    lat = np.array(lat)
    p50, p99 = float(np.percentile(lat, 50)), float(np.percentile(lat, 99))
    # roofline of the dominant kernel (algorithmic bytes per launch / avg launch time)
    W = H = 256
    gc_n, gc_ms = prof["goal_connection"]
    it_n, it_ms = prof["iteration"]
    cand = []
    if gc_n:
        # per-plan setup: 1 B obstacle + 12 B (flag, heading, cost) per cell
        cand.append(("k_setup_goal_connection", gc_ms / gc_n, 13.0 * W * H, gc_n))
    if it_n:
        # one loop iteration: 6 B/cell + 8 B/rollout + 24 B/new node
        b = 6.0 * W * H + (8.0 * exp_total[0] + 24.0 * exp_total[1]) / it_n
        cand.append(("loop_iteration", it_ms / it_n, b, it_n))
    dom = max(cand, key=lambda c: c[1] * c[3])
    achieved = dom[2] / (dom[1] * 1e-3) / 1e9
    roof = {"bound": "hbm", "kernel": dom[0], "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": None, "avg_launch_ms": round(dom[1], 4),
            "launches": dom[3], "algorithmic_bytes_per_launch": dom[2]}
    cpu = None
    if args.cpu_sample > 0:
        sample = [s for s in steps[args.warmup]][:args.cpu_sample]
        v, times = cpu_baseline(sample, args)
        cpu = {"value": round(v, 4), "unit": "plans/s", "cores": 1, "kind": "port",
               "sample": f"{len(sample)} scenes of the 256x256 workload through the oracle (1 thread); "
                         f"{sum(times):.1f} s"}
    out = {
        "metric": "plans/s at p99<=20 ms (served/batched vs plain loop), 256^2 grid, success %",
        "value": round(value, 3), "unit": "plans/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(total_s * 1e3 / args.steps, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32+int32", "data": "synthetic",
        "config": {"workload": "plain: one plan per launch sequence, 256x256, coverage 5/15/30%, "
                               "start (128,128,0), goal x~U[64,192] row 0, heading~U[-pi/4,pi/4], "
                               "min_dist 22x40, P=40, r=20",
                   "mode": "plain", "plans_per_step": len(steps[0]), "max_iter": args.max_iter,
                   "opt_rounds": args.opt_rounds},
        "p50_ms": round(p50, 3), "p99_ms": round(p99, 3), "success": round(float(np.mean(oks)), 4),
        "mean_iterations": round(float(np.mean(iters)), 2),
        "kernel_ms": {k: {"launches": v[0], "avg_ms": round(v[1] / v[0], 4) if v[0] else None}
                      for k, v in prof.items()},
        "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
