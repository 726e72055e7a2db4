"""Served and mixed-size batched planning on one GPU.

served  - a Poisson stream of plan requests of mixed grid size arrives in
          real time; each request is submitted to one Batch as it arrives
          (continuous batching: it joins the running plans at the next
          iteration as slots and the cell budget free up) with its own
          deadline, measured from its arrival.  Reported: completed plans/s,
          p50/p99 latency from arrival, deadline-miss rate, success rate.
batched - a fixed set of plans of uneven grid size submitted together and
          advanced together until every plan is done.
1024    - the 1024x1024 configuration (rectangles plus random-walk strips,
          smart expansion with 32 px density regions, 300 ms per plan), plain
          loop and batched form on the same scenes.

Frames are drawn from a pool of seeded scenes per size (uploaded once, as a
deployment keeps its frames on the device); every request has its own
start/goal-bearing scene and its own Philox stream.

    python -m fastrrt_amd.serve --mode served --duration 10 --rate 2000
    python -m fastrrt_amd.serve --mode batched --plans 512
    python -m fastrrt_amd.serve --mode 1024 --plans 24
"""
import argparse
import json
import math
import time

import numpy as np

from . import scenes
from .batch import Batch, PlanSpec
from .pyfastrrt import SearchFrame

SERVED_SIZES = (128, 192, 256, 384, 512, 768, 1024)  # log-spaced: log-uniform over 128..1024
BATCHED_SIZES = (128, 256, 384, 512)
DEADLINES_MS = (10.0, 50.0, 100.0)


class ScenePool:
    """`per_size` seeded scenes per grid size, coverage ~ U[5%, 30%], start and
    goal at least 60% of the frame height apart, frames on the device."""

    def __init__(self, sizes, per_size, seed, coverage=(0.05, 0.30), min_dist=(22, 40)):
        rng = np.random.default_rng(seed)
        self.by_size = {}
        for size in sizes:
            lst = []
            for k in range(per_size):
                cov = float(rng.uniform(coverage[0], coverage[1]))
                sc = scenes.scene_sized(int(seed * 1000 + size * 10 + k), size, cov, min_dist=min_dist)
                lst.append((sc, SearchFrame(sc.frame, sc.class_costs)))
            self.by_size[size] = lst

    def pick(self, size, rng):
        lst = self.by_size[size]
        return lst[int(rng.integers(0, len(lst)))]


def validate_all(items, dist_tol=20.0):
    """(scene, path[, tag]) of every found plan -> {"paths", "valid",
    "failed"}: each path checked by fastrrt_amd.validate (chain, goal
    tolerance, every edge's footprints or Hermite connection, heading
    continuity); the first failures with the checks they fail, the plan's
    tag, its scene's seed, size and coverage, and the path."""
    from . import validate
    failed, n_ok = [], 0
    for item in items:
        sc, path = item[0], item[1]
        f, _ = validate.check_path(sc.frame, sc.class_costs, path, sc.start, sc.goal, sc.min_dist,
                                   (sc.ego_lb, sc.ego_ub), dist_tol, math.radians(10))
        n_ok += not f
        if f and len(failed) < 8:
            failed.append({"checks": f, "tag": int(item[2]) if len(item) > 2 else None, "scene_seed": int(sc.seed),
                           "size": [int(sc.frame.shape[1]), int(sc.frame.shape[0])],
                           "coverage": float(sc.coverage), "path": np.asarray(path).tolist()})
    return {"paths": len(items), "valid": n_ok, "failed": failed}


def _spec(sc, frame, tag, stream, deadline_ms, max_iter, opt_rounds, seed):
    return PlanSpec(frame, sc.start, sc.goal, min_dist=sc.min_dist, ego_lb=sc.ego_lb, ego_ub=sc.ego_ub,
                    max_iterations=max_iter, optimize_rounds=opt_rounds, deadline_ms=deadline_ms, seed=seed,
                    stream=stream, tag=tag)


def run_served(duration_s=10.0, rate=2000.0, sizes=SERVED_SIZES, deadlines=DEADLINES_MS, seed=11, per_size=8,
               max_plans=256, max_cells=64 << 20, max_iter=200, opt_rounds=20, pool=None):
    rng = np.random.default_rng(seed)
    pool = pool or ScenePool(sizes, per_size, seed)
    b = Batch(max_plans=max_plans, max_cells=max_cells, launch="auto")
    # arrival schedule (ms from the start of the stream)
    arrivals, t = [], 0.0
    while True:
        t += float(rng.exponential(1000.0 / rate))
        if t >= duration_s * 1000.0:
            break
        arrivals.append(t)
    reqs = []
    for i, a in enumerate(arrivals):
        size = sizes[int(rng.integers(0, len(sizes)))]
        sc, frame = pool.pick(size, rng)
        reqs.append((a, size, sc, frame, float(deadlines[int(rng.integers(0, len(deadlines)))])))
    results = {}
    base = b.clock_ms()
    i = 0
    n = len(reqs)
    while i < n or len(results) < i:
        now = b.clock_ms() - base
        while i < n and reqs[i][0] <= now:
            a, size, sc, frame, dl = reqs[i]
            left = dl - (now - a)
            b.submit(_spec(sc, frame, i, 500000 + i, max(0.05, left), max_iter, opt_rounds, seed))
            i += 1
        if b.step(1) == 0 and i < n:
            wait = reqs[i][0] - (b.clock_ms() - base)
            if wait > 0.05:
                time.sleep(min(wait, 1.0) / 1000.0)
        for r in b.poll():
            results[r.tag] = r
    wall = (b.clock_ms() - base) / 1000.0
    lat, ok, miss, by_size = [], [], [], {}
    for tag, r in results.items():
        a, size, sc, frame, dl = reqs[tag]
        latency = r.done_ms - (base + a)
        lat.append(latency)
        ok.append(r.ok)
        late = latency > dl
        miss.append(late or (not r.ok and r.iterations < max_iter))
        d = by_size.setdefault(size, [0, 0, 0])
        d[0] += 1
        d[1] += int(r.ok)
        d[2] += int(miss[-1])
    lat = np.array(lat)
    valid = validate_all([(reqs[t][2], r.path, t) for t, r in results.items() if r.ok])
    return {"mode": "served", "requests": n, "returned": len(results), "validation": valid,
            "duration_s": duration_s, "offered_rate": rate,
            "completed_per_s": n / wall, "found_per_s": float(np.sum(ok)) / wall,
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "deadline_miss_rate": float(np.mean(miss)), "success_rate": float(np.mean(ok)),
            "late_rate": float(np.mean(lat > np.array([reqs[t][4] for t in results]))),
            "by_size": {str(k): {"requests": v[0], "success": v[1] / v[0], "deadline_miss": v[2] / v[0]}
                        for k, v in sorted(by_size.items())},
            "sizes": list(sizes), "deadlines_ms": list(deadlines), "max_plans": max_plans, "max_cells": max_cells,
            "hbm": b.memory()}


def run_batched(n_plans=512, sizes=BATCHED_SIZES, seed=12, max_iter=200, opt_rounds=20, deadline_ms=0.0,
                max_cells=None, coverage=(0.05, 0.30), min_dist=(22, 40)):
    rng = np.random.default_rng(seed)
    plans = []
    for i in range(n_plans):
        size = sizes[int(rng.integers(0, len(sizes)))]
        cov = float(rng.uniform(coverage[0], coverage[1]))
        sc = scenes.scene_sized(int(seed * 100000 + i), size, cov, min_dist=min_dist)
        plans.append((size, sc, SearchFrame(sc.frame, sc.class_costs)))
    cells = sum(s * s for s, _, _ in plans)
    b = Batch(max_plans=n_plans, max_cells=max_cells or cells, launch="auto")
    t0 = time.perf_counter()
    for i, (size, sc, frame) in enumerate(plans):
        b.submit(_spec(sc, frame, i, 700000 + i, deadline_ms, max_iter, opt_rounds, seed))
    res = b.drain()
    wall = time.perf_counter() - t0
    lat = np.array([r.latency_ms for r in res])
    by_size = {}
    for r in res:
        d = by_size.setdefault(plans[r.tag][0], [0, 0, 0])
        d[0] += 1
        d[1] += int(r.ok)
        d[2] += r.iterations
    valid = validate_all([(plans[r.tag][1], r.path, r.tag) for r in res if r.ok])
    return {"mode": "batched", "plans": n_plans, "returned": len(res), "validation": valid,
            "statuses": sorted(set(int(r.status) for r in res)),
            "cells": cells, "wall_ms": wall * 1e3, "plans_per_s": n_plans / wall,
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "success_rate": float(np.mean([r.ok for r in res])),
            "mean_iterations": float(np.mean([r.iterations for r in res])),
            "by_size": {str(k): {"plans": v[0], "success": v[1] / v[0], "mean_iterations": v[2] / v[0]}
                        for k, v in sorted(by_size.items())},
            "max_iter": max_iter, "deadline_ms": deadline_ms}


def run_1024(n_plans=24, seed=13, timeout_ms=300.0, smart=True, region_side=32, opt_rounds=20, coverage=0.20,
             forms=("plain", "batched")):
    """The 1024x1024 configuration: `n_plans` scenes (20 % rectangles plus
    strips), smart expansion with 32 px density regions, a 300 ms budget per
    plan.  plain: one plan at a time through the FastRRT handle (timeout
    = the budget, the reference's protocol); batched: all plans submitted
    together with deadline = the budget."""
    from . import pyfastrrt as P
    scs = [scenes.scene_1024(seed * 1000 + i, coverage) for i in range(n_plans)]
    frames = [SearchFrame(sc.frame, sc.class_costs) for sc in scs]
    out = {"mode": "1024", "plans": n_plans, "budget_ms": timeout_ms, "smart": smart, "region_side": region_side,
           "coverage": coverage}
    if "plain" in forms:
        sc0 = scs[0]
        ego = P.EgoParams(1024, 1024, sc0.perception_m, sc0.perception_m, 40.0, 5.412658773, sc0.ego_lb, sc0.ego_ub,
                          -1.0, sc0.class_costs)
        rrt = P.FastRRT(ego)
        rrt.set_density_region(region_side)
        lat, ok, iters, loop_ms, paths = [], [], [], 0.0, []
        t_all = time.perf_counter()
        for i, sc in enumerate(scs):
            sp = P.SearchParams(frames[i], P.Waypoint(sc.start[0], sc.start[1], P.angle(sc.start[2])),
                                P.Waypoint(sc.goal[0], sc.goal[1], P.angle(sc.goal[2])), 1.0, sc.min_dist,
                                int(timeout_ms), 40.0, 20.0, P.angle.new_deg(10))
            rrt.set_seed(seed, 900000 + i)
            t0 = time.perf_counter()
            rrt.set_plan_data(sp)
            rrt.search_init()
            t1 = time.perf_counter()
            it = 0
            while not rrt.goal_reached() and rrt.loop(smart):
                it += 1
            t2 = time.perf_counter()
            good = rrt.goal_reached()
            if good:
                r = 0
                while r < opt_rounds and rrt.path_optimize():
                    r += 1
            path = rrt.get_planned_path() if good else None
            t3 = time.perf_counter()
            lat.append((t3 - t0) * 1e3)
            ok.append(bool(good and path))
            if good and path:
                paths.append((sc, np.array([[w.x, w.z, w.heading.rad()] for w in path], np.float32)))
            iters.append(it)
            loop_ms += (t2 - t1) * 1e3
        wall = time.perf_counter() - t_all
        out["plain"] = {"plans_per_s": n_plans / wall, "p50_ms": float(np.percentile(lat, 50)),
                        "p99_ms": float(np.percentile(lat, 99)), "success": float(np.mean(ok)),
                        "mean_iterations": float(np.mean(iters)),
                        "iteration_ms": loop_ms / max(1, sum(iters)),
                        "validation": validate_all(paths)}
    if "batched" in forms:
        b = Batch(max_plans=n_plans, max_cells=n_plans * 1024 * 1024, launch="auto")
        res = None
        for rep in range(2):  # the first pass allocates the slabs
            t0 = time.perf_counter()
            for i, sc in enumerate(scs):
                b.submit(PlanSpec(frames[i], sc.start, sc.goal, min_dist=sc.min_dist, ego_lb=sc.ego_lb,
                                  ego_ub=sc.ego_ub, max_iterations=100000, optimize_rounds=opt_rounds,
                                  deadline_ms=timeout_ms, seed=seed, stream=900000 + i, tag=i, smart=smart,
                                  region_side=region_side))
            res = b.drain()
            wall = time.perf_counter() - t0
        lat = np.array([r.latency_ms for r in res])
        out["batched"] = {"plans_per_s": n_plans / wall, "p50_ms": float(np.percentile(lat, 50)),
                          "p99_ms": float(np.percentile(lat, 99)), "success": float(np.mean([r.ok for r in res])),
                          "mean_iterations": float(np.mean([r.iterations for r in res])),
                          "validation": validate_all([(scs[r.tag], r.path, r.tag) for r in res if r.ok])}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mode", choices=("served", "batched", "1024"), default="served")
    ap.add_argument("--duration", type=float, default=10.0)
    ap.add_argument("--rate", type=float, default=2000.0)
    ap.add_argument("--plans", type=int, default=512)
    ap.add_argument("--max-iter", type=int, default=200)
    ap.add_argument("--deadline-ms", type=float, default=0.0, help="batched mode: per-plan deadline (0: none)")
    ap.add_argument("--per-size", type=int, default=8)
    ap.add_argument("--max-plans", type=int, default=256)
    ap.add_argument("--sizes", type=str, default="")
    args = ap.parse_args()
    if args.mode == "served":
        sizes = tuple(int(s) for s in args.sizes.split(",")) if args.sizes else SERVED_SIZES
        out = run_served(args.duration, args.rate, sizes=sizes, per_size=args.per_size, max_plans=args.max_plans,
                         max_iter=args.max_iter)
    elif args.mode == "batched":
        sizes = tuple(int(s) for s in args.sizes.split(",")) if args.sizes else BATCHED_SIZES
        out = run_batched(args.plans, sizes=sizes, max_iter=args.max_iter, deadline_ms=args.deadline_ms)
    else:
        out = run_1024(args.plans)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
