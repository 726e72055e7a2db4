"""Benchmark: FastRRT planning on one MI355X, 256x256 scenes.

Metric: plans/s at p99 <= 20 ms (batched vs the plain loop), with success %.

A "step" is one batch of `--batch` plans on synthetic 256x256 scenes
(rectangles 4-30 px at 5/15/30 % coverage, a third each), planned end to
end: goal-connection planes, loop() until the goal is reached or the plan
fails, path_optimize rounds, path extraction.  Every plan has its own Philox
key (seed, stream), so a plan's search is the same in both forms.

  batched (value) - the step's plans submitted together to one Batch and
                    advanced together (one hipGraph replay per loop iteration
                    of every running plan), each with deadline = the budget;
  plain           - the plans of the first `--plain-steps` timed steps one at
                    a time through the FastRRT handle (the reference's loop
                    protocol: one host sync per loop() call), timeout = the
                    budget - 2 ms (set_plan_data and path_optimize fall
                    outside the reference's timeout, fastrrt.cpp:91,
                    :110-114).

Success is compared plan by plan on the plans both forms ran (by default all
20 timed steps, 1920 plans: about 30 of them are solved by one form only, so
the comparison's standard error is about 0.3 pp).  96 plans per step: the
same plans solved in groups of 120, 96, 72 and 48 gave 628, 634, 636 and 643
of 1920 (each smaller group gets more loop iterations per plan before the
deadline and loses none that a larger one solved), and the plain loop solves
about 633-645 of them, so 96 keeps the success gap within about 0.5 pp at
about 50 x 96 plans/s.  Frames are uploaded before the timed region
(device-resident inputs).

roofline: the batched form's unit of work, one launch of the captured
iteration graph (one loop iteration of every running plan: 11 kernels, of
which the two lazy goal-plane kernels and k_expand take about half the time).
The timed region retires plans on wall-clock deadlines, so its work depends
on timing (and a profiler's overhead changes it); the roofline is therefore
measured on a deterministic pass after it: the plans of step 1 of the
sequence (the first timed step at the default single warm-up step, the same
plans whatever `--warmup` is, so the committed profiles cover the very pass a
run reports), batched, `--roofline-iters` iterations, no deadline, no
optimisation rounds.
The graph is timed with HIP events recorded on the engine's stream around
each launch; k_expand on a second, identical pass with the iteration captured
as three graphs and events between them.  `python bench.py --roofline-only` runs that pass alone (the
command the committed rocprofv3 summaries profile).  Algorithmic bytes per
plan-iteration (the path's per-unit figure): 6 B per grid cell (4 B node
type, 1 B obstacle plane, 1 B goal-connection flag) + 8 B per expansion
(source heading and cost) + 24 B per new node, with new nodes counted as
expansions (an upper bound; the term is < 0.2 % of the total at 256^2).
`traffic` is the HBM bytes per iteration measured by rocprofv3 --pmc
FETCH_SIZE / WRITE_SIZE on that command, summed over the iteration's
kernels and scaled by the same counters' reading of a copy kernel of known
size (profiles/iteration_traffic.json).  The k_expand figures of the same
pass are reported beside it.  The path is instruction-bound more than
HBM-bound, so `instruction_side` gives the same pass's work: footprint
samples/s (the reference's footprint test reads every sample of every
footprint it tests: at each cell a rollout enters and at each point of a
goal curve whose footprints are tested), FP64 Hermite points/s of the goal
curves, and the VALU issue rate against MI355X's peak (256 CUs x 4 SIMDs x
2.4 GHz / 4 cycles per wave64 instruction), from the SQ_INSTS_VALU count of
the same command (profiles/iteration_counters.json).

cpu_baseline: the CPU oracle on the host cores, run before the GPU is
touched: one process per core (`--cpu-workers`, at most the cores present)
on as many plans of the first timed step (all coverages), with the
one-thread figure on `--cpu-sample` plans beside it.  The oracle runs the
same loop without a deadline (up to --max-iter iterations).

Prints one JSON line.  One GPU only: --gpus N > 1 exits with a message.
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# VALU issue peak: 256 CUs x 4 SIMDs, one wave64 VALU instruction per SIMD
# every 4 cycles (16 lanes a cycle) at 2.4 GHz
VALU_PEAK_GWI = 256 * 4 * 2.4 / 4.0
CELL_BYTES = 6.0     # per grid cell per plan-iteration: node type 4 B, obstacle 1 B, goal flag 1 B
EXP_BYTES = 8.0      # per expansion: source heading + cost
NEW_BYTES = 24.0     # per new node: parent, heading, cost, type, derive, region
ROLLOUT_BYTES = 76.0 # k_expand per rollout: queue entry 20, source 8, result 12, claim 16, centre pixels 20
SEED = 1234


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=96, help="plans per step")
    ap.add_argument("--budget-ms", type=float, default=20.0, help="per-plan latency budget (batched deadline)")
    ap.add_argument("--plain-timeout-ms", type=int, default=-1,
                    help="FastRRT handle timeout of the plain form (-1: budget - 2 ms, so that its p99 stays "
                         "within the budget once set_plan_data and path_optimize are added)")
    ap.add_argument("--max-iter", type=int, default=200)
    ap.add_argument("--opt-rounds", type=int, default=20)
    ap.add_argument("--plain-plans", type=int, default=-1, help="plans of the first timed step run plain (-1: all)")
    ap.add_argument("--plain-steps", type=int, default=20,
                    help="timed steps whose plans also run plain (success compared plan by plan)")
    ap.add_argument("--cpu-sample", type=int, default=2, help="plans timed on the CPU oracle, one thread (0 = skip)")
    ap.add_argument("--cpu-workers", type=int, default=16, help="processes of the all-core CPU oracle run")
    ap.add_argument("--c1-plans", type=int, default=32,
                    help="C1 plans (10 %% rectangles, goal (128,0,0)) compared between the CPU oracle and the GPU")
    ap.add_argument("--c1-timeout-ms", type=int, default=100, help="the reference's timeout of the C1 comparison")
    ap.add_argument("--plain-profile-plans", type=int, default=48,
                    help="plans of a separate plain pass timed kernel by kernel with HIP events")
    ap.add_argument("--iters-per-poll", type=int, default=1)
    ap.add_argument("--roofline-iters", type=int, default=32)
    ap.add_argument("--roofline-only", action="store_true", help="run only the deterministic roofline pass")
    ap.add_argument("--mode", choices=("batched", "served"), default="batched",
                    help="batched: each step's plans submitted together and drained; served: the steps' plans "
                         "admitted continuously, a new plan whenever --admit-group slots are free")
    ap.add_argument("--admit-group", type=int, default=16, help="served mode: free slots that trigger admission")
    return ap.parse_args()


class Plan:
    __slots__ = ("sc", "stream", "tag")

    def __init__(self, sc, stream, tag):
        self.sc, self.stream, self.tag = sc, stream, tag


def make_steps(args, scenes):
    n_steps = args.steps + args.warmup
    per_cov = max(1, args.batch // 3)
    pool = scenes.scenes_256_sweep(per_cov * n_steps, base_seed=5000)
    by_cov = {}
    for sc in pool:
        by_cov.setdefault(sc.coverage, []).append(sc)
    steps, tag = [], 0
    for k in range(n_steps):
        step = []
        for cov in sorted(by_cov):
            for sc in by_cov[cov][k * per_cov:(k + 1) * per_cov]:
                step.append(Plan(sc, 100000 + tag, tag))
                tag += 1
        steps.append(step)
    return steps


def plain_plan(P, rrt, frame, pl, args, timeout_ms=None):
    """set_plan_data -> search_init -> loop() until goal/timeout -> path_optimize
    rounds -> get_planned_path, through the FastRRT handle.  Returns the
    latency, the path (None when none), the iterations and the phase times."""
    sc = pl.sc
    budget = args.plain_timeout_ms if timeout_ms is None else timeout_ms
    sp = P.SearchParams(frame, P.Waypoint(sc.start[0], sc.start[1], P.angle(sc.start[2])),
                        P.Waypoint(sc.goal[0], sc.goal[1], P.angle(sc.goal[2])), 1.0, sc.min_dist, budget, 40.0,
                        20.0, P.angle.new_deg(10))
    rrt.set_seed(SEED, pl.stream)
    t0 = time.perf_counter()
    rrt.set_plan_data(sp)
    rrt.search_init()
    t_setup = time.perf_counter()
    it = 0
    while not rrt.goal_reached() and rrt.loop(False):
        it += 1
        if it >= args.max_iter:
            break
    ok = rrt.goal_reached()
    t_loop = time.perf_counter()
    if ok:
        r = 0
        while r < args.opt_rounds and rrt.path_optimize():
            r += 1
    t_opt = time.perf_counter()
    path = rrt.get_planned_path() if ok else None
    t1 = time.perf_counter()
    phases = ((t_setup - t0) * 1e3, (t_loop - t_setup) * 1e3, (t_opt - t_loop) * 1e3, (t1 - t_opt) * 1e3)
    arr = np.array([[w.x, w.z, w.heading.rad()] for w in path], np.float32) if (ok and path) else None
    return (t1 - t0) * 1e3, arr, it, phases


def plain_path_cost(L, rrt, path):
    """cost field of the path's last node (read after the plan is timed)"""
    cost = np.zeros((256, 256), np.float32)
    L.fastrrt_export_state(rrt._ptr, None, None, None, None, None, cost.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                           None, None, None)
    return float(cost[int(path[-1, 1]), int(path[-1, 0])])


def run_plain(P, L, plans, args, frame_of, timeout_ms=None):
    """The plans one at a time through one FastRRT handle.  Latency covers
    set_plan_data to the extracted path; the path cost is read after the
    latency clock stops, and its read time is left out of plans/s."""
    sc0 = plans[0].sc
    ego = P.EgoParams(256, 256, sc0.perception_m, sc0.perception_m, 40.0, 5.412658773, sc0.ego_lb, sc0.ego_ub, -1.0,
                      sc0.class_costs)
    rrt = P.FastRRT(ego)
    for pl in plans[:2]:  # warm-up
        plain_plan(P, rrt, frame_of(pl.sc), pl, args, timeout_ms)
    lat, paths, iters, phases, costs = [], {}, [], [], {}
    extra = 0.0
    t0 = time.perf_counter()
    for pl in plans:
        ms, path, it, ph = plain_plan(P, rrt, frame_of(pl.sc), pl, args, timeout_ms)
        lat.append(ms)
        paths[pl.tag] = path
        iters.append(it)
        phases.append(ph)
        if path is not None:
            tc = time.perf_counter()
            costs[pl.tag] = plain_path_cost(L, rrt, path)
            extra += time.perf_counter() - tc
    t1 = time.perf_counter()
    wall = t1 - t0 - extra
    ph = np.array(phases)
    phase_ms = {name: [round(float(np.percentile(ph[:, k], q)), 3) for q in (50, 99, 100)]
                for k, name in enumerate(("setup", "loop", "optimize", "path"))}
    found = sum(1 for v in paths.values() if v is not None)
    out = {"plans_per_s": len(plans) / wall, "found_plans_per_s": found / wall,
           "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
           "max_ms": float(np.max(lat)), "success": found / len(plans),
           "mean_iterations": float(np.mean(iters)), "plans": len(plans),
           "mean_path_cost": float(np.mean(list(costs.values()))) if costs else None,
           "phase_p50_p99_max_ms": phase_ms}
    return out, paths, costs


def validate_paths(items):
    """items: (tag, scene, path) of every found plan -> counts of valid plans
    and of each failed check (fastrrt_amd.validate), with the first failing
    plans' tags."""
    from fastrrt_amd import validate
    fails, n_ok, worst, short_last, bad = {}, 0, 0.0, 0, []
    for tag, sc, path in items:
        short_last += math.hypot(float(path[-1, 0] - path[-2, 0]), float(path[-1, 1] - path[-2, 1])) < 2.0
        f, rate = validate.check_path(sc.frame, sc.class_costs, path, sc.start, sc.goal, sc.min_dist,
                                      (sc.ego_lb, sc.ego_ub), 20.0, math.radians(10))
        n_ok += not f
        if f and len(bad) < 8:
            bad.append([int(tag)] + f)
        for k in f:
            fails[k] = fails.get(k, 0) + 1
        if rate.size:
            worst = max(worst, float(rate[:-1].max()) if rate.size > 1 else 0.0)
    return {"paths": len(items), "valid": n_ok, "failed_checks": fails,
            "max_inner_turning_rate_per_px": round(worst, 4), "last_edges_under_2px": short_last,
            "failed_plans": bad}


def plan_spec(pl, args, frame_of, max_iter=None, opt_rounds=None, deadline=None):
    from fastrrt_amd.batch import PlanSpec
    sc = pl.sc
    return PlanSpec(frame_of(sc), sc.start, sc.goal, min_dist=sc.min_dist, ego_lb=sc.ego_lb, ego_ub=sc.ego_ub,
                    max_iterations=args.max_iter if max_iter is None else max_iter,
                    optimize_rounds=args.opt_rounds if opt_rounds is None else opt_rounds,
                    deadline_ms=args.budget_ms if deadline is None else deadline, seed=SEED,
                    stream=pl.stream, tag=pl.tag)


def run_batched(b, plans, args, frame_of, max_iter=None, opt_rounds=None, deadline=None):
    """The requests are built, then submitted together (each plan's deadline
    counts from its own submission); the wall time covers both."""
    t0 = time.perf_counter()
    specs = [plan_spec(pl, args, frame_of, max_iter, opt_rounds, deadline) for pl in plans]
    for spec in specs:
        b.submit(spec)
    res = []
    while b.step(args.iters_per_poll) > 0:
        res += b.poll()
    res += b.poll()
    t1 = time.perf_counter()
    return res, (t1 - t0) * 1e3


def run_served(b, plans, args, frame_of, capacity):
    """Continuous admission: at most `capacity` plans in flight; whenever at
    least --admit-group slots are free (or the queue's last plans remain),
    the next plans are submitted.  Each plan's deadline counts from its own
    submission."""
    t0 = time.perf_counter()
    queue = [plan_spec(pl, args, frame_of) for pl in plans]
    queue.reverse()
    res, inflight = [], 0
    group = max(1, min(args.admit_group, capacity))
    while queue or inflight:
        if queue and (capacity - inflight >= group or len(queue) <= capacity - inflight):
            while queue and inflight < capacity:
                b.submit(queue.pop())
                inflight += 1
        b.step(args.iters_per_poll)
        got = b.poll()
        res += got
        inflight -= len(got)
    t1 = time.perf_counter()
    return res, (t1 - t0) * 1e3


def roofline_pass(L, batch, plans, args, frame_of):
    """The iteration graph (and k_expand) over a deterministic batched pass,
    run twice: the graph timed whole as it runs unprofiled, then split in
    three to time k_expand alone (the same work: the results must match)."""
    def timed_pass(split):
        L.fastrrt_profile_reset()
        L.fastrrt_profile_split(split)
        L.fastrrt_profile_enable(True)
        r, _ = run_batched(batch, plans, args, frame_of, max_iter=args.roofline_iters, opt_rounds=0, deadline=0.0)
        L.fastrrt_profile_enable(False)
        L.fastrrt_profile_split(False)
        return r, read_prof(L)

    res, prof = timed_pass(False)
    res2, prof2 = timed_pass(True)
    key = lambda rs: sorted((r.tag, r.status, r.iterations, r.expansions) for r in rs)
    if key(res) != key(res2):
        raise RuntimeError("the roofline pass gave different results in its two runs")
    it_n, it_ms = prof["iteration"]
    ex_n, ex_ms = prof2["expand"]
    if not it_n or not ex_n:
        return None
    cells = sum(pl.sc.width * pl.sc.height for pl in plans) / len(plans)
    plan_iters = sum(r.iterations for r in res)
    rollouts = sum(r.expansions for r in res)
    per_launch = (CELL_BYTES * cells * plan_iters + (EXP_BYTES + NEW_BYTES) * rollouts) / it_n
    avg_ms = it_ms / it_n
    achieved = per_launch / (avg_ms * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "iteration_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            tr = json.load(fh)
        if tr.get("batch") == len(plans) and tr.get("roofline_iters") == args.roofline_iters:
            traffic = tr["hbm_bytes_per_iteration"]
    ex_avg = ex_ms / ex_n
    ex_bytes = ROLLOUT_BYTES * rollouts / ex_n
    # instruction side: the reference's footprint test reads every one of its
    # (2 mz + 1)(2 mx + 1) samples, at each cell a rollout enters and at each
    # point of a goal curve whose footprints are tested; FP64 Hermite points
    # of the goal curves (centre pass and footprint pass)
    md = plans[0].sc.min_dist
    F = (2 * (md[1] // 2) + 1) * (2 * (md[0] // 2) + 1)
    cells = sum(r.work[0] for r in res)
    gpts = sum(r.work[1] for r in res)
    fpts = sum(r.work[2] for r in res)
    t_s = it_n * avg_ms * 1e-3
    instr = {"footprint_samples_per_s": round((cells + fpts) * F / t_s, 1),
             "footprint_tests_per_launch": round((cells + fpts) / it_n, 1), "samples_per_footprint": F,
             "hermite_points_per_s": round((gpts + fpts) / t_s, 1),
             "hermite_points_per_launch": round((gpts + fpts) / it_n, 1)}
    cpath = os.path.join(ROOT, "profiles", "iteration_counters.json")
    if os.path.exists(cpath):
        with open(cpath) as fh:
            cn = json.load(fh)
        if cn.get("batch") == len(plans) and cn.get("roofline_iters") == args.roofline_iters:
            wi = cn["valu_instructions_per_iteration"]
            ach = wi / (avg_ms * 1e-3) / 1e9
            instr["valu"] = {"bound": "valu", "achieved": round(ach, 2), "peak": VALU_PEAK_GWI,
                             "unit": "G wave64 VALU instructions/s", "frac": round(ach / VALU_PEAK_GWI, 4),
                             "instructions_per_launch": wi,
                             "busiest_kernel": cn.get("busiest_kernel"),
                             "source": "profiles/iteration_counters.json (SQ_INSTS_VALU of the same pass)"}
    return {"bound": "hbm", "kernel": "iteration graph (11 kernels, one launch per loop iteration of every plan)",
            "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": traffic,
            "avg_launch_ms": round(avg_ms, 5), "launches": it_n, "plans": len(plans),
            "plan_iterations_per_launch": round(plan_iters / it_n, 1),
            "rollouts_per_launch": round(rollouts / it_n, 1),
            "algorithmic_bytes_per_launch": round(per_launch, 1),
            "k_expand": {"avg_launch_ms": round(ex_avg, 5), "launches": ex_n,
                         "algorithmic_bytes_per_launch": round(ex_bytes, 1),
                         "achieved_gbs": round(ex_bytes / (ex_avg * 1e-3) / 1e9, 4)},
            "instruction_side": instr,
            "pass": f"{len(plans)} plans of step 1 of the sequence, batched, {args.roofline_iters} iterations, "
                    "no deadline, no optimisation rounds"}


def read_prof(L):
    """cumulative (launches, total ms) per timed kernel"""
    prof = {}
    for k, name in enumerate(("goal_connection", "expand", "connect", "iteration")):
        n, ms = ctypes.c_long(), ctypes.c_double()
        L.fastrrt_profile_read(k, ctypes.byref(n), ctypes.byref(ms))
        prof[name] = (n.value, ms.value)
    return prof


def prof_diff(a, b):
    return {k: (b[k][0] - a[k][0], b[k][1] - a[k][1]) for k in b}


def prof_json(p):
    return {k: {"launches": v[0], "avg_ms": round(v[1] / v[0], 5) if v[0] else None} for k, v in p.items()}


def _cpu_plan(item):
    """One plan through the CPU oracle, the same loop as the GPU forms
    (no deadline: up to max_iter iterations, then the optimisation rounds);
    returns its seconds."""
    sc, stream, max_iter, opt_rounds = item
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from _parity import OraclePlanner, Params
    t0 = time.perf_counter()
    o = OraclePlanner(sc, Params(), SEED, stream)
    o.search_init()
    it = 0
    while not o.goal_reached() and o.loop():
        it += 1
        if it >= max_iter:
            break
    if o.goal_reached():
        r = 0
        while r < opt_rounds and o.path_optimize():
            r += 1
        o.path()
    return time.perf_counter() - t0


def c1_scenes(n):
    """C1: 256x256, 10 % rectangles, start (128,128,0), goal (128,0,0)."""
    from fastrrt_amd import scenes
    return [Plan(scenes.scene_256(20000 + k, 0.10, 128, 0.0), 300000 + k, 900000 + k) for k in range(n)]


def _cpu_c1_plan(item):
    """One C1 plan through the CPU oracle with the reference's protocol:
    set_plan_data, search_init (the timeout clock starts, fastrrt.cpp:91),
    loop() until the goal or the timeout, path_optimize rounds until the
    timeout or no gain.  Returns (found, path cost, iterations, seconds)."""
    sc, stream, timeout_ms, opt_rounds = item
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from _parity import OraclePlanner, Params
    t0 = time.perf_counter()
    o = OraclePlanner(sc, Params(), SEED, stream, timeout_ms=timeout_ms)
    o.search_init()
    it = 0
    while not o.goal_reached() and o.loop():
        it += 1
    cost = None
    if o.goal_reached():
        r = 0
        while r < opt_rounds and o.path_optimize():
            r += 1
        path = o.path()
        if path.shape[0]:
            x, z = int(path[-1, 0]), int(path[-1, 1])
            cost = float(o.L.og_get_cost(o.L.og_rrt_graph(o.r), x, z))
    return cost is not None, cost, it, time.perf_counter() - t0


def cpu_baseline(steps, args):
    """The CPU oracle on the host cores, run before the GPU is touched: one
    thread on `--cpu-sample` plans, then one process per core (at most
    `--cpu-workers`) on as many plans, spread over the three coverages."""
    first = steps[args.warmup]
    single = first[:args.cpu_sample]
    t_single = sum(_cpu_plan((pl.sc, pl.stream, args.max_iter, args.opt_rounds)) for pl in single)
    workers = max(1, min(args.cpu_workers, os.cpu_count() or 1))
    stride = max(1, len(first) // workers)
    many = first[::stride][:workers]
    import multiprocessing as mp
    t0 = time.perf_counter()
    with mp.get_context("fork").Pool(workers) as pool:
        pool.map(_cpu_plan, [(pl.sc, pl.stream, args.max_iter, args.opt_rounds) for pl in many])
    wall = time.perf_counter() - t0
    c1 = None
    if args.c1_plans > 0:
        plans = c1_scenes(args.c1_plans)
        t0 = time.perf_counter()
        got = []
        with mp.get_context("fork").Pool(workers) as pool:
            items = [(pl.sc, pl.stream, args.c1_timeout_ms, args.opt_rounds) for pl in plans]
            for g in pool.imap(_cpu_c1_plan, items):
                got.append(g)
                if len(got) % 64 == 0:  # progress on stderr for long comparisons
                    print(f"[bench] C1 on the CPU oracle: {len(got)}/{len(items)} plans", file=sys.stderr, flush=True)
        c1_wall = time.perf_counter() - t0
        costs = [g[1] for g in got if g[0]]
        c1 = {"plans": len(plans), "success": round(float(np.mean([g[0] for g in got])), 4),
              "mean_path_cost": round(float(np.mean(costs)), 2) if costs else None,
              "mean_iterations": round(float(np.mean([g[2] for g in got])), 1),
              "per_plan": {pl.tag: (g[0], g[1]) for pl, g in zip(plans, got)},
              "cores": workers, "wall_s": round(c1_wall, 1)}
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    return {"value": round(len(many) / wall, 4), "unit": "plans/s", "cores": workers, "kind": "port",
            "host": {"cpu_model": model, "nproc": os.cpu_count()},
            "sample": f"{len(many)} plans of the timed workload (one in {stride} of the first timed step, all "
                      f"coverages) through the CPU oracle, one process per core, {wall:.1f} s wall",
            "single_thread": {"value": round(len(single) / t_single, 4), "cores": 1,
                              "sample": f"the first {len(single)} plans (5% coverage), {t_single:.1f} s"},
            "c1": c1}


def c1_compare(P, L, Batch, cpu_c1, args, frame_of):
    """C1 on the GPU beside the CPU oracle's run: the same plans (same
    Philox keys, so the same searches) through the FastRRT handle with the
    reference's timeout, and batched (`--batch` plans at a time) with that
    timeout as each plan's deadline; like the CPU run (and the reference's
    test, test_rrt.cpp:170-203), both loop until the goal or the timeout,
    with no iteration cap.  A plan's search is the same on both sides, so a
    plan the CPU solves within its timeout is one the GPU solves too unless
    the GPU ran out of time first.  The stated bounds: GPU success >= CPU
    success on the same plans, and on the plans both solve a mean path cost
    within 2 % of the CPU's (the FastRRT handle's carried cost plane can
    change a path); the plans solved only by the CPU are reported."""
    import copy
    plans = c1_scenes(cpu_c1["plans"])
    uncapped = copy.copy(args)
    uncapped.max_iter = 1 << 30
    plain, paths, costs = run_plain(P, L, plans, uncapped, frame_of, timeout_ms=args.c1_timeout_ms)
    n = min(args.batch, len(plans))
    b = Batch(max_plans=n, max_cells=n * 256 * 256, launch="auto")
    res = []
    for k in range(0, len(plans), n):
        got, _ = run_batched(b, plans[k:k + n], args, frame_of, max_iter=0, deadline=float(args.c1_timeout_ms))
        res += got
    ok_b = {r.tag: r.ok for r in res}
    cost_b = [r.path_cost for r in res if r.ok]
    cpu_ok = {int(t): v[0] for t, v in cpu_c1["per_plan"].items()}
    cpu_cost = {int(t): v[1] for t, v in cpu_c1["per_plan"].items() if v[0]}
    only_cpu_plain = sum(1 for pl in plans if cpu_ok[pl.tag] and paths.get(pl.tag) is None)
    only_cpu_batched = sum(1 for pl in plans if cpu_ok[pl.tag] and not ok_b.get(pl.tag, False))
    bcost = {r.tag: r.path_cost for r in res if r.ok}

    def same(gpu_cost):
        """mean path cost of the CPU and of a GPU form on the plans both solved"""
        both = [t for t in cpu_cost if t in gpu_cost]
        if not both:
            return None
        return {"plans": len(both), "cpu": round(float(np.mean([cpu_cost[t] for t in both])), 2),
                "gpu": round(float(np.mean([gpu_cost[t] for t in both])), 2)}
    return {"workload": "256x256, 10% rectangles 4-30 px, start (128,128,0), goal (128,0,0), min_dist 22x40, "
                        f"P=40, r=20, timeout {args.c1_timeout_ms} ms (loop and path_optimize), then path_optimize",
            "plans": len(plans),
            "cpu": {k: cpu_c1[k] for k in ("success", "mean_path_cost", "mean_iterations", "cores")},
            "gpu_plain": {"success": round(plain["success"], 4), "mean_path_cost": plain["mean_path_cost"],
                          "mean_iterations": round(plain["mean_iterations"], 1), "p99_ms": round(plain["p99_ms"], 2),
                          "solved_by_cpu_not_gpu": only_cpu_plain, "cost_on_plans_both_solved": same(costs)},
            "gpu_batched": {"success": round(float(np.mean(list(ok_b.values()))), 4),
                            "mean_path_cost": float(np.mean(cost_b)) if cost_b else None,
                            "mean_iterations": round(float(np.mean([r.iterations for r in res])), 1),
                            "plans_per_batch": n, "solved_by_cpu_not_gpu": only_cpu_batched,
                            "cost_on_plans_both_solved": same(bcost)},
            "bound": "GPU success >= CPU success on the same plans; on the plans both solve, mean path cost "
                     "within 2 % of the CPU's (plans solved by the CPU only are reported: a plan that needs more "
                     "iterations than a form gets within the timeout)"}


def main():
    args = parse()
    if args.plain_timeout_ms < 0:
        args.plain_timeout_ms = max(1, int(args.budget_ms) - 2)
    if args.gpus != 1:
        print(f"bench.py: measures one GPU; --gpus {args.gpus} is not supported "
              "(plans are independent: more GPUs would shard requests)", file=sys.stderr)
        sys.exit(2)
    from fastrrt_amd import scenes
    steps = make_steps(args, scenes)
    cpu = cpu_baseline(steps, args) if args.cpu_sample > 0 and not args.roofline_only else None
    from fastrrt_amd import pyfastrrt as P
    from fastrrt_amd.batch import Batch
    L = P._lib.load()
    frames = {}

    def frame_of(sc):
        f = frames.get(sc.seed)
        if f is None:
            f = P.SearchFrame(sc.frame, sc.class_costs)
            frames[sc.seed] = f
        return f

    for st in steps:
        for pl in st:
            frame_of(pl.sc)
    B = len(steps[0])
    batch = Batch(max_plans=B, max_cells=B * 256 * 256, launch="auto")
    roof_plans = steps[min(1, len(steps) - 1)]
    if args.roofline_only:
        # warm-up on the same plans: both passes are the same deterministic work
        run_batched(batch, roof_plans, args, frame_of, max_iter=args.roofline_iters, opt_rounds=0, deadline=0.0)
        print(json.dumps({"roofline": roofline_pass(L, batch, roof_plans, args, frame_of)}), flush=True)
        return
    results, walls = [], []
    if args.mode == "served":
        for k in range(args.warmup):
            run_served(batch, steps[k], args, frame_of, B)
        timed = [pl for k in range(args.warmup, args.warmup + args.steps) for pl in steps[k]]
        t0 = time.perf_counter()
        res, wall = run_served(batch, timed, args, frame_of, B)
        t1 = time.perf_counter()
        results.append(res)
        walls.append(wall / args.steps)
    else:
        for k in range(args.warmup):
            run_batched(batch, steps[k], args, frame_of)
        # batched form, timed (no kernel timing inside)
        t0 = time.perf_counter()
        for k in range(args.warmup, args.warmup + args.steps):
            res, wall = run_batched(batch, steps[k], args, frame_of)
            results.append(res)
            walls.append(wall)
        t1 = time.perf_counter()
    allres = [r for rs in results for r in rs]
    scene_of = {pl.tag: pl.sc for st in steps for pl in st}
    lat = np.array([r.latency_ms for r in allres])
    n_b = len(allres)
    lat_ok = [r.latency_ms for r in allres if r.ok]
    lat_fail = [r.latency_ms for r in allres if not r.ok]
    found_b = [r for r in allres if r.ok]
    batched = {"plans_per_s": n_b / (t1 - t0), "found_plans_per_s": len(found_b) / (t1 - t0),
               "p50_ms": float(np.percentile(lat, 50)),
               "p99_ms": float(np.percentile(lat, 99)), "success": float(np.mean([r.ok for r in allres])),
               "mean_iterations": float(np.mean([r.iterations for r in allres])), "plans": n_b,
               "mean_path_cost": float(np.mean([r.path_cost for r in found_b])) if found_b else None,
               "wall_ms_per_step": float(np.mean(walls)),
               "found_p50_p99_max_ms": [round(float(np.percentile(lat_ok, q)), 3) for q in (50, 99, 100)] if lat_ok else None,
               "failed_p50_p99_max_ms": [round(float(np.percentile(lat_fail, q)), 3) for q in (50, 99, 100)] if lat_fail else None,
               # plans whose result came back after their deadline (submit -> result > budget)
               "deadline_miss_rate": round(float(np.mean(lat > args.budget_ms)), 5),
               "validation": validate_paths([(r.tag, scene_of[r.tag], r.path) for r in found_b]),
               "hbm": batch.memory()}
    # plain form on the plans of the timed steps (no kernel timing: the
    # comparison's latencies are the handle's own)
    plain_set = [pl for k in range(args.warmup, args.warmup + min(max(args.plain_steps, 1), args.steps))
                 for pl in steps[k]]
    if args.plain_plans >= 0:
        plain_set = plain_set[:args.plain_plans]
    plain, paths_p = None, {}
    prof_p = None
    if plain_set:
        plain, paths_p, costs_p = run_plain(P, L, plain_set, args, frame_of)
        plain["validation"] = validate_paths([(t, scene_of[t], p) for t, p in paths_p.items() if p is not None])
        same = {r.tag: r.ok for rs in results for r in rs if r.tag in paths_p}
        if same:
            oks = {t: paths_p[t] is not None for t in same}
            plain["success_same_plans_batched"] = float(np.mean(list(same.values())))
            plain["success_same_plans_plain"] = float(np.mean(list(oks.values())))
            plain["success_gap_pp"] = round(100.0 * (plain["success_same_plans_plain"] -
                                                     plain["success_same_plans_batched"]), 3)
            plain["plans_solved_plain_not_batched"] = int(sum(1 for t in same if oks[t] and not same[t]))
            plain["plans_solved_batched_not_plain"] = int(sum(1 for t in same if same[t] and not oks[t]))
        # kernel timing of the plain loop, on a separate pass
        L.fastrrt_profile_reset()
        L.fastrrt_profile_enable(True)
        run_plain(P, L, plain_set[:max(2, args.plain_profile_plans)], args, frame_of)
        L.fastrrt_profile_enable(False)
        prof_p = read_prof(L)
    c1 = None
    if cpu and cpu.get("c1"):
        c1 = c1_compare(P, L, Batch, cpu["c1"], args, frame_of)
        cpu["c1"].pop("per_plan", None)
    roof = roofline_pass(L, batch, roof_plans, args, frame_of)
    out = {
        "metric": "plans/s at p99<=20 ms (served/batched vs plain loop), 256^2 grid, success %",
        "value": round(batched["plans_per_s"], 3), "unit": "plans/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round((t1 - t0) * 1e3 / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32+fp64+int32", "data": "synthetic",
        "config": {"workload": "256x256 frames, rectangles 4-30 px at 5/15/30% coverage, start (128,128,0), "
                               "goal x~U[64,192] on row 0, heading~U[-pi/4,pi/4]; min_dist 22x40, P=40, r=20, "
                               "steer 40 deg, L=5.41; loop to goal, then path_optimize rounds",
                   "mode": args.mode, "plans_per_step": B, "budget_ms": args.budget_ms,
                   "plain_timeout_ms": args.plain_timeout_ms,
                   "max_iter": args.max_iter, "opt_rounds": args.opt_rounds,
                   "launch_grid_waves_tiles_items": list(batch.launch())},
        "p50_ms": round(batched["p50_ms"], 3), "p99_ms": round(batched["p99_ms"], 3),
        "success": round(batched["success"], 4),
        "batched": batched, "plain": plain,
        "batched_vs_plain": (round(batched["plans_per_s"] / plain["plans_per_s"], 3) if plain else None),
        "kernel_ms_plain": prof_json(prof_p) if prof_p else None,
        "c1": c1,
        "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
