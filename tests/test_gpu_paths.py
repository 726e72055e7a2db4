"""Every path the GPU returns is a valid plan on its frame, checked by the
oracle: it starts at the start cell, ends within the distance tolerance of
the goal, every waypoint lies in the grid and its footprint (at the
waypoint's heading) is free of obstacles (og_feasible, path_feasible.cu:13-78),
and it has no repeated cell (the parent walk is a chain, not a cycle).

Batched form and plain FastRRT handle, over the 256x256 sweep workload."""
import math

import numpy as np
import pytest

import _oracle as O
from fastrrt_amd import pyfastrrt as P
from fastrrt_amd import scenes
from fastrrt_amd.batch import Batch, PlanSpec

pytestmark = pytest.mark.gpu

DIST_TOL = 20.0


def _checker(sc):
    g = O.Graph(sc.width, sc.height)
    g.physical(sc.perception_m, sc.perception_m, 40.0, 5.412658773)
    g.search(sc.min_dist[0], sc.min_dist[1], sc.ego_lb[0], sc.ego_lb[1], sc.ego_ub[0], sc.ego_ub[1])
    g.costs(sc.class_costs)
    return g


def _check(sc, path):
    assert path.shape[0] >= 2, "a found plan has at least the start and one more waypoint"
    g = _checker(sc)
    frame = np.ascontiguousarray(sc.frame, np.float32)
    xs, zs = path[:, 0].astype(np.int64), path[:, 1].astype(np.int64)
    assert (xs[0], zs[0]) == (sc.start[0], sc.start[1])
    assert math.hypot(xs[-1] - sc.goal[0], zs[-1] - sc.goal[1]) <= DIST_TOL
    assert len(set(zip(xs.tolist(), zs.tolist()))) == len(xs), "repeated cell in the path"
    for (x, z, h) in zip(xs, zs, path[:, 2]):
        assert 0 <= x < sc.width and 0 <= z < sc.height
        assert O.lib().og_feasible(g.g, O.fptr(frame), int(x), int(z), float(h)) == 1, (x, z, h)


def test_batched_paths_are_valid():
    scs = scenes.scenes_256_sweep(8, base_seed=7100)
    frames = [P.SearchFrame(sc.frame, sc.class_costs) for sc in scs]
    b = Batch(max_plans=len(scs), max_cells=len(scs) * 256 * 256)
    for i, sc in enumerate(scs):
        b.submit(PlanSpec(frames[i], sc.start, sc.goal, min_dist=sc.min_dist, ego_lb=sc.ego_lb, ego_ub=sc.ego_ub,
                          max_iterations=80, optimize_rounds=10, seed=4242, stream=i, tag=i))
    res = b.drain()
    assert len(res) == len(scs)
    found = [r for r in res if r.ok]
    assert all(r.status in (2, 3, 4) for r in res)
    assert all(r.path.shape[0] == 0 for r in res if not r.ok)
    print(f"[paths] batched: {len(found)}/{len(res)} found, "
          f"{sum(1 for r in res if r.status == 4)} goal reached without a feasible end node", flush=True)
    assert len(found) >= 3
    for r in found:
        _check(scs[r.tag], r.path)


def test_plain_paths_are_valid():
    scs = scenes.scenes_256_sweep(3, base_seed=7200)
    sc0 = scs[0]
    ego = P.EgoParams(256, 256, sc0.perception_m, sc0.perception_m, 40.0, 5.412658773, sc0.ego_lb, sc0.ego_ub, -1.0,
                      sc0.class_costs)
    rrt = P.FastRRT(ego)
    n_found = 0
    for i, sc in enumerate(scs):
        frame = P.SearchFrame(sc.frame, sc.class_costs)
        rrt.set_seed(4343, i)
        rrt.set_plan_data(P.SearchParams(frame, P.Waypoint(sc.start[0], sc.start[1], P.angle(sc.start[2])),
                                         P.Waypoint(sc.goal[0], sc.goal[1], P.angle(sc.goal[2])), 1.0, sc.min_dist, -1,
                                         40.0, DIST_TOL, P.angle.new_deg(10)))
        rrt.search_init()
        it = 0
        while not rrt.goal_reached() and rrt.loop(False) and it < 80:
            it += 1
        if not rrt.goal_reached():
            continue
        k = 0
        while k < 10 and rrt.path_optimize():
            k += 1
        wp = rrt.get_planned_path()
        path = np.array([[w.x, w.z, w.heading.rad()] for w in wp], np.float32)
        _check(sc, path)
        n_found += 1
    print(f"[paths] plain: {n_found}/{len(scs)} found", flush=True)
    assert n_found >= 2
