"""Side-by-side drivers: the oracle planner and the HIP planner on one scene.

Test infrastructure (imports the oracle).
"""
import ctypes

import numpy as np

import _oracle as O

STATE_KEYS = ("type", "parent_x", "parent_z", "child_count", "heading", "cost")


class Params:
    def __init__(self, min_dist=(22, 40), max_path=40.0, dist_tol=20.0, herr_deg=10.0, v=1.0, steer_deg=40.0,
                 vehicle_len=5.412658773, max_curv=-1.0):
        self.min_dist = min_dist
        self.max_path = max_path
        self.dist_tol = dist_tol
        self.herr = np.float32(np.radians(herr_deg))
        self.v = v
        self.steer_deg = steer_deg
        self.vehicle_len = vehicle_len
        self.max_curv = max_curv


class OraclePlanner:
    def __init__(self, sc, prm, seed, stream, region_side=None, timeout_ms=-1):
        self.L = O.lib()
        self.sc = sc
        self.r = self.L.og_rrt_new(sc.width, sc.height, sc.perception_m, sc.perception_m, prm.steer_deg,
                                   prm.vehicle_len, sc.ego_lb[0], sc.ego_lb[1], sc.ego_ub[0], sc.ego_ub[1],
                                   prm.max_curv)
        self.L.og_rrt_set_seed(self.r, seed, stream)
        self.frame = np.ascontiguousarray(sc.frame, np.float32)
        self.costs = np.ascontiguousarray(sc.class_costs, np.float32)
        self.L.og_rrt_set_plan_data(self.r, O.fptr(self.frame), O.fptr(self.costs), len(self.costs), sc.start[0],
                                    sc.start[1], sc.start[2], sc.goal[0], sc.goal[1], sc.goal[2], prm.v,
                                    prm.min_dist[0], prm.min_dist[1], timeout_ms, prm.max_path, prm.dist_tol,
                                    prm.herr)
        if region_side:
            self.L.og_set_region_side(self.L.og_rrt_graph(self.r), region_side)

    def __del__(self):
        try:
            self.L.og_rrt_free(self.r)
        except Exception:
            pass

    def graph(self):
        return self.L.og_rrt_graph(self.r).contents

    def plane(self, name, dtype):
        g = self.graph()
        return np.ctypeslib.as_array(getattr(g, name), shape=(self.sc.height, self.sc.width)).view(dtype).copy()

    def state(self):
        return {"type": self.plane("type", np.int32), "parent_x": self.plane("px", np.int32),
                "parent_z": self.plane("pz", np.int32), "child_count": self.plane("derive", np.int32),
                "heading": self.plane("heading", np.float32), "cost": self.plane("cost", np.float32)}

    def goal_planes(self):
        return {"goal_flag": (self.plane("gc_flag", np.float32) > 0).astype(np.uint8),
                "goal_heading": self.plane("gc_heading", np.float32),
                "goal_cost": self.plane("gc_cost", np.float32)}

    def search_init(self):
        self.L.og_rrt_search_init(self.r)

    def loop(self, smart=False):
        return bool(self.L.og_rrt_loop(self.r, int(smart)))

    def goal_reached(self):
        return bool(self.L.og_rrt_goal_reached(self.r))

    def path(self):
        buf = np.zeros(3 * 65536, np.float32)
        n = self.L.og_rrt_get_planned_path(self.r, O.fptr(buf), 65536)
        return buf[:3 * n].reshape(n, 3)

    def interpolated(self):
        buf = np.zeros(3 * 262144, np.float32)
        n = self.L.og_rrt_interpolate_planned_path(self.r, O.fptr(buf), 262144)
        return buf[:3 * n].reshape(n, 3)

    def path_optimize(self):
        return bool(self.L.og_rrt_path_optimize(self.r))


class GpuPlanner:
    def __init__(self, sc, prm, seed, stream, region_side=None):
        from fastrrt_amd import pyfastrrt as P
        self.P = P
        self.sc = sc
        ego = P.EgoParams(sc.width, sc.height, sc.perception_m, sc.perception_m, prm.steer_deg, prm.vehicle_len,
                          sc.ego_lb, sc.ego_ub, prm.max_curv, sc.class_costs)
        self.rrt = P.FastRRT(ego)
        self.frame = P.SearchFrame(sc.frame, sc.class_costs)
        sp = P.SearchParams(self.frame, P.Waypoint(sc.start[0], sc.start[1], P.angle(sc.start[2])),
                            P.Waypoint(sc.goal[0], sc.goal[1], P.angle(sc.goal[2])), prm.v, prm.min_dist, -1,
                            prm.max_path, prm.dist_tol, P.angle(float(prm.herr)))
        self.rrt.set_seed(seed, stream)
        self.rrt.set_plan_data(sp)
        if region_side:
            self.rrt.set_density_region(region_side)

    def state(self):
        s = self.rrt.export_state(goal=False)
        return {k: s[k] for k in STATE_KEYS}

    def goal_planes(self):
        s = self.rrt.export_state()
        return {k: s[k] for k in ("goal_flag", "goal_heading", "goal_cost")}

    def search_init(self):
        self.rrt.search_init()

    def loop(self, smart=False):
        return bool(self.rrt.loop(smart))

    def goal_reached(self):
        return bool(self.rrt.goal_reached())

    def path(self, interpolate=False):
        p = self.rrt.get_planned_path(interpolate)
        if not p:
            return np.zeros((0, 3), np.float32)
        return np.array([[w.x, w.z, w.heading.rad()] for w in p], np.float32)

    def path_optimize(self):
        return bool(self.rrt.path_optimize())


def diff_states(a, b, mask_null_payload=True):
    """Return a list of mismatching plane names (bit-exact comparison)."""
    bad = []
    for k in STATE_KEYS:
        x, y = a[k], b[k]
        if x.dtype == np.float32:
            eq = x.view(np.int32) == y.view(np.int32)
        else:
            eq = x == y
        if not eq.all():
            bad.append((k, int((~eq).sum()), np.argwhere(~eq)[:5].tolist()))
    return bad


def diff_goal_planes(a, b):
    bad = []
    if not (a["goal_flag"] == b["goal_flag"]).all():
        bad.append(("goal_flag", int((a["goal_flag"] != b["goal_flag"]).sum())))
    m = a["goal_flag"] > 0
    for k in ("goal_heading", "goal_cost"):
        eq = a[k][m].view(np.int32) == b[k][m].view(np.int32)
        if not eq.all():
            bad.append((k, int((~eq).sum())))
    return bad


def run_side_by_side(sc, prm, seed, stream, max_iter, smart=False, goal_planes_first=False, region_side=None):
    """Oracle and GPU planner in lock step, grids compared bit for bit after
    every loop().  The handle evaluates the goal-connection planes lazily; they
    are compared before the search (which finishes the pass up front) or after
    it (the search ran on lazily evaluated cells)."""
    print(f"[parity] seed={seed} size={sc.width} cov={sc.coverage} smart={smart}", flush=True)
    o = OraclePlanner(sc, prm, seed, stream, region_side)
    g = GpuPlanner(sc, prm, seed, stream, region_side)
    if goal_planes_first:
        bad = diff_goal_planes(o.goal_planes(), g.goal_planes())
        assert not bad, f"goal-connection planes differ: {bad}"
    o.search_init()
    g.search_init()
    assert o.goal_reached() == g.goal_reached()
    it = 0
    while True:
        go = (not o.goal_reached()) and o.loop(smart)
        gg = (not g.goal_reached()) and g.loop(smart)
        it += 1
        bad = diff_states(o.state(), g.state())
        assert not bad, f"iteration {it}: grid differs {bad}"
        assert go == gg, f"iteration {it}: loop() returned {go} (oracle) vs {gg} (gpu)"
        if not go or it >= max_iter:
            break
    if not goal_planes_first:
        bad = diff_goal_planes(o.goal_planes(), g.goal_planes())
        assert not bad, f"goal-connection planes differ after the search: {bad}"
    return o, g, it
