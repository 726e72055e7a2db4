"""Checks of a returned plan, in numpy, independent of the HIP kernels.

A plan is a list of waypoints (x, z, heading) in grid pixels, first the
start.  check_path() tests what every plan the reference returns satisfies:

* the chain: it starts at the start cell, stays in the grid and visits no
  cell twice (a parent walk, src/fastrrt.cpp:204-232, is a chain);
* the goal test: the last waypoint lies within the distance tolerance of the
  goal and its heading within the heading tolerance
  (findBestNode, src/cuda/graph_best_node.cu:134-178);
* every edge is drivable: its end waypoint's footprint at the waypoint's
  heading is free of obstacles (the (2 mz + 1) x (2 mx + 1) rotated lattice
  of src/cuda/path_feasible.cu:18-78, samples outside the grid or in the ego
  box skipped, bit 0 of channel 2 marking a pixel feasible at every
  heading, sampled in float32 from float64 cos/sin rounded once), or the
  edge is a Hermite connection the reference accepts: the curve of
  checkDirectConnectionToGoal (src/cuda/graph_direct_goal_connection.cu:
  16-167) between the two waypoints and their headings, whose accepted
  points pass the curvature bound and have free footprints at the curve's
  own heading.  Goal connections and optimisation splices are such curves,
  and they store the end tangent as the node's heading, while the cell
  itself was tested at the heading of the first curve point that entered
  it: a goal-connection cell can be blocked at its stored heading;
* the heading continuity of every edge: its mean turning rate
  2 |sin(dh / 2)| / chord (chord shortened by the sqrt(2) px the cell
  truncation of its two ends can take off) is at most the curvature its
  kind of edge allows.  A rollout is an arc of constant curvature, for which
  this rate is the curvature itself, at most cos(beta) tan(steer_max) / (2 lr)
  (src/cuda/kinematics.cu:160-191); a Hermite edge whose curvature stays
  below k (and turns less than pi) has rate <= k: node-to-cell connections
  and optimisation splices use the physical bound
  2 tan(steer_max) / (0.5 L sqrt(4 + tan(steer_max))) (src/cuda/graph.cu:300),
  and only the last edge can be a cell-to-goal connection, bounded by 0.8
  (src/fastrrt.cpp:79).  The reference tests a Hermite edge's curvature at
  its in-grid sample points only, so an edge that passes the Hermite check
  (short curves near the grid border have few such points) carries no
  turning bound beyond that check; a cell less than 2 px from the goal
  connects to it with a one-point curve whose parameter is 0/0
  (src/cuda/graph_direct_goal_connection.cu:57: the tests see NaN and
  pass).

Each check returns its verdict; check_path() returns the names of the checks
a plan fails (empty: valid) and the edges' turning rates.
"""
import math

import numpy as np


def footprint_free(obstacle, trav_all, x, z, heading, mx, mz, ego=None):
    """path_feasible.cu:18-78 on an obstacle mask (H, W) of bools."""
    H, W = obstacle.shape
    if trav_all is not None and trav_all[z, x]:
        return True
    c = np.float32(math.cos(float(heading)))
    s = np.float32(math.sin(float(heading)))
    i = np.arange(-mz, mz + 1, dtype=np.float32)[:, None]
    j = np.arange(-mx, mx + 1, dtype=np.float32)[None, :]
    fx = (j * c - i * s) + np.float32(x)
    fz = (j * s + i * c) + np.float32(z)
    xl = np.trunc(fx).astype(np.int64)
    zl = np.trunc(fz).astype(np.int64)
    ok = (xl >= 0) & (xl < W) & (zl >= 0) & (zl < H)
    if ego is not None:
        (lbx, lbz), (ubx, ubz) = ego
        ok &= ~((xl >= lbx) & (xl <= ubx) & (zl >= ubz) & (zl <= lbz))
    return not bool(obstacle[zl[ok], xl[ok]].any())


def border_rollout_end_ok(obstacle, trav_all, prev, cur, mx, mz, ego=None, samples=512):
    """A rollout that leaves the grid stops in its last in-grid cell, a cell
    on the border, and returns the heading after the step that left
    (kinematics.cu:183-202: the heading is advanced before the grid test
    breaks the loop), while that cell's footprint was tested with the
    heading it was entered with, up to a cell's worth of 0.1 px steps
    earlier on the same arc.  Along a rollout the heading moves
    monotonically from the parent's to the end's, so the tested heading lies
    in that range: True when `cur` is on the border and its footprint is
    free at one of `samples` headings spread over the range."""
    H, W = obstacle.shape
    x, z = int(cur[0]), int(cur[1])
    if not (x == 0 or z == 0 or x == W - 1 or z == H - 1):
        return False
    for h in np.linspace(float(prev[2]), float(cur[2]), samples, dtype=np.float32):
        if footprint_free(obstacle, trav_all, x, z, np.float32(h), mx, mz, ego):
            return True
    return False


def border_turn_ok(prev, cur, shape, bound, step_px=0.1):
    """Heading continuity of a rollout that left the grid (kinematics.cu:
    183-202): its end heading also carries the turn of the steps taken in
    its last cell and of the step that left, at most sqrt(2) + 2 steps of
    arc at the rollout's curvature, which the chord between the cells does
    not see.  True when `cur` is on the border and the edge's turning rate,
    that turn taken off, is within `bound` (step_px = v dt, 0.1 px at
    1 m/s)."""
    H, W = shape
    x, z = int(cur[0]), int(cur[1])
    if not (x == 0 or z == 0 or x == W - 1 or z == H - 1):
        return False
    dh = abs(float(cur[2]) - float(prev[2]))
    chord = math.hypot(float(cur[0]) - float(prev[0]), float(cur[1]) - float(prev[1]))
    dh = max(dh - bound * (math.sqrt(2.0) + 2.0 * step_px), 0.0)
    return 2.0 * abs(math.sin(dh / 2.0)) / max(chord - math.sqrt(2.0), 1.0) <= bound + 1e-9


def obstacle_mask(frame, class_costs):
    """classCost[TO_INT(frame.x)] < 0; a class outside the table is an obstacle."""
    cls = np.trunc(frame[:, :, 0]).astype(np.int64)
    costs = np.asarray(class_costs, np.float32)
    inside = (cls >= 0) & (cls < len(costs))
    return ~inside | (costs[np.clip(cls, 0, len(costs) - 1)] < 0)


def max_turning_rates(max_steer_deg=40.0, vehicle_length=5.412658773, goal_curvature=0.8):
    """(bound of an inner edge, bound of the last edge)"""
    t = math.tan(math.radians(max_steer_deg))
    lr = vehicle_length / 2.0
    rollout = math.cos(math.atan(t / 2.0)) * t / (2.0 * lr)
    physical = 2.0 * t / (0.5 * vehicle_length * math.sqrt(4.0 + t))
    inner = max(rollout, physical)
    return inner, max(inner, goal_curvature)


def hermite_edge_ok(obstacle, trav_all, a, b, max_curv, mx, mz, ego=None):
    """checkDirectConnectionToGoal (graph_direct_goal_connection.cu:16-167)
    from waypoint a = (x, z, heading) to b, in float64 as the reference
    (tangents and headings in float32): True when the curve has points, every
    accepted point passes the curvature bound and has a free footprint at the
    curve's heading, and it ends in b's row or column (the reference's &&)."""
    H, W = obstacle.shape
    x, z = int(a[0]), int(a[1])
    gx, gz = int(b[0]), int(b[1])
    f32 = np.float32
    dx, dz = gx - x, gz - z
    distance = f32(math.sqrt(float(f32(dx * dx + dz * dz))))
    n = int(distance)
    if n <= 0:
        return False
    if n == 1:
        return True  # t = 0 / 0: the reference's tests see NaN and pass
    a1 = f32(float(f32(a[2])) - math.pi / 2)
    a2 = f32(float(f32(b[2])) - math.pi / 2)
    t1x, t1y = float(distance * f32(math.cos(float(a1)))), float(distance * f32(math.sin(float(a1))))
    t2x, t2y = float(distance * f32(math.cos(float(a2)))), float(distance * f32(math.sin(float(a2))))
    last_x = last_z = -1
    for i in range(n):
        t = i / (n - 1)
        t2, t3 = t * t, t * t * t
        h00, h10, h01, h11 = 2 * t3 - 3 * t2 + 1, t3 - 2 * t2 + t, -2 * t3 + 3 * t2, t3 - t2
        px = h00 * x + h10 * t1x + h01 * gx + h11 * t2x
        pz = h00 * z + h10 * t1y + h01 * gz + h11 * t2y
        if px < 0 or px >= W or pz < 0 or pz >= H:
            continue
        cx, cz = int(px), int(pz)
        if cx == last_x and cz == last_z:
            continue
        d00, d10, d01, d11 = 6 * t2 - 6 * t, 3 * t2 - 4 * t + 1, -6 * t2 + 6 * t, 3 * t2 - 2 * t
        ddx = d00 * x + d10 * t1x + d01 * gx + d11 * t2x
        ddz = d00 * z + d10 * t1y + d01 * gz + d11 * t2y
        heading = f32(math.atan2(float(f32(ddz)), float(f32(ddx))) + math.pi / 2)
        if max_curv > 0:
            e00, e10, e01, e11 = 12 * t - 6, 6 * t - 4, -12 * t + 6, 6 * t - 2
            d2x = e00 * x + e10 * t1x + e01 * gx + e11 * t2x
            d2z = e00 * z + e10 * t1y + e01 * gz + e11 * t2y
            v = ddx * ddx + ddz * ddz
            if v <= 0 or f32(abs(ddx * d2z - ddz * d2x) / (v * math.sqrt(v))) > f32(max_curv):
                return False
        last_x, last_z = cx, cz
        if not footprint_free(obstacle, trav_all, cx, cz, heading, mx, mz, ego):
            return False
    return not (last_x != gx and last_z != gz)


def check_path(frame, class_costs, path, start, goal, min_dist, ego=None, dist_tol=20.0,
               heading_tol=math.radians(10), max_steer_deg=40.0, vehicle_length=5.412658773):
    """Returns (failed check names, turning rate per edge)."""
    path = np.asarray(path, np.float64).reshape(-1, 3)
    H, W = frame.shape[:2]
    fails = []
    if path.shape[0] < 2:
        return ["length"], np.zeros(0)
    xs = path[:, 0].astype(np.int64)
    zs = path[:, 1].astype(np.int64)
    if (xs[0], zs[0]) != (int(start[0]), int(start[1])):
        fails.append("start")
    if ((xs < 0) | (xs >= W) | (zs < 0) | (zs >= H)).any():
        return fails + ["grid"], np.zeros(0)
    if len(set(zip(xs.tolist(), zs.tolist()))) != len(xs):
        fails.append("repeated_cell")
    if math.hypot(float(xs[-1] - goal[0]), float(zs[-1] - goal[1])) > dist_tol:
        fails.append("goal_distance")
    if abs(np.float32(path[-1, 2]) - np.float32(goal[2])) > np.float32(heading_tol):
        fails.append("goal_heading")
    obst = obstacle_mask(frame, class_costs)
    trav = (np.trunc(frame[:, :, 2]).astype(np.int64) & 1) != 0
    mx, mz = int(min_dist[0]) // 2, int(min_dist[1]) // 2
    inner, last = max_turning_rates(max_steer_deg, vehicle_length)
    if not footprint_free(obst, trav, int(xs[0]), int(zs[0]), np.float32(path[0, 2]), mx, mz, ego):
        fails.append("footprint")
    hermite = np.zeros(len(xs) - 1, bool)
    for k in range(1, len(xs)):
        if footprint_free(obst, trav, int(xs[k]), int(zs[k]), np.float32(path[k, 2]), mx, mz, ego):
            continue
        hermite[k - 1] = hermite_edge_ok(obst, trav, path[k - 1], path[k], last, mx, mz, ego)
        if not hermite[k - 1] and not border_rollout_end_ok(obst, trav, path[k - 1], path[k], mx, mz, ego):
            fails.append("footprint")
            break
    dh = np.diff(path[:, 2])
    chord = np.hypot(np.diff(xs), np.diff(zs)).astype(np.float64)
    rate = 2.0 * np.abs(np.sin(dh / 2.0)) / np.maximum(chord - math.sqrt(2.0), 1.0)
    bound = np.full(rate.shape, inner)
    bound[-1] = last
    over = rate > bound + 1e-9
    for k in np.nonzero(over)[0]:
        # an edge whose Hermite curve the reference accepts was tested at
        # its sample points, which bound its turning
        if not hermite[k]:
            hermite[k] = hermite_edge_ok(obst, trav, path[k], path[k + 1], bound[k], mx, mz, ego)
        if not hermite[k] and not border_turn_ok(path[k], path[k + 1], (H, W), bound[k]):
            fails.append("heading_continuity")
            break
    return fails, rate
