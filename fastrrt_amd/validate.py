"""Checks of a returned plan, in numpy, independent of the HIP kernels.

A plan is a list of waypoints (x, z, heading) in grid pixels, first the
start.  check_path() tests what every plan the reference returns satisfies:

* the chain: it starts at the start cell, stays in the grid and visits no
  cell twice (a parent walk, src/fastrrt.cpp:204-232, is a chain);
* the goal test: the last waypoint lies within the distance tolerance of the
  goal and its heading within the heading tolerance
  (findBestNode, src/cuda/graph_best_node.cu:134-178);
* the footprint of every waypoint at its heading is free of obstacles:
  the (2 mz + 1) x (2 mx + 1) rotated lattice of
  src/cuda/path_feasible.cu:18-78 (samples outside the grid or in the ego
  box skipped, bit 0 of channel 2 marking a pixel feasible at every
  heading), sampled in float32 from float64 cos/sin rounded once;
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
  (src/fastrrt.cpp:79).  A cell less than 2 px from the goal connects to it
  with a one-point curve whose parameter is 0/0
  (src/cuda/graph_direct_goal_connection.cu:57: the curvature and footprint
  tests see NaN and pass), so a last edge shorter than 2 px carries no bound.

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
    if not all(footprint_free(obst, trav, int(x), int(z), np.float32(h), mx, mz, ego)
               for x, z, h in zip(xs, zs, path[:, 2])):
        fails.append("footprint")
    dh = np.diff(path[:, 2])
    chord = np.hypot(np.diff(xs), np.diff(zs)).astype(np.float64)
    rate = 2.0 * np.abs(np.sin(dh / 2.0)) / np.maximum(chord - math.sqrt(2.0), 1.0)
    inner, last = max_turning_rates(max_steer_deg, vehicle_length)
    if (rate[:-1] > inner + 1e-9).any() or (chord[-1] >= 2.0 and rate[-1] > last + 1e-9):
        fails.append("heading_continuity")
    return fails, rate
