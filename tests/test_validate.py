"""fastrrt_amd.validate (the bench's path checker) against the oracle: the
footprint verdict equals og_feasible (path_feasible.cu:13-78) on random
poses of benchmark frames, oracle-planned paths pass every check, and
corrupted paths fail the check they break."""
import ctypes
import math

import numpy as np

import _oracle as O
from _parity import OraclePlanner, Params
from fastrrt_amd import scenes
from fastrrt_amd import validate as V


def test_footprint_equals_oracle():
    L = O.lib()
    rng = np.random.default_rng(8)
    n = 0
    for seed, cov in ((1, 0.1), (2, 0.3)):
        sc = scenes.scene_256(seed, cov)
        frame = np.ascontiguousarray(sc.frame, np.float32)
        frame[100:104, 100:110, 2] = 1.0  # pixels feasible at every heading
        g = O.Graph(256, 256)
        g.search(22, 40, sc.ego_lb[0], sc.ego_lb[1], sc.ego_ub[0], sc.ego_ub[1])
        g.costs(sc.class_costs)
        obst = V.obstacle_mask(frame, sc.class_costs)
        trav = (np.trunc(frame[:, :, 2]).astype(np.int64) & 1) != 0
        for _ in range(1500):
            x, z = int(rng.integers(0, 256)), int(rng.integers(0, 256))
            h = np.float32(rng.uniform(-math.pi, math.pi))
            exp = L.og_feasible(g.g, O.fptr(frame), x, z, float(h)) == 1
            got = V.footprint_free(obst, trav, x, z, h, 11, 20, (sc.ego_lb, sc.ego_ub))
            assert got == exp, (seed, x, z, h)
            n += exp
    assert 100 < n < 2900


def test_oracle_paths_pass_and_corruptions_fail():
    checked = 0
    for seed in (3, 11, 19, 29):
        sc = scenes.scene_sized(seed, 160, 0.08, min_dist=(10, 16))
        prm = Params(min_dist=(10, 16), max_path=14.0, dist_tol=6.0)
        o = OraclePlanner(sc, prm, 7 + seed, seed)
        o.search_init()
        for _ in range(200):
            if o.goal_reached() or not o.loop():
                break
        if not o.goal_reached():
            continue
        for r in range(3):
            path = o.path()
            fails, rate = V.check_path(sc.frame, sc.class_costs, path, sc.start, sc.goal, (10, 16), None,
                                       6.0, float(prm.herr))
            assert not fails, (seed, r, fails)
            checked += 1
            if not o.path_optimize():
                break
        blocked = np.array(sc.frame, copy=True)
        blocked[int(path[1, 1]), int(path[1, 0]), 0] = scenes.OBSTACLE_CLASS
        assert "footprint" in V.check_path(blocked, sc.class_costs, path, sc.start, sc.goal, (10, 16), None, 6.0,
                                           float(prm.herr))[0]
        far = np.array(sc.goal, dtype=float)
        far[1] += 30
        assert "goal_distance" in V.check_path(sc.frame, sc.class_costs, path, sc.start, far, (10, 16), None, 6.0,
                                               float(prm.herr))[0]
    assert checked >= 4


def test_heading_continuity_on_a_straight_path():
    """Waypoints 6 px apart in a column, heading 0, on a free frame: valid;
    a middle heading turned by 2 rad gives its edges a turning rate of
    2 sin(1) / (6 - sqrt 2) = 0.37, above the 0.28 an inner edge allows."""
    frame = np.zeros((128, 64, 3), np.float32)
    frame[:, :, 0] = scenes.FREE_CLASS
    path = np.array([[32, 100 - 6 * k, 0.0] for k in range(6)], np.float32)
    goal = (32, 70, 0.0)
    fails, rate = V.check_path(frame, scenes.SEGMENTATION_CLASS_COST, path, (32, 100), goal, (10, 16))
    assert not fails and rate.max() == 0.0
    bad = path.copy()
    bad[2, 2] = 2.0
    fails, rate = V.check_path(frame, scenes.SEGMENTATION_CLASS_COST, bad, (32, 100), goal, (10, 16))
    assert fails == ["heading_continuity"] and abs(rate[1] - 2 * math.sin(1.0) / (6 - math.sqrt(2))) < 1e-12


def test_hermite_edge_equals_oracle_connection():
    """validate.hermite_edge_ok restates checkDirectConnectionToGoal
    (graph_direct_goal_connection.cu:16-167): its verdict equals the
    oracle's og_check_direct_connection (cost >= 0) on random pairs of a
    benchmark frame, curvature bounds 0.8 and the physical one."""
    L = O.lib()
    rng = np.random.default_rng(21)
    sc = scenes.scene_256(4, 0.12)
    frame = np.ascontiguousarray(sc.frame, np.float32)
    g = O.Graph(256, 256)
    g.physical(34.64, 34.64, 40.0, 5.412658773)
    g.search(22, 40, sc.ego_lb[0], sc.ego_lb[1], sc.ego_ub[0], sc.ego_ub[1])
    g.costs(sc.class_costs)
    obst = V.obstacle_mask(frame, sc.class_costs)
    trav = (np.trunc(frame[:, :, 2]).astype(np.int64) & 1) != 0
    inner, last = V.max_turning_rates()
    n_ok = 0
    for k in range(600):
        x, z = int(rng.integers(0, 256)), int(rng.integers(0, 256))
        gx = int(np.clip(x + rng.integers(-40, 41), 0, 255))
        gz = int(np.clip(z + rng.integers(-40, 41), 0, 255))
        h1 = float(np.float32(rng.uniform(-math.pi, math.pi)))
        h2 = float(np.float32(rng.uniform(-math.pi, math.pi))) if k % 2 else h1
        mc = float(np.float32(last if k % 3 else inner))
        g.add(x, z, h1, -1, -1, 0)
        c = L.og_check_direct_connection(g.g, O.fptr(frame), mc, x, z, h1, gx, gz, h2, 0, 0)
        got = V.hermite_edge_ok(obst, trav, (x, z, h1), (gx, gz, h2), mc, 11, 20, (sc.ego_lb, sc.ego_ub))
        assert got == (c >= 0), (x, z, h1, gx, gz, h2, mc, c)
        n_ok += got
    assert 20 < n_ok < 580


def test_rollouts_leaving_the_grid_pass_the_border_rule():
    """The oracle's rollouts (kinematics.cu:123-231) that leave the grid stop
    in a border cell with the heading after the step that left: every such
    end the oracle accepts passes border_rollout_end_ok, including ends whose
    footprint at the returned heading collides; a border waypoint whose
    footprint collides at every heading of its range is rejected."""
    L = O.lib()
    sc = scenes.scene_256(41, 0.10)
    frame = np.ascontiguousarray(sc.frame, np.float32)
    # free top rows with thin posts that only footprints turned away from
    # the vertical reach
    frame[0:24, :, 0] = scenes.FREE_CLASS
    frame[12:20, :, 0] = np.where(np.arange(256) % 48 < 4, scenes.OBSTACLE_CLASS, scenes.FREE_CLASS)
    ego = (sc.ego_lb, sc.ego_ub)
    g = O.Graph(256, 256)
    g.physical(256, 256, 40.0, 5.412658773)
    g.costs(sc.class_costs)
    g.search(22, 40, sc.ego_lb[0], sc.ego_lb[1], sc.ego_ub[0], sc.ego_ub[1])
    obst = V.obstacle_mask(frame, sc.class_costs)
    trav = (np.trunc(frame[:, :, 2]).astype(np.int64) & 1) != 0
    rng = np.random.default_rng(12)
    out = (ctypes.c_float * 4)()
    max_steer = 40.0 * math.pi / 180.0
    n_border = n_collide = 0
    for _ in range(3000):
        x, z = int(rng.integers(30, 226)), int(rng.integers(1, 6))
        h0 = float(np.float32(rng.uniform(-0.6, 0.6)))
        steer = float(np.float32(rng.uniform(-max_steer, max_steer)))
        g.add(x, z, h0, -1, -1, 0)
        L.og_kinematic_new_path(g.g, O.fptr(frame), x, z, steer, float(rng.integers(5, 41)), 1.0, out)
        if out[0] < 0:
            continue
        ex, ez, eh = int(out[0]), int(out[1]), float(out[3])
        if ez != 0:
            continue
        n_border += 1
        n_collide += not V.footprint_free(obst, trav, ex, ez, np.float32(eh), 11, 20, ego)
        assert V.border_rollout_end_ok(obst, trav, (x, z, h0), (ex, ez, eh), 11, 20, ego), (x, z, h0, steer)
    assert n_border > 500 and n_collide > 20, (n_border, n_collide)
    # a border cell inside an obstacle block collides at every heading
    blocked = np.array(frame, copy=True)
    blocked[0:6, 100:140, 0] = scenes.OBSTACLE_CLASS
    ob2 = V.obstacle_mask(blocked, sc.class_costs)
    assert not V.border_rollout_end_ok(ob2, trav, (120, 3, 0.0), (120, 0, 0.3), 11, 20, ego)
    # an interior waypoint is not covered by the rule
    assert not V.border_rollout_end_ok(obst, trav, (120, 30, 0.0), (120, 20, 0.3), 11, 20, ego)


def test_border_rollout_turns_pass_the_continuity_rule():
    """Rollouts of the oracle that leave the grid through any side: the
    turning rate of the edge to their end exceeds the inner bound for some
    (the end heading carries the steps taken after the last cell was
    entered), border_turn_ok accepts every one, no rollout that stays inside
    exceeds the bound, and a short border edge turning 1.2 rad is rejected."""
    L = O.lib()
    inner, _ = V.max_turning_rates()
    g = O.Graph(256, 256)
    g.physical(256, 256, 40.0, 5.412658773)
    g.costs([0, 1, 2, 3, 4, -1])
    g.search(22, 40)
    frame = O.empty_frame(256, 256, scenes.FREE_CLASS)
    out = (ctypes.c_float * 4)()
    rng = np.random.default_rng(100)
    n_border = n_over = 0
    for _ in range(12000):
        side = int(rng.integers(0, 4))
        x, z = int(rng.integers(1, 255)), int(rng.integers(1, 40))
        if side == 1:
            z = 255 - z
        elif side == 2:
            x, z = z, x
        elif side == 3:
            x, z = 255 - z, x
        h0 = float(np.float32(rng.uniform(-math.pi, math.pi)))
        steer = float(np.float32(rng.uniform(-0.698, 0.698)))
        g.add(x, z, h0, -1, -1, 0)
        L.og_kinematic_new_path(g.g, O.fptr(frame), x, z, steer, float(rng.integers(5, 41)), 1.0, out)
        if out[0] < 0:
            continue
        a, b = (x, z, h0), (float(out[0]), float(out[1]), float(out[3]))
        chord = math.hypot(b[0] - a[0], b[1] - a[1])
        over = 2 * abs(math.sin((b[2] - a[2]) / 2)) / max(chord - math.sqrt(2), 1.0) > inner + 1e-9
        if int(b[0]) in (0, 255) or int(b[1]) in (0, 255):
            n_border += 1
            n_over += over
            assert V.border_turn_ok(a, b, (256, 256), inner), (a, b)
        else:
            assert not over, (a, b)
    assert n_border > 1000 and n_over > 10, (n_border, n_over)
    # a short border edge turning 1.2 rad, and the rule off the border
    assert not V.border_turn_ok((120, 3, 0.0), (121, 0, 1.2), (256, 256), inner)
    assert not V.border_turn_ok((120, 30, 0.0), (121, 27, 0.3), (256, 256), inner)
