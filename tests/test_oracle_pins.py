"""Independent pins of the oracle functions the reference's own tests leave
open: path optimisation (graph_optimization.cu:25-195 behind
fastrrt.cpp:183-196), the goal-window argmin of the loop's direct goal
connection (graph_check_goal_reached.cu:17-188), and the FP32 rollout
(kinematics.cu:123-231) against an FP64 recomputation.

* path optimisation: hand-worked chains on an obstacle-free frame whose
  Hermite connections are straight lines, so every candidate cost is the
  closed form  cost_i + |z_i - z_j|  (all points of a straight connection
  fall in distinct cells; asserted first).  The expected splice is worked
  out by hand from the reference's rules: per i the LAST improving j
  (newCost > 0 and newCost < cost_j, :40-54), the FIRST i of maximal gain
  (maxGain < gain, :128-137), removal of the nodes strictly between, and
  cost_j := newCost (:150-163).  A rule read the other way (best j, first
  j, >= on the gain, <= on the cost, no cost update) gives a different
  chain.  On solved random scenes every round is also checked against a
  Python enumeration of the same rules over the host-truncated connection
  check, and the rounds keep every waypoint footprint-feasible.
* goal window: brute force on small grids enumerates the half-open window
  [trunc(z - r), z + r) x [trunc(x - r), x + r) in Python and takes the
  floor(100 * total) argmin with the documented tie rule (first node in
  cell order, last window cell inside it); known answers pin the window
  edges (fractional radius), the floor (two totals 0.002 apart tie) and
  both tie rules.
* rollout: the oracle's final FP32 pose after the same number of steps
  agrees with a float64 re-run of the recurrence in numpy/libm within
  1e-5 relative (position: relative to the pose's distance from the map
  origin; heading: relative to a full turn), and the end cells agree except
  where the float64 pose lies within that tolerance of a cell edge.

Test infrastructure only (imports the oracle)."""
import ctypes
import math

import numpy as np
import pytest

import _oracle as O
from _parity import OraclePlanner, Params
from fastrrt_amd import scenes

COSTS = scenes.SEGMENTATION_CLASS_COST
LEN = 5.412658773


def _free_frame(W, H):
    f = np.zeros((H, W, 3), np.float32)
    f[:, :, 0] = scenes.FREE_CLASS
    return f


class _Rrt:
    """An oracle planner on a free W x H frame whose tree is written by hand."""

    def __init__(self, W, H, start, goal, dist_tol=10.0):
        self.L = O.lib()
        self.W, self.H = W, H
        self.frame = _free_frame(W, H)
        self.costs = np.ascontiguousarray(COSTS, np.float32)
        self.r = self.L.og_rrt_new(W, H, 34.64, 34.64, 40.0, LEN, -1, -1, -1, -1, -1.0)
        self.L.og_rrt_set_plan_data(self.r, O.fptr(self.frame), O.fptr(self.costs), len(self.costs), start[0],
                                    start[1], start[2], goal[0], goal[1], goal[2], 1.0, 10, 16, -1, 40.0, dist_tol,
                                    float(np.radians(10)))
        self.L.og_rrt_search_init(self.r)
        self.g = self.L.og_rrt_graph(self.r)

    def __del__(self):
        try:
            self.L.og_rrt_free(self.r)
        except Exception:
            pass

    def chain(self, nodes):
        """nodes: [(x, z, heading, cost)], the first is the start (already added)."""
        px, pz = nodes[0][0], nodes[0][1]
        self.L.og_add(self.g, px, pz, nodes[0][2], -1, -1, nodes[0][3])
        for (x, z, h, c) in nodes[1:]:
            self.L.og_add(self.g, x, z, h, px, pz, c)
            px, pz = x, z

    def path(self):
        buf = np.zeros(3 * 4096, np.float32)
        n = self.L.og_rrt_get_planned_path(self.r, O.fptr(buf), 4096)
        return [(int(buf[3 * k]), int(buf[3 * k + 1]), float(buf[3 * k + 2])) for k in range(n)]

    def cost(self, x, z):
        return float(self.L.og_get_cost(self.g, x, z))

    def optimize(self):
        return bool(self.L.og_rrt_path_optimize(self.r))

    def connect(self, a, b):
        """host-truncated connection a -> b (x, z, heading); the cost includes a's cost"""
        return float(self.L.og_check_direct_connection_host(self.g, O.fptr(self.frame),
                                                            float(self.g.contents.max_curv), a[0], a[1], a[2], b[0],
                                                            b[1], b[2], 1))


def _straight_chain(costs, x=32, z0=200, step=20):
    return [(x, z0 - step * k, 0.0, float(c)) for k, c in enumerate(costs)]


def test_straight_connection_closed_form():
    """On a free frame, a connection between two cells of one column with
    both headings 0 is a straight line: cost_i + distance (every point of
    the curve lands in its own cell, the first on the start cell)."""
    r = _Rrt(64, 256, (32, 200, 0.0), (32, 100, 0.0))
    r.chain([(32, 200, 0.0, 7.0)])
    for d in (2, 3, 5, 20, 37, 60, 99):
        assert r.connect((32, 200, 0.0), (32, 200 - d, 0.0)) == 7.0 + d


def test_path_optimize_known_answer_last_improving_j_and_max_gain():
    """Chain P0..P4 at z = 200, 180, 160, 140, 120 (x = 32, heading 0) with
    costs 0, 30, 70, 90, 85.  Candidates (newCost < cost_j):
      i=0: j=2 40<70 (gain 30), j=3 60<90 (30), j=4 80<85 (5) -> last j = 4, gain 5
      i=1: j=3 70<90 (gain 20), j=4 90<85 no                    -> j = 3, gain 20
      i=2: j=4 110<85 no
    Max gain: i=1 -> remove P2, cost(P3) = 70.  Round 2 on P0 P1 P3 P4
    (0, 30, 70, 85): i=0: j=2 60<70 (10), j=3 80<85 (5) -> j=3, gain 5;
    i=1: j=3 90<85 no -> remove P1, P3, cost(P4) = 80.  Round 3: two nodes,
    nothing to splice."""
    r = _Rrt(64, 256, (32, 200, 0.0), (32, 120, 0.0))
    r.chain(_straight_chain([0, 30, 70, 90, 85]))
    assert [p[:2] for p in r.path()] == [(32, 200), (32, 180), (32, 160), (32, 140), (32, 120)]
    assert r.optimize()
    assert [p[:2] for p in r.path()] == [(32, 200), (32, 180), (32, 140), (32, 120)]
    assert [r.cost(32, z) for z in (200, 180, 140, 120)] == [0.0, 30.0, 70.0, 85.0]
    assert r.L.og_get_type(r.g, 32, 160) == O.NULL
    assert r.optimize()
    assert [p[:2] for p in r.path()] == [(32, 200), (32, 120)]
    assert r.cost(32, 120) == 80.0
    assert r.L.og_check_graph_consistent(r.g) == 1
    assert not r.optimize()


def test_path_optimize_ties_and_strict_comparisons():
    """Chains at z = 200, 180, 160, 140 (straight, step 20):
    * costs 0, 10, 45, 60: i=0: j=2 40<45 (gain 5), j=3 60<60 is no
      improvement; i=1: j=3 50<60 (gain 10) -> i=1 wins, remove P2,
      cost(P3) = 50;
    * costs 0, 10, 45, 55: i=0 gain 5 (j=2); i=1: j=3 50<55 gain 5: equal
      gains, the first i wins -> remove P1, cost(P2) = 40;
    * costs 0, 25, 40, 60: every candidate equals the current cost (40, 60,
      65 vs 40, 60, 60): no improvement, false, graph unchanged."""
    r = _Rrt(64, 256, (32, 200, 0.0), (32, 140, 0.0))
    r.chain(_straight_chain([0, 10, 45, 60]))
    assert r.optimize()
    assert [p[:2] for p in r.path()] == [(32, 200), (32, 180), (32, 140)]
    assert [r.cost(32, z) for z in (200, 180, 140)] == [0.0, 10.0, 50.0]
    r = _Rrt(64, 256, (32, 200, 0.0), (32, 140, 0.0))
    r.chain(_straight_chain([0, 10, 45, 55]))
    assert r.optimize()
    assert [p[:2] for p in r.path()] == [(32, 200), (32, 160), (32, 140)]
    assert [r.cost(32, z) for z in (200, 160, 140)] == [0.0, 40.0, 55.0]
    r = _Rrt(64, 256, (32, 200, 0.0), (32, 140, 0.0))
    r.chain(_straight_chain([0, 25, 40, 60]))
    assert not r.optimize()
    assert [p[:2] for p in r.path()] == [(32, 200), (32, 180), (32, 160), (32, 140)]
    assert [r.cost(32, z) for z in (200, 180, 160, 140)] == [0.0, 25.0, 40.0, 60.0]


def _enumerate_round(rr, path):
    """The reference's selection rules restated over the host connection
    check: returns the expected (x, z, cost) chain after one round, or None."""
    n = len(path)
    cost = [rr.cost(x, z) for (x, z, _h) in path]
    best_i, best_gain, best_j, best_nc = -1, -1.0, -1, 0.0
    for i in range(n - 2):
        cand = None
        for j in range(i + 2, n):
            nc = rr.connect(path[i], path[j])
            if nc > 0 and nc < cost[j]:
                cand = (j, nc, np.float32(cost[j]) - np.float32(nc))
        if cand is not None and best_gain < cand[2]:
            best_i, best_j, best_nc, best_gain = i, cand[0], cand[1], cand[2]
    if best_i < 0:
        return None
    keep = list(range(best_i + 1)) + list(range(best_j, n))
    out = []
    for k in keep:
        c = best_nc if k == best_j else cost[k]
        out.append((path[k][0], path[k][1], c))
    return out


def _optimize_rounds_against_enumeration(seed):
    sc = scenes.scene_sized(seed, 160, 0.08, min_dist=(10, 16))
    prm = Params(min_dist=(10, 16), max_path=14.0, dist_tol=6.0)
    o = OraclePlanner(sc, prm, 7 + seed, seed)
    o.search_init()
    for _ in range(150):
        if o.goal_reached() or not o.loop():
            break
    if not o.goal_reached():
        return 0, 0
    rr = _Rrt.__new__(_Rrt)
    rr.L, rr.r, rr.frame = o.L, None, o.frame
    rr.g = o.L.og_rrt_graph(o.r)
    rr.W, rr.H = sc.width, sc.height
    frame = np.ascontiguousarray(sc.frame, np.float32)
    rounds = 0
    for _ in range(12):
        path = [tuple(p) for p in o.path().tolist()]
        path = [(int(x), int(z), float(h)) for (x, z, h) in path]
        expect = _enumerate_round(rr, path)
        got = o.path_optimize()
        assert got == (expect is not None)
        if not got:
            break
        rounds += 1
        st = o.state()
        nodes = sorted((int(x), int(z)) for z, x in np.argwhere(st["type"] == O.NODE))
        assert nodes == sorted((x, z) for (x, z, _c) in expect)
        for k, (x, z, c) in enumerate(expect):
            assert np.float32(st["cost"][z, x]) == np.float32(c)
            if k:
                assert (st["parent_x"][z, x], st["parent_z"][z, x]) == expect[k - 1][:2]
        assert o.L.og_check_graph_consistent(rr.g) == 1
        g = O.Graph(sc.width, sc.height)
        g.search(10, 16)
        g.costs(sc.class_costs)
        for (x, z) in nodes:
            assert o.L.og_feasible(g.g, O.fptr(frame), x, z, float(st["heading"][z, x])) == 1
    return 1, rounds


def test_path_optimize_rounds_match_enumeration():
    solved = rounds = 0
    for seed in (7, 11, 19, 23, 29):
        a, b = _optimize_rounds_against_enumeration(seed)
        solved += a
        rounds += b
    print(f"[pins] path_optimize: {rounds} rounds on {solved} solved scenes match the enumeration", flush=True)
    assert solved >= 4 and rounds >= 4


# ---------------------------------------------------------------------------
# goal window argmin
# ---------------------------------------------------------------------------

def _window_graph(W, H, max_curv=-1.0):
    g = O.Graph(W, H)
    g.physical(34.64, 34.64, 40.0, LEN, max_curv)
    g.search(2, 2)
    g.costs(COSTS)
    return g


def _set_gc(g, x, z, heading, cost):
    s = g.g.contents
    pos = z * g.W + x
    s.gc_flag[pos] = 1.0
    s.gc_heading[pos] = heading
    s.gc_cost[pos] = cost


def _brute_force_window(g, frame, radius):
    """Independent enumeration of graph_check_goal_reached.cu:17-139."""
    L = O.lib()
    s = g.g.contents
    W, H = g.W, g.H
    typ = np.ctypeslib.as_array(s.type, shape=(H * W,))
    head = np.ctypeslib.as_array(s.heading, shape=(H * W,))
    flag = np.ctypeslib.as_array(s.gc_flag, shape=(H * W,))
    gch = np.ctypeslib.as_array(s.gc_heading, shape=(H * W,))
    gcc = np.ctypeslib.as_array(s.gc_cost, shape=(H * W,))
    r = np.float32(radius)
    best = None
    for pos in range(W * H):
        if typ[pos] != O.NODE:
            continue
        z, x = divmod(pos, W)
        zp = int(np.float32(z) - r)          # int zp = z - searchRadius (truncation)
        while np.float32(zp) < np.float32(z) + r:
            xp = int(np.float32(x) - r)
            while np.float32(xp) < np.float32(x) + r:
                if 0 <= zp < H and 0 <= xp < W and flag[zp * W + xp] > 0:
                    q = zp * W + xp
                    c = L.og_check_direct_connection(g.g, O.fptr(frame), float(s.max_curv), x, z, float(head[pos]),
                                                     xp, zp, float(gch[q]), 1, 0)
                    if c >= 0:
                        total = np.float32(np.float32(c) + gcc[q])
                        key = math.floor(float(np.float32(np.float32(100.0) * total)))
                        # lowest key; equal key: first node, last window cell of it
                        if best is None or key < best[0] or (key == best[0] and pos == best[1]):
                            best = (key, pos, xp, zp, float(gch[q]), float(c), float(total))
                xp += 1
            zp += 1
    return best


def _oracle_window(g, frame, radius):
    L = O.lib()
    ok = L.og_find_best_goal_direct_connection(g.g, O.fptr(frame), radius, 1)
    s = g.g.contents
    return ok, list(s.best_parent), list(s.best_child)


@pytest.mark.parametrize("seed", range(6))
def test_goal_window_brute_force(seed):
    rng = np.random.default_rng(500 + seed)
    W, H = 40, 36
    g = _window_graph(W, H)
    frame = _free_frame(W, H)
    obst = rng.random((H, W)) < 0.04
    frame[:, :, 0][obst] = scenes.OBSTACLE_CLASS
    for _ in range(8):
        x, z = int(rng.integers(0, W)), int(rng.integers(0, H))
        g.add(x, z, float(rng.uniform(-1.0, 1.0)), -1, -1, float(rng.integers(0, 30)))
    for _ in range(60):
        x, z = int(rng.integers(0, W)), int(rng.integers(0, H))
        # integer-ish costs make equal keys (ties) common
        _set_gc(g, x, z, float(rng.uniform(-0.8, 0.8)), float(rng.integers(0, 12)) + float(rng.choice([0.0, 0.004])))
    for radius in (3.0, 6.5, 9.0):
        exp = _brute_force_window(g, frame, radius)
        ok, parent, child = _oracle_window(g, frame, radius)
        if exp is None:
            assert ok == 0
            continue
        key, pos, xp, zp, lih, c, total = exp
        assert ok == 1
        assert (int(parent[0]), int(parent[1])) == (pos % W, pos // W), (radius, exp, parent, child)
        assert (int(child[0]), int(child[1])) == (xp, zp)
        assert np.float32(parent[3]) == np.float32(c) and np.float32(child[3]) == np.float32(total)
        assert np.float32(child[2]) == np.float32(lih)


def test_goal_window_known_answers():
    """Node at (10, 20) heading pi/2 (tangent +x) on a free 48 x 40 frame:
    connections to cells of row 20 ahead of it are straight, cost = node
    cost + distance."""
    W, H = 48, 40
    frame = _free_frame(W, H)
    hp = float(np.float32(math.pi / 2))
    # half-open window: with r = 10 the cell at x + r = 20 is outside,
    # with r = 10.5 it is inside (20 < 20.5)
    g = _window_graph(W, H)
    g.add(10, 20, hp, -1, -1, 5.0)
    _set_gc(g, 20, 20, hp, -1000.0)
    _set_gc(g, 15, 20, hp, 50.0)
    ok, parent, child = _oracle_window(g, frame, 10.0)
    assert ok == 1 and (child[0], child[1]) == (15.0, 20.0) and child[3] == 5.0 + 5.0 + 50.0
    ok, parent, child = _oracle_window(g, frame, 10.5)
    assert ok == 1 and (child[0], child[1]) == (20.0, 20.0) and child[3] == 5.0 + 10.0 - 1000.0
    # the same along z: heading pi is the tangent +z; z + r is outside
    # with r = 10 and inside with r = 10.5
    hpi = float(np.float32(math.pi))
    g = _window_graph(W, H)
    g.add(20, 10, hpi, -1, -1, 5.0)
    _set_gc(g, 20, 20, hpi, -1000.0)
    _set_gc(g, 20, 16, hpi, 50.0)
    ok, parent, child = _oracle_window(g, frame, 10.0)
    assert ok == 1 and (child[0], child[1]) == (20.0, 16.0) and child[3] == 5.0 + 6.0 + 50.0
    ok, parent, child = _oracle_window(g, frame, 10.5)
    assert ok == 1 and (child[0], child[1]) == (20.0, 20.0)
    # the lower edge is the truncation of x - r: heading -pi/2 is the tangent
    # -x; from x = 30, r = 9.5 starts the window at trunc(20.5) = 20 (the
    # cell at x = 20 is inside although 20 < x - r), r = 9 starts it at 21
    hm = float(np.float32(-math.pi / 2))
    g = _window_graph(W, H)
    g.add(30, 20, hm, -1, -1, 5.0)
    _set_gc(g, 20, 20, hm, -1000.0)
    _set_gc(g, 25, 20, hm, 50.0)
    ok, parent, child = _oracle_window(g, frame, 9.5)
    assert ok == 1 and (child[0], child[1]) == (20.0, 20.0)
    ok, parent, child = _oracle_window(g, frame, 9.0)
    assert ok == 1 and (child[0], child[1]) == (25.0, 20.0)
    # floor(100 * total): totals 60.000 and 60.006 share key 6000 (rounding
    # would give 6000 and 6001), so the tie rule (last window cell in scan
    # order) picks the later cell although its float total is larger
    g = _window_graph(W, H)
    g.add(10, 20, hp, -1, -1, 5.0)
    _set_gc(g, 13, 20, hp, 52.0)        # 5 + 3 + 52 = 60.000
    _set_gc(g, 15, 20, hp, 50.006)      # 5 + 5 + 50.006 = 60.006
    ok, parent, child = _oracle_window(g, frame, 10.0)
    assert ok == 1 and (child[0], child[1]) == (15.0, 20.0)
    assert math.floor(100 * np.float32(child[3])) == 6000 and child[3] > 60.005
    # ties across nodes: the first node in cell order wins
    g = _window_graph(W, H)
    g.add(10, 20, hp, -1, -1, 5.0)      # cell 20*48+10
    g.add(10, 10, hp, -1, -1, 5.0)      # cell 10*48+10: earlier in cell order
    _set_gc(g, 16, 20, hp, 50.0)        # from (10,20): 5 + 6 + 50 = 61
    _set_gc(g, 16, 10, hp, 50.0)        # from (10,10): 5 + 6 + 50 = 61
    ok, parent, child = _oracle_window(g, frame, 8.0)
    assert ok == 1 and (parent[0], parent[1]) == (10.0, 10.0) and (child[0], child[1]) == (16.0, 10.0)
    # no flagged cell in any window: false
    g = _window_graph(W, H)
    g.add(10, 20, hp, -1, -1, 5.0)
    _set_gc(g, 40, 5, hp, 1.0)
    ok, _, _ = _oracle_window(g, frame, 10.0)
    assert ok == 0


# ---------------------------------------------------------------------------
# rollout against a float64 recomputation
# ---------------------------------------------------------------------------

def _rollout_fp64(sx, sz, h0, steer, steps, lr, max_steer, og=(128.0, 128.0)):
    x = float(np.float32(og[1] - sz))
    y = float(np.float32(sx - og[0]))
    st = max(-max_steer, min(max_steer, steer))
    t = math.tan(st)
    b = math.atan(t / 2.0)
    ds = 0.1
    hif = ds * math.cos(b) * t / (2.0 * lr)
    h = float(h0)
    for _ in range(steps):
        x += ds * math.cos(h + b)
        y += ds * math.sin(h + b)
        h += hif
    return x, y, h


def test_rollout_fp32_within_1e5_of_fp64():
    L = O.lib()
    g = O.Graph(256, 256)
    g.physical(256, 256, 40.0, LEN)
    g.costs([0, 1, 2, 3, 4, -1])
    g.search(0, 0)
    frame = O.empty_frame(256, 256)
    lr = float(np.float32(LEN) / np.float32(2.0))
    max_steer = 40.0 * math.pi / 180.0
    out = (ctypes.c_float * 4)()
    pose = (ctypes.c_double * 4)()
    rng = np.random.default_rng(77)
    n, over_rel_h, flips, worst_pos, worst_h = 0, 0, 0, 0.0, 0.0
    for _ in range(2000):
        h0 = np.float32(rng.uniform(-math.pi, math.pi))
        steer = np.float32(rng.uniform(-max_steer, max_steer))
        psize = float(rng.integers(5, 41))
        g.add(100, 200, float(h0), -1, -1, 0)
        L.og_kinematic_new_path_pose(g.g, O.fptr(frame), 100, 200, float(steer), psize, 1.0, out, pose)
        L.og_clear(g.g)
        steps = int(pose[3])
        x, y, h = _rollout_fp64(100, 200, h0, float(steer), steps, lr, max_steer)
        rel_pos = math.hypot(pose[0] - x, pose[1] - y) / math.hypot(x, y)
        rel_h = abs(pose[2] - h) / (2.0 * math.pi)
        worst_pos, worst_h = max(worst_pos, rel_pos), max(worst_h, rel_h)
        assert rel_pos <= 1e-5, (h0, steer, psize, rel_pos)
        assert rel_h <= 1e-5, (h0, steer, psize, rel_h)
        over_rel_h += abs(pose[2] - h) > 1e-5 * max(abs(h), abs(float(h0)))
        # end cell (kinematics.cu:25-46): FP32 vs FP64 pose; a flip is allowed
        # only where the FP64 coordinate is within the tolerance of an edge
        tol = 1e-5 * math.hypot(x, y)
        if out[0] >= 0:
            cx64, cz64 = math.trunc(128.0 + y), math.trunc(128.0 - x)
            if (cx64, cz64) != (int(out[0]), int(out[1])):
                near = min(abs(128.0 + y - round(128.0 + y)), abs(128.0 - x - round(128.0 - x)))
                assert near <= tol, ((h0, steer, psize), (cx64, cz64), tuple(out))
                flips += 1
        n += 1
    print(f"[fp64] {n} rollouts: worst position {worst_pos:.3g} (relative to |pose|), worst heading "
          f"{worst_h:.3g} (relative to 2 pi); {over_rel_h} headings beyond 1e-5 relative to |h|; "
          f"{flips} end-cell flips at cell edges")


# ---------------------------------------------------------------------------
# path feasibility: check_kinematic_connection_start_end (kinematics.cu:233-306)
# ---------------------------------------------------------------------------

def _feasible(g, frame, s, e, v):
    fh = ctypes.c_double()
    pc = ctypes.c_float()
    r = O.lib().og_check_feasible_connection_ex(g.g, O.fptr(frame), s[0], s[1], e[0], e[1], v, ctypes.byref(fh),
                                                ctypes.byref(pc))
    return r, fh.value, pc.value


def test_feasible_connection_closed_forms():
    """Heading 0 points to -z.  An end straight ahead gives path heading =
    heading, so the steering stays 0 and the walk is the straight column: it
    reaches the end after distance / ds steps without moving away (true,
    out parameters untouched).  An obstacle row across that column rejects
    it.  An end behind the start: the first new cell is farther from the end
    than the start was, so the walk stops there (false: 10 px > 2 px), with
    cost 1 and the heading after that step."""
    W, H = 256, 256
    frame = O.empty_frame(W, H, cls=1.0)
    g = O.Graph(W, H)
    g.physical(256, 256, 40.0, LEN)
    g.costs([0, 0, 2, 3, 4, -1])
    g.search(4, 4)
    g.add(128, 200, 0.0, -1, -1, 0)
    assert _feasible(g, frame, (128, 200), (128, 150), 1.0) == (1, 0.0, 0.0)
    assert _feasible(g, frame, (128, 200), (128, 199), 1.0) == (1, 0.0, 0.0)
    blocked = frame.copy()
    blocked[170, :, 0] = 5.0
    assert _feasible(g, blocked, (128, 200), (128, 150), 1.0)[0] == 0
    r, fh, pc = _feasible(g, frame, (128, 200), (128, 210), 1.0)
    assert r == 0 and pc == 1.0 and abs(fh) <= 0.1 * math.tan(math.radians(40)) / LEN * 20
    # a start outside the grid answers false
    assert _feasible(g, frame, (-1, 200), (128, 150), 1.0)[0] == 0


def test_feasible_connection_gtest():
    """unittests/src/test_kinematics.cpp:51-94: a node derived at 20 deg,
    P = 50, v = 2 from (128, 128, 0) is reachable from the start."""
    L = O.lib()
    frame = O.empty_frame(256, 256)
    g = O.Graph(256, 256)
    g.physical(256, 256, 40.0, LEN)
    g.costs([0, 1, 2, 3, 4, 5])
    g.search(0, 0)
    g.add(128, 128, 0.0, -1, -1, 0)
    out = (ctypes.c_int * 2)()
    L.og_derivate_node(g.g, O.fptr(frame), float(np.float32(math.radians(20))), 50.0, 2.0, 128, 128, out)
    assert out[0] > 128 and out[1] < 128
    assert _feasible(g, frame, (128, 128), (out[0], out[1]), 2.0)[0] == 1


def test_goal_curve_sample_certificate_bound():
    """The centre pass's certificate from 64 samples of a goal curve of n > 64
    points (sample l at point l * s, s = ceil(n / 64)) relies on every point
    lying within delta = M g / (n - 1) of a sample, with M = |A| + |B| + |C|
    bounding |P'| of the Hermite curve P' = A t^2 + B t + C and g = max(s / 2
    + 1, n - 1 - last sample) the index gap.  Checked on the curves of random
    cells to random goals (tangents of length `distance`, headings from the
    cell and at the goal), points evaluated in FP64 from the Hermite basis."""
    rng = np.random.default_rng(31)
    worst = 0.0
    for _ in range(400):
        x, z = rng.integers(0, 1024, 2)
        gx, gz = rng.integers(0, 1024, 2)
        dist = math.sqrt(float((gx - x) ** 2 + (gz - z) ** 2))
        n = int(dist)
        if n <= 64:
            continue
        h1, h2 = rng.uniform(-math.pi, math.pi, 2)
        p0, p1 = np.array([x, z], float), np.array([gx, gz], float)
        m0 = dist * np.array([math.cos(h1), math.sin(h1)])
        m1 = dist * np.array([math.cos(h2), math.sin(h2)])
        t = np.arange(n) / (n - 1)
        t2, t3 = t * t, t * t * t
        pts = (np.outer(2 * t3 - 3 * t2 + 1, p0) + np.outer(t3 - 2 * t2 + t, m0) + np.outer(-2 * t3 + 3 * t2, p1)
               + np.outer(t3 - t2, m1))
        A = 6 * (p0 - p1) + 3 * (m0 + m1)
        B = 6 * (p1 - p0) - 4 * m0 - 2 * m1
        M = np.linalg.norm(A) + np.linalg.norm(B) + np.linalg.norm(m0)
        s = (n + 63) // 64
        last = s * min((n - 1) // s, 63)
        g = max(s // 2 + 1, n - 1 - last)
        delta = M * g / (n - 1)
        samples = pts[np.arange(0, last + 1, s)]
        d = np.min(np.linalg.norm(pts[:, None, :] - samples[None, :, :], axis=2), axis=1)
        assert d.max() <= delta, (x, z, gx, gz, d.max(), delta)
        worst = max(worst, float(d.max() / delta))
    assert worst > 0.05  # the bound is exercised, not vacuous
