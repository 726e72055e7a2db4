"""The oracle's canConnectToGoal / cudaHermite restatement
(graph_expansion.cu:246-261, graph_hermite_connection.cu:11-136), pinned by
the reference's three gtests (test_graph_can_connect_hermite.cpp:31-109),
the closed form of a straight connection, the pinned direct-connection
check on the same curves, and the curvature measure derived numerically
from the curve itself."""
import math

import numpy as np

import _oracle as O

MAX_STEER = np.float32(np.radians(40.0))  # canConnectToGoal passes the max steering angle as the bound


def _graph(W=256, H=256, costs=(0, 1, 2, 3, 4, 5), md=(0, 0)):
    g = O.Graph(W, H)
    g.physical(256, 256, 40.0, 5.412658773)
    g.costs(list(costs))
    g.search(md[0], md[1])
    return g


def _dist_channel(frame, gx, gz):
    """Distance-to-goal channel: Euclidean pixel distance to the goal."""
    H, W = frame.shape[:2]
    zz, xx = np.mgrid[0:H, 0:W]
    frame[:, :, 1] = np.sqrt(((gx - xx) ** 2 + (gz - zz) ** 2).astype(np.float32))


def _hook(g, frame, x, z, gx, gz, gh, max_curv=MAX_STEER):
    blocked = O.ctypes.c_int(0)
    r = g.L.og_can_connect_hermite(g.g, O.fptr(frame), float(max_curv), x, z, gx, gz, int(gh),
                                   O.ctypes.byref(blocked))
    return bool(blocked.value), float(r)


def test_straight_connection_success():
    """StraightConnectionSuccess: node (128, 128) at 0 rad, goal (128, 0) at 0,
    distance channel processed: the curve is the straight segment, one new
    cell per sample after the node's own, so the cost is 127 cells."""
    g = _graph()
    g.add(128, 128, 0.0, -1, -1, 0)
    f = O.empty_frame(256, 256)
    _dist_channel(f, 128, 0)
    assert _hook(g, f, 128, 128, 128, 0, 0) == (False, 127.0)


def test_goal_obstacle_answers_false():
    """StraightConnectionFail and NoPreProcessFail: class 0 costs -1, so the
    goal pixel is an obstacle and the answer is false before any curve."""
    g = _graph(costs=(-1, 1, 2, 3, 4, 5))
    g.add(128, 128, 0.0, -1, -1, 0)
    f = O.empty_frame(256, 256)
    assert _hook(g, f, 128, 128, 128, 0, 0)[0]
    _dist_channel(f, 128, 0)
    assert _hook(g, f, 128, 128, 128, 0, 0)[0]


def test_unprocessed_distance_walks_nothing():
    """A free goal with the distance channel left at 0: no curve points, -1
    (which the reference's bool conversion turns into true)."""
    g = _graph()
    g.add(128, 128, 0.0, -1, -1, 0)
    assert _hook(g, O.empty_frame(256, 256), 128, 128, 128, 0, 0) == (False, -1.0)


def test_matches_direct_connection_without_curvature():
    """With no curvature bound, a curve the pinned direct-connection check
    accepts costs the same here less the node's own cell (this walk starts
    from it); a curve it rejects for a collision is rejected here too unless
    only the node's own cell or the endpoint test failed there."""
    rng = np.random.default_rng(21)
    W = H = 128
    f = np.zeros((H, W, 3), np.float32)
    f[:, :, 0] = 1
    for _ in range(25):
        x0, z0 = rng.integers(0, W - 8), rng.integers(0, H - 8)
        f[z0:z0 + rng.integers(2, 8), x0:x0 + rng.integers(2, 8), 0] = 0
    costs = (-1, 0)
    gx, gz = 64, 2
    f[gz, gx, 0] = 1
    _dist_channel(f, gx, gz)
    g = _graph(W, H, costs, md=(6, 10))
    g.plane("intrinsic")[:] = rng.uniform(0, 0.5, (H, W)).astype(np.float32)
    agree = 0
    for _ in range(300):
        x, z = int(rng.integers(0, W)), int(rng.integers(40, H))
        if f[z, x, 0] == 0:
            continue
        h = float(np.float32(rng.uniform(-0.6, 0.6)))
        gh = int(rng.integers(-1, 2))
        g.add(x, z, h, -1, -1, float(np.float32(rng.uniform(0, 5))))
        d = g.L.og_check_direct_connection(g.g, O.fptr(f), 0.0, x, z, h, gx, gz, float(gh), 0, 1)
        blocked, r = _hook(g, f, x, z, gx, gz, gh, 0.0)
        assert not blocked
        own = np.float32(g.plane("intrinsic")[z, x]) + np.float32(1)
        if d >= 0:
            assert r >= 0
            # the direct check adds the node's own cell first: compare the sums from the node's cost
            base = np.float32(g.plane("cost")[z, x])
            assert math.isclose(r - base, (d - base) - own, rel_tol=0, abs_tol=1e-3), (x, z, r, d)
            agree += 1
        g.L.og_clear(g.g)
    assert agree >= 20


def _numeric_curvature_measure(x, z, h, gx, gz, gh, d, n, power=2):
    """max over the sample points of |P' x P''| / |P'|^power (power 2: the
    reference's pow(v, 3/2) with integer 3/2; power 3: the curvature), from
    finite differences of the curve."""
    a1, a2 = h - math.pi / 2, gh - math.pi / 2
    t1 = np.array([d * math.cos(a1), d * math.sin(a1)])
    t2 = np.array([d * math.cos(a2), d * math.sin(a2)])
    p0, p1 = np.array([x, z], float), np.array([gx, gz], float)

    def P(t):
        return ((2 * t ** 3 - 3 * t ** 2 + 1)[:, None] * p0 + (t ** 3 - 2 * t ** 2 + t)[:, None] * t1 +
                (-2 * t ** 3 + 3 * t ** 2)[:, None] * p1 + (t ** 3 - t ** 2)[:, None] * t2)

    t = np.linspace(0, 1, n)
    e = 1e-4
    d1 = (P(t + e) - P(t - e)) / (2 * e)
    d2 = (P(t + e) - 2 * P(t) + P(t - e)) / (e * e)
    cr = np.abs(d1[:, 0] * d2[:, 1] - d1[:, 1] * d2[:, 0])
    return (cr / np.sqrt((d1 ** 2).sum(1)) ** power)[1:].max()


def test_curvature_bound_is_cross_over_squared_speed():
    """Sideways goals on a free frame: the walk fails with -2 exactly when the
    numerically differentiated measure exceeds the bound (cases within 5% of
    the bound are skipped); the true curvature |P' x P''| / |P'|^3 would let
    every one of these gentle curves pass."""
    W = H = 256
    f = O.empty_frame(W, H)
    g = _graph()
    n_fail = n_pass = 0
    for off in range(-60, 61, 6):
        gx, gz = 128 + off, 10
        _dist_channel(f, gx, gz)
        x, z = 128, 200
        g.add(x, z, 0.0, -1, -1, 0)
        d = float(f[z, x, 1])
        m = _numeric_curvature_measure(x, z, 0.0, gx, gz, 0, d, int(d))
        if abs(m - MAX_STEER) < 0.05 * MAX_STEER:
            continue
        blocked, r = _hook(g, f, x, z, gx, gz, 0)
        assert not blocked
        if m > MAX_STEER:
            assert r == -2.0, (off, m, r)
            assert _numeric_curvature_measure(x, z, 0.0, gx, gz, 0, d, int(d), power=3) < MAX_STEER
            n_fail += 1
        else:
            assert r > 0, (off, m, r)
            n_pass += 1
        g.L.og_clear(g.g)
    assert n_fail >= 3 and n_pass >= 3
