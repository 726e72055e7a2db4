"""The graph ABI's Hermite hook can_connect_to_goal_using_hermite
(python_cuda_graph_bind.cpp:147-153 -> graph_expansion.cu:246-261 ->
graph_hermite_connection.cu:11-136) on the GPU: the reference's gtests
(test_graph_can_connect_hermite.cpp:31-109) and the curve's value bit for bit
against the CPU oracle."""
import numpy as np
import pytest

import _oracle as O
from fastrrt_amd import pyfastrrt as P
from fastrrt_amd import scenes

pytestmark = pytest.mark.gpu


def _dist_channel(frame, gx, gz):
    H, W = frame.shape[:2]
    zz, xx = np.mgrid[0:H, 0:W]
    frame[:, :, 1] = np.sqrt(((gx - xx) ** 2 + (gz - zz) ** 2).astype(np.float32))


def test_frame_distance_to_goal_channel():
    """process_distance_to_goal fills channel 1 with float32 sqrt of the
    integer squared distance and leaves channels 0 and 2 alone."""
    rng = np.random.default_rng(4)
    px = rng.uniform(0, 30, (70, 90, 3)).astype(np.float32)
    frame = P.SearchFrame(px, [0, 1])
    frame.process_distance_to_goal(33, -5)
    got = frame.read()
    want = px.copy()
    _dist_channel(want, 33, -5)
    assert np.array_equal(got.view(np.int32), want.view(np.int32))


def test_reference_hermite_gtests():
    costs_ok = [0, 1, 2, 3, 4, 5]
    costs_blocked = [-1, 1, 2, 3, 4, 5]
    g = P.CudaGraph(256, 256, 256, 256, 40, 5.412658773, 0, 0, -1, -1, -1, -1, np.array(costs_ok))
    g.add(128, 128, P.angle(0.0), -1, -1, 0)
    # StraightConnectionSuccess, with the frame processed as the gtest does
    fp = P.SearchFrame(O.empty_frame(256, 256), costs_ok)
    fp.process_distance_to_goal(128, 0)
    assert g.can_connect_to_goal_using_hermite(fp, 128, 128, 128, 0, 0.0)
    assert g.can_connect_hermite_value(fp, 128, 128, 128, 0, 0.0) == (False, 127.0)
    f = O.empty_frame(256, 256)
    _dist_channel(f, 128, 0)
    assert g.can_connect_to_goal_using_hermite(P.SearchFrame(f, costs_ok), 128, 128, 128, 0, 0.0)
    assert g.can_connect_hermite_value(P.SearchFrame(f, costs_ok), 128, 128, 128, 0, 0.0) == (False, 127.0)
    # StraightConnectionFail (processed) and NoPreProcessFail: the goal pixel is an obstacle
    assert not g.can_connect_to_goal_using_hermite(P.SearchFrame(f, costs_blocked), 128, 128, 128, 0, 0.0)
    f0 = O.empty_frame(256, 256)
    assert not g.can_connect_to_goal_using_hermite(P.SearchFrame(f0, costs_blocked), 128, 128, 128, 0, 0.0)
    # a free goal without the distance channel: no points, -1, which converts to true
    assert g.can_connect_hermite_value(P.SearchFrame(f0, costs_ok), 128, 128, 128, 0, 0.0) == (False, -1.0)
    assert g.can_connect_to_goal_using_hermite(P.SearchFrame(f0, costs_ok), 128, 128, 128, 0, 0.0)


@pytest.mark.parametrize("min_dist", [(10, 16), (22, 40)])
def test_hermite_hook_matches_oracle(min_dist):
    sc = scenes.scene_256(140 + min_dist[0], coverage=0.05, min_dist=min_dist)
    f = np.ascontiguousarray(sc.frame).copy()
    gx, gz = sc.goal[0], sc.goal[1]
    _dist_channel(f, gx, gz)
    frame = P.SearchFrame(f, sc.class_costs)
    g = P.CudaGraph(256, 256, sc.perception_m, sc.perception_m, 40, 5.412658773, min_dist[0], min_dist[1], -1, -1,
                    -1, -1, np.asarray(sc.class_costs, np.float32))
    g.compute_apf_repulsion(frame, 0.5, 12)
    o = O.Graph(256, 256)
    o.physical(sc.perception_m, sc.perception_m, 40.0, 5.412658773)
    o.search(min_dist[0], min_dist[1])
    o.costs(sc.class_costs)
    o.apf_repulsion(f, 0.5, 12)
    bound = float(np.float32(np.radians(40.0)))
    rng = np.random.default_rng(min_dist[0])
    seen = {}
    print(f"[hermite] min_dist={min_dist}", flush=True)
    for k in range(160):
        if k % 2:
            x, z = int(rng.integers(-2, 258)), int(rng.integers(20, 258))
            h = float(np.float32(rng.uniform(-1.2, 1.2)))
            gh = float(rng.uniform(-1.9, 1.9))
        else:  # nearly straight curves below the goal: the ones that can pass the bound
            x, z = gx + int(rng.integers(-4, 5)), int(rng.integers(60, 256))
            h = float(np.float32(rng.uniform(-0.03, 0.03)))
            gh = float(rng.uniform(0.0, 0.99))
        tx, tz = (gx, gz) if k % 8 else (int(rng.integers(-3, 259)), int(rng.integers(-3, 40)))
        if 0 <= x < 256 and 0 <= z < 256:
            g.add(x, z, P.angle(h), -1, -1, float(k))
            o.add(x, z, h, -1, -1, float(k))
        got = g.can_connect_hermite_value(frame, x, z, tx, tz, gh)
        blocked = O.ctypes.c_int(0)
        want = o.L.og_can_connect_hermite(o.g, O.fptr(f), bound, x, z, tx, tz, int(gh), O.ctypes.byref(blocked))
        assert got[0] == bool(blocked.value), (k, x, z, tx, tz)
        if not got[0]:
            assert np.float32(got[1]).view(np.int32) == np.float32(want).view(np.int32), (k, x, z, got, want)
        key = "blocked" if got[0] else ("cost" if got[1] >= 0 else str(int(got[1])))
        seen[key] = seen.get(key, 0) + 1
    print(f"[hermite] outcomes {seen}", flush=True)
    assert seen.get("cost", 0) >= 5 and seen.get("-1", 0) >= 5 and seen.get("-2", 0) >= 5
    assert seen.get("blocked", 0) >= 1
