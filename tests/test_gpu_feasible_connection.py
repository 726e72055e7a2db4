"""Path feasibility on the GPU: CudaGraph::checkFeasibleConnection
(kinematics.cu:308-326) -> check_kinematic_connection_start_end (:233-306)
through the graph ABI (cudagraph_check_feasible_connection[s]), against the
oracle.

The walk is FP64 with the device's double-precision libm on one side and
glibc's on the other, so the verdicts and path costs (cell counts plus
intrinsic costs) must agree exactly and the final headings within 1e-12
relative."""
import ctypes
import math

import numpy as np
import pytest

import _oracle as O
from fastrrt_amd import pyfastrrt as P
from fastrrt_amd import scenes

pytestmark = pytest.mark.gpu

LEN = 5.412658773


def _pair(W, H, costs, min_dist, ego=((-1, -1), (-1, -1))):
    lb, ub = ego
    g = P.CudaGraph(W, H, 34.64, 34.64, 40, LEN, min_dist[0], min_dist[1], lb[0], lb[1], ub[0], ub[1],
                    np.asarray(costs, np.float32))
    o = O.Graph(W, H)
    o.physical(34.64, 34.64, 40.0, LEN)
    o.search(min_dist[0], min_dist[1], lb[0], lb[1], ub[0], ub[1])
    o.costs(costs)
    return g, o


def test_gtest_check_kinematic_path():
    """unittests/src/test_kinematics.cpp:51-94 through the graph ABI."""
    frame = P.SearchFrame(O.empty_frame(256, 256), [0, 1, 2, 3, 4, 5])
    g = P.CudaGraph(256, 256, 256, 256, 40, LEN, 0, 0, -1, -1, -1, -1, np.array([0, 1, 2, 3, 4, 5]))
    g.add(128, 128, P.angle(0.0), -1, -1, 0)
    node = g.derivate_node_end(frame, float(np.float32(math.radians(20))), 50.0, 2.0, 128, 128)
    assert node[0] > 128 and node[1] < 128
    assert g.check_feasible_connection(frame, (128, 128), node, 2)


@pytest.mark.parametrize("cov,seed,velocity,apf", [(0.10, 1, 1, False), (0.20, 2, 2, True), (0.05, 3, 1, True)])
def test_random_queries_match_oracle(cov, seed, velocity, apf):
    sc = scenes.scene_256(100 + seed, coverage=cov)
    md = (10, 16)
    g, o = _pair(256, 256, sc.class_costs, md, (sc.ego_lb, sc.ego_ub))
    frame = P.SearchFrame(sc.frame, sc.class_costs)
    fr = np.ascontiguousarray(sc.frame, np.float32)
    if apf:
        g.compute_apf_repulsion(frame, 1.5, 12)
        o.apf_repulsion(fr, 1.5, 12)
    rng = np.random.default_rng(seed)
    starts = []
    for _ in range(40):
        x, z = int(rng.integers(0, 256)), int(rng.integers(0, 256))
        h = float(np.float32(rng.uniform(-math.pi, math.pi)))
        g.add(x, z, P.angle(h), -1, -1, 0)
        o.add(x, z, h, -1, -1, 0)
        starts.append((x, z))
    q = []
    for _ in range(1500):
        sx, sz = starts[int(rng.integers(0, len(starts)))]
        ex = int(np.clip(sx + rng.integers(-70, 71), -5, 260))
        ez = int(np.clip(sz + rng.integers(-70, 71), -5, 260))
        q.append((sx, sz, ex, ez))
    q.append((-3, 10, 20, 20))     # start outside the grid
    q.append((starts[0][0], starts[0][1], starts[0][0], starts[0][1]))  # zero distance
    print(f"[feasible] cov={cov} v={velocity} apf={apf}: {len(q)} queries", flush=True)
    ok, fh, pc = g.check_feasible_connections(frame, q, velocity)
    L = O.lib()
    n_true = n_stop = 0
    for k, (sx, sz, ex, ez) in enumerate(q):
        h = ctypes.c_double()
        c = ctypes.c_float()
        r = L.og_check_feasible_connection_ex(o.g, O.fptr(fr), sx, sz, ex, ez, float(velocity), ctypes.byref(h),
                                              ctypes.byref(c))
        assert bool(r) == bool(ok[k]), (k, q[k], r, ok[k])
        assert np.float32(c.value) == pc[k], (k, q[k], c.value, pc[k])
        assert abs(h.value - fh[k]) <= 1e-12 * max(1.0, abs(h.value)), (k, q[k], h.value, fh[k])
        n_true += bool(r)
        n_stop += c.value > 0
    print(f"[feasible] {n_true} feasible, {n_stop} stopped on moving away", flush=True)
    assert 0 < n_true < len(q) and n_stop > 0
    # the single-query entry point agrees
    for k in range(0, len(q), 97):
        assert g.check_feasible_connection(frame, q[k][:2], q[k][2:], velocity) == bool(ok[k])
