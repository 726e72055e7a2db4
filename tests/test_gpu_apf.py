"""Potential-field intrinsic costs (src/cuda/apf.cu) on the GPU through the
graph ABI (compute_apf_repulsion / compute_apf_attraction /
get_intrinsic_costs), bit for bit against the CPU oracle, and their effect
on the planner's costs (a rollout through the graph ABI adds intrinsic + 1
per cell, kinematics.cu:206)."""
import numpy as np
import pytest

import _oracle as O
from fastrrt_amd import pyfastrrt as P
from fastrrt_amd import scenes

pytestmark = pytest.mark.gpu


def _pair(W, H, costs, ego, min_dist=(22, 40)):
    lb, ub = ego
    g = P.CudaGraph(W, H, 34.641016151377535, 34.641016151377535, 40, 5.412658773, min_dist[0], min_dist[1], lb[0],
                    lb[1], ub[0], ub[1], np.asarray(costs, np.float32))
    o = O.Graph(W, H)
    o.physical(34.641016151377535, 34.641016151377535, 40.0, 5.412658773)
    o.search(min_dist[0], min_dist[1], lb[0], lb[1], ub[0], ub[1])
    o.costs(costs)
    return g, o


def _same(g, o, what):
    a = g.get_intrinsic_costs()
    b = o.plane("intrinsic")
    bad = np.argwhere(a.view(np.int32) != b.view(np.int32))
    assert bad.shape[0] == 0, f"{what}: {bad.shape[0]} cells differ, first (z, x) {bad[:3].tolist()}"


@pytest.mark.parametrize("radius", [1, 5, 17, 50])
def test_scene_repulsion_then_attraction(radius):
    sc = scenes.scene_256(31 + radius, coverage=0.15)
    print(f"[apf] 256x256 radius={radius}", flush=True)
    frame = P.SearchFrame(sc.frame, sc.class_costs)
    g, o = _pair(256, 256, sc.class_costs, (sc.ego_lb, sc.ego_ub))
    g.compute_apf_repulsion(frame, 2.0, radius)
    o.apf_repulsion(sc.frame, 2.0, radius)
    _same(g, o, "repulsion")
    g.compute_apf_attraction(frame, 0.05, sc.goal[0], sc.goal[1])
    o.apf_attraction(sc.frame, 0.05, sc.goal[0], sc.goal[1])
    _same(g, o, "attraction")
    assert g.get_intrinsic_costs().max() > 0


def test_window_edges_and_repeated_calls():
    """Non-square grid; a radius wider than the grid, radius 0 and a negative
    radius (only the obstacle set), and sums on top of a previous call."""
    sc = scenes.scene_sized(77, 96, 0.2, min_dist=(10, 16))
    f = np.ascontiguousarray(sc.frame[:45, :70])
    frame = P.SearchFrame(f, sc.class_costs)
    g, o = _pair(70, 45, sc.class_costs, ((30, 40), (40, 30)))
    for kr, radius in ((2.0, 300), (1.0, 0), (0.5, -3), (1.5, 9), (-2.0, 4)):
        print(f"[apf] 70x45 kr={kr} radius={radius}", flush=True)
        g.compute_apf_repulsion(frame, kr, radius)
        o.apf_repulsion(f, kr, radius)
        _same(g, o, f"repulsion kr={kr} r={radius}")
        g.compute_apf_attraction(frame, 0.25, 3, 60)
        o.apf_attraction(f, 0.25, 3, 60)
        _same(g, o, f"attraction after r={radius}")


def test_large_grid():
    sc = scenes.scene_1024(5)
    print("[apf] 1024x1024 radius=40", flush=True)
    frame = P.SearchFrame(sc.frame, sc.class_costs)
    g, o = _pair(1024, 1024, sc.class_costs, (sc.ego_lb, sc.ego_ub))
    g.compute_apf_repulsion(frame, 2.0, 40)
    o.apf_repulsion(sc.frame, 2.0, 40)
    _same(g, o, "repulsion 1024")


def test_rollout_costs_carry_intrinsic_costs():
    """derivate_node on a graph with potential-field costs: the end cell's
    cost (sum of intrinsic + 1 over the rollout's cells) equals the oracle's."""
    sc = scenes.scene_256(92, coverage=0.10, min_dist=(10, 16))
    frame = P.SearchFrame(sc.frame, sc.class_costs)
    g, o = _pair(256, 256, sc.class_costs, (sc.ego_lb, sc.ego_ub), (10, 16))
    g.compute_apf_repulsion(frame, 2.0, 20)
    o.apf_repulsion(sc.frame, 2.0, 20)
    L = O.lib()
    out = (O.ctypes.c_int * 2)()
    n_end, top = 0, 0.0
    for a in range(-60, 61, 15):
        for s in (-30, 0, 30):
            h, st = float(np.float32(np.radians(a))), float(np.float32(np.radians(s)))
            g.add(128, 128, P.angle(h), -1, -1, 0)
            o.add(128, 128, h, -1, -1, 0)
            end = g.derivate_node_end(frame, st, 30.0, 1.0, 128, 128)
            L.og_derivate_node(o.g, O.fptr(sc.frame), st, 30.0, 1.0, 128, 128, out)
            assert end == (out[0], out[1]), (a, s)
            if end[0] >= 0:
                n_end += 1
                gc = np.float32(g.get_cost(*end))
                oc = np.float32(o.plane("cost")[end[1], end[0]])
                assert gc.view(np.int32) == oc.view(np.int32), (a, s, gc, oc)
                top = max(top, float(gc))
            g.clear()
            L.og_clear(o.g)
    assert n_end >= 10
    assert top > 31.05  # above the 31 cells' plain count: the field is in the costs
