"""Output formats, checkpoint formats and the graph test API through the C
ABI, against the oracle and the reference's fixtures.

* dump_graph_to_file / read_from_dump_file (graph.cu:496-546,
  python_cuda_graph_bind.cpp:188-197): the three adds of
  test_graph_dump.cpp:8-33 dump byte-identically to unittests/test.dat; a
  dump read back gives the same grid.
* save_current_graph_state / load_graph_state (python_bind.cpp:188-198): a
  search's state dumps byte-identically to the oracle's dump of the same
  state; a fresh handle and a fresh oracle that load the file hold the same
  grid (the original's, to the six decimals of the text format) and continue
  the search identically.
* Node-list fixtures written by the reference (unittests/src/test_data/:
  before_collision.txt, and before_error_1/2.txt, the trees
  fastrrt.cpp:120,141 dumped inside loop()) loaded through add /
  add_temporary / set_collision: solve_collisions leaves 165 nodes of
  before_collision, and every fixture's grid, list_all, count_all, DAG
  verdict and dump equal the oracle's.
* export_graph_nodes (python_bind.cpp:89-107), list_all
  (python_cuda_graph_bind.cpp:121-140), ideal_curve (fastrrt.cpp:270-281),
  interpolate_planned_path_p (python_bind.cpp:157-171).
* process_direct_goal_connection and is_directly_connected_to_goal /
  direct_connection_to_goal_cost / _heading, get_cost and count_all
  (python_cuda_graph_bind.cpp:161-217) against the oracle's planes."""
import gzip
import os

import numpy as np
import pytest

import _oracle as O
from _parity import GpuPlanner, OraclePlanner, Params, diff_states

from fastrrt_amd import pyfastrrt as P
from fastrrt_amd import scenes

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
LEN = 5.412658773
SMALL = Params(min_dist=(10, 16), max_path=30.0, dist_tol=12.0)


def _graph(W=256, H=256, costs=(0, 1, 2, 3, 4, -1), md=(0, 0)):
    return P.CudaGraph(W, H, 34.64, 34.64, 40, LEN, md[0], md[1], -1, -1, -1, -1, np.asarray(costs, np.float32))


def _ograph(W=256, H=256, costs=(0, 1, 2, 3, 4, -1), md=(0, 0)):
    g = O.Graph(W, H)
    g.physical(34.64, 34.64, 40.0, LEN)
    g.search(md[0], md[1])
    g.costs(list(costs))
    return g


def _gstate(g):
    s = g.export_state()
    return {k: s[k] for k in ("type", "parent_x", "parent_z", "child_count", "heading", "cost")}


def _ostate(o):
    return {"type": o.plane("type").copy(), "parent_x": o.plane("px").copy(), "parent_z": o.plane("pz").copy(),
            "child_count": o.plane("derive").copy(), "heading": o.plane("heading").copy(),
            "cost": o.plane("cost").copy()}


def _read(path):
    with open(path, "rb") as f:
        return f.read()


def test_dump_is_test_dat_and_reads_back(tmp_path):
    g = _graph(100, 100)
    g.add(50, 50, P.angle(float(np.float32(0.0))), -1, -1, 0)
    g.add(40, 40, P.angle(float(np.float32(np.radians(-11)))), 50, 50, 10)
    g.add_temporary(30, 30, P.angle(float(np.float32(np.radians(4)))), 40, 40, -20)
    path = str(tmp_path / "g.dat")
    g.dump_graph_to_file(path)
    with gzip.open(os.path.join(GOLDEN, "test.dat.gz"), "rb") as f:
        assert _read(path) == f.read()
    g2 = _graph(100, 100)
    g2.read_from_dump_file(path)
    assert g2.count_all() == 3 and g2.get_parent(30, 30) == (40, 40) and g2.get_type(30, 30) == O.TEMP
    assert g2.get_cost(30, 30) == -20.0
    _same_as_text(g, g2, path)


def _same_as_text(g, g2, path):
    """g2 was read from g's dump: integers equal, floats as printed by %f
    (six decimals), and bit for bit what the oracle parses from the file."""
    a, b = _gstate(g), _gstate(g2)
    for k in ("type", "parent_x", "parent_z", "child_count"):
        assert (a[k] == b[k]).all(), k
    for k in ("heading", "cost"):
        assert np.allclose(a[k], b[k], rtol=0, atol=5e-7 * max(1.0, float(np.abs(a[k]).max()))), k
    o = _ograph(g2._w, g2._h)
    O.lib().og_read_dump(o.g, path.encode())
    c = _ostate(o)
    bad = [k for k in b if not (b[k].view(np.int32) == c[k].view(np.int32)).all()]
    assert not bad, bad


def test_save_and_load_graph_state_continue_the_search(tmp_path):
    sc = scenes.scene_sized(41, 128, 0.08, min_dist=(10, 16))
    o = OraclePlanner(sc, SMALL, 55, 3)
    g = GpuPlanner(sc, SMALL, 55, 3)
    o.search_init()
    g.search_init()
    for _ in range(6):
        o.loop()
        g.loop()
    assert not diff_states(o.state(), g.state())
    gp, op = str(tmp_path / "gpu.dat"), str(tmp_path / "oracle.dat")
    g.rrt.save_current_graph_state(gp)
    O.lib().og_dump_graph(o.L.og_rrt_graph(o.r), op.encode())
    assert _read(gp) == _read(op)
    # a fresh handle and a fresh oracle on the same plan, both loaded from the
    # file (fastrrt.cpp:293-297), continue the search identically
    g2 = GpuPlanner(sc, SMALL, 55, 3)
    o2 = OraclePlanner(sc, SMALL, 55, 3)
    g2.search_init()
    o2.search_init()
    g2.rrt.load_graph_state(gp)
    o2.L.og_rrt_load_graph_state(o2.r, gp.encode())
    a, b = g.state(), g2.state()
    for k in ("type", "parent_x", "parent_z", "child_count"):
        assert (a[k] == b[k]).all(), k
    for k in ("heading", "cost"):
        assert np.allclose(a[k], b[k], rtol=0, atol=5e-7 * max(1.0, float(np.abs(a[k]).max()))), k
    assert not diff_states(o2.state(), g2.state())
    for it in range(5):
        ra, rb = o2.loop(), g2.loop()
        assert ra == rb
        assert not diff_states(o2.state(), g2.state()), it


def _load_nodes(g, o, text):
    """assertGraphRemainsDAGAfterCollision (test_graph_solve_collisions.cpp:231-263)"""
    L = O.lib()
    for line in text.strip().splitlines():
        p = line.split()
        x, z, h, t, px, pz, c = int(p[0]), int(p[1]), float(p[2]), int(p[3]), int(p[4]), int(p[5]), float(p[7])
        if t == 2:
            g.add_temporary(x, z, P.angle(h), px, pz, c)
            o.add_temporary(x, z, h, px, pz, c)
        else:
            g.add(x, z, P.angle(h), px, pz, c)
            o.add(x, z, h, px, pz, c)
            if t == 4:
                g.set_collision(x, z, px, pz, h, c)
                L.og_set_collision(o.g, x, z, px, pz, h, c)


def _same_graph(g, o, what, tmp_path):
    a, b = _gstate(g), _ostate(o)
    bad = [k for k in a if not (a[k].view(np.int32) == b[k].view(np.int32)).all()]
    assert not bad, f"{what}: planes differ {bad}"
    L = O.lib()
    assert g.count_all() == L.og_count_all(o.g)
    assert g.check_graph_is_consistent() == bool(L.og_check_graph_consistent(o.g))
    lst = g.list_all()
    zs, xs = np.nonzero(b["type"])
    exp = np.stack([xs, zs, b["heading"][zs, xs], b["type"][zs, xs]], 1).astype(np.float32)
    assert lst.shape == exp.shape and (lst.view(np.int32) == exp.view(np.int32)).all(), what
    gp, op = str(tmp_path / "g.dat"), str(tmp_path / "o.dat")
    g.dump_graph_to_file(gp)
    L.og_dump_graph(o.g, op.encode())
    assert _read(gp) == _read(op), what


@pytest.mark.parametrize("name", ["before_collision.txt", "before_error_1.txt", "before_error_2.txt"])
def test_reference_node_list_fixtures(name, tmp_path):
    text = open(os.path.join(GOLDEN, name)).read()
    g, o = _graph(), _ograph()
    _load_nodes(g, o, text)
    _same_graph(g, o, f"{name} loaded", tmp_path)
    g.solve_collisions()
    O.lib().og_solve_collisions(o.g)
    _same_graph(g, o, f"{name} after solve_collisions", tmp_path)
    if name == "before_collision.txt":
        assert g.count_all() == 165 and g.count(O.COLLISION) == 0 and g.check_graph_is_consistent()
    g.accept_derived_nodes()
    O.lib().og_accept_derived_nodes(o.g)
    _same_graph(g, o, f"{name} after accept", tmp_path)


def test_export_graph_nodes_ideal_curve_and_interpolation():
    sc = scenes.scene_sized(7, 128, 0.05, min_dist=(10, 16))
    o = OraclePlanner(sc, SMALL, 14, 7)
    g = GpuPlanner(sc, SMALL, 14, 7)
    o.search_init()
    g.search_init()
    for _ in range(80):
        ro = (not o.goal_reached()) and o.loop()
        rg = (not g.goal_reached()) and g.loop()
        assert ro == rg
        if not ro:
            break
    st = o.state()
    zs, xs = np.nonzero(st["type"])
    exp = np.stack([xs, zs, st["type"][zs, xs]], 1).astype(np.int32)
    got = g.rrt.export_graph_nodes()
    assert got is not None and got.shape == exp.shape and (got == exp).all()
    # ideal_curve: the Hermite curve from the coordinate origin (W/2, H/2, 0)
    out = np.zeros(3 * 4096, np.float32)
    for (gx, gz, gh) in ((64, 0, 0.3), (100, 20, -0.7), (5, 120, 1.2)):
        n = O.lib().og_interpolate_hermite(128, 128, 64, 64, 0.0, gx, gz, float(np.float32(gh)), O.fptr(out), 4096)
        wp = g.rrt.build_ideal_curve(gx, gz, float(np.float32(gh)))
        arr = np.array([[w.x, w.z, w.heading.rad()] for w in wp], np.float32)
        assert arr.shape == (n, 3) and (arr.view(np.int32) == out[:3 * n].reshape(n, 3).view(np.int32)).all()
    assert o.goal_reached(), "scene 7 reaches its goal"
    path = g.path()
    inter = g.path(interpolate=True)
    wp = g.rrt.interpolate_planned_path_p(path)
    arr = np.array([[w.x, w.z, w.heading.rad()] for w in wp], np.float32)
    assert arr.shape == inter.shape and (arr.view(np.int32) == inter.view(np.int32)).all()
    # a trailing partial triple is ignored, never read past the buffer
    flat = np.concatenate([path.reshape(-1), np.float32([1.0, 2.0])])
    L = P._lib.load()
    ptr = L.interpolate_planned_path_p(g.rrt._ptr, flat.ctypes.data_as(P._lib.fp), flat.size)
    wp2 = P._path_from_ptr(L, ptr)
    arr2 = np.array([[w.x, w.z, w.heading.rad()] for w in wp2], np.float32)
    assert (arr2.view(np.int32) == inter.view(np.int32)).all()


def test_graph_goal_connection_api():
    sc = scenes.scene_256(17, coverage=0.12, goal_x=150, goal_heading=0.4)
    frame = P.SearchFrame(sc.frame, sc.class_costs)
    fr = np.ascontiguousarray(sc.frame, np.float32)
    g = _graph(costs=sc.class_costs, md=(22, 40))
    o = _ograph(costs=sc.class_costs, md=(22, 40))
    gh = float(np.float32(sc.goal[2]))
    g.process_direct_goal_connection(frame, sc.goal[0], sc.goal[1], P.angle(gh), -1)
    O.lib().og_process_direct_goal_connection(o.g, O.fptr(fr), sc.goal[0], sc.goal[1], gh, -1.0)
    s = g.export_state()
    flag = o.plane("gc_flag") > 0
    assert (s["goal_flag"].astype(bool) == flag).all()
    assert (s["goal_cost"][flag].view(np.int32) == o.plane("gc_cost")[flag].view(np.int32)).all()
    assert (s["goal_heading"][flag].view(np.int32) == o.plane("gc_heading")[flag].view(np.int32)).all()
    assert 0 < int(flag.sum()) < flag.size
    rng = np.random.default_rng(3)
    zs, xs = np.nonzero(flag)
    for k in rng.choice(len(xs), 40, replace=False):
        x, z = int(xs[k]), int(zs[k])
        assert g.is_directly_connected_to_goal(x, z)
        assert np.float32(g.direct_connection_to_goal_cost(x, z)) == o.plane("gc_cost")[z, x]
        assert np.float32(g.direct_connection_to_goal_heading(x, z).rad()) == o.plane("gc_heading")[z, x]
    zs, xs = np.nonzero(~flag)
    for k in rng.choice(len(xs), 40, replace=False):
        assert not g.is_directly_connected_to_goal(int(xs[k]), int(zs[k]))
    # get_cost / count_all through the same handle
    g.add(128, 128, P.angle(0.0), -1, -1, 0)
    g.add(120, 100, P.angle(0.1), 128, 128, 33.5)
    g.add_temporary(125, 90, P.angle(0.2), 120, 100, 40.0)
    assert g.count_all() == 3 and g.get_cost(120, 100) == 33.5 and g.get_cost(125, 90) == 40.0
    assert g.get_cost(999, 0) == -1.0 and g.get_cost(0, 0) == 0.0
