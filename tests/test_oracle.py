"""The CPU oracle, pinned against the reference's own known answers, golden
fixtures and closed forms (no GPU needed).

Every case cites the reference test or source it restates.  Cases that
retype a formula would pin nothing; these instead use the reference's
expected values, fixture files, libm, and geometric/closed-form facts.
"""
import ctypes
import gzip
import math
import os

import numpy as np
import pytest

import _oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
L = O.lib()
L.og_sinf.restype = L.og_cosf.restype = L.og_tanf.restype = L.og_atanf.restype = ctypes.c_float
L.og_atan2f.restype = ctypes.c_float
for f in (L.og_sinf, L.og_cosf, L.og_tanf, L.og_atanf):
    f.argtypes = [ctypes.c_float]
L.og_atan2f.argtypes = [ctypes.c_float, ctypes.c_float]
L.og_trunc_d.argtypes = [ctypes.c_double]
L.og_philox.argtypes = [ctypes.POINTER(ctypes.c_uint32), ctypes.c_uint32, ctypes.c_uint32,
                        ctypes.POINTER(ctypes.c_uint32)]
L.og_expansion_draws.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32,
                                 ctypes.POINTER(ctypes.c_float)]

PHYS = 34.641016151377535
L_VEH = 5.412658773


def test_graph(W=256, H=256, mdx=0, mdz=0, costs=(0, 1, 2, 3, 4, -1), perc=256.0):
    g = O.Graph(W, H)
    g.physical(perc, perc, 40.0, L_VEH)
    g.costs(list(costs))
    g.search(mdx, mdz)
    return g


test_graph.__test__ = False


# ----------------------------------------------------------------------------
# shared scalar primitives vs libm
# ----------------------------------------------------------------------------

def _ulp_diff(a, b):
    a = np.asarray(a, np.float32).view(np.int32).astype(np.int64)
    b = np.asarray(b, np.float32).view(np.int32).astype(np.int64)
    a = np.where(a < 0, -(a & 0x7FFFFFFF), a)
    b = np.where(b < 0, -(b & 0x7FFFFFFF), b)
    return np.abs(a - b)


@pytest.mark.parametrize("fn,ref,lo,hi", [
    ("og_sinf", np.sin, -700.0, 700.0), ("og_cosf", np.cos, -700.0, 700.0),
    ("og_tanf", np.tan, -1.5, 1.5), ("og_atanf", np.arctan, -50.0, 50.0)])
def test_trig_within_one_ulp_of_libm(fn, ref, lo, hi):
    rng = np.random.default_rng(7)
    xs = np.concatenate([rng.uniform(lo, hi, 20000), rng.uniform(-1, 1, 5000), np.float32(
        [0.0, -0.0, 1e-30, 1.5707963705062866, -1.5707963705062866, 3.1415927410125732, 0.6981317])]).astype(
        np.float32)
    f = getattr(L, fn)
    got = np.array([f(float(x)) for x in xs], np.float32)
    want = ref(xs.astype(np.float64)).astype(np.float32)
    assert _ulp_diff(got, want).max() <= 1
    # the float nearest pi/2 gives cos ~ -4.371139e-08, not 0 (matters for
    # the Hermite tangents of src/cuda/graph_direct_goal_connection.cu:46-47)
    assert np.float32(L.og_cosf(-1.5707963705062866)) == np.float32(np.cos(np.float64(np.float32(-np.pi / 2))))


def test_atan2_quadrants_and_specials():
    rng = np.random.default_rng(3)
    ys = rng.uniform(-300, 300, 8000).astype(np.float32)
    xs = rng.uniform(-300, 300, 8000).astype(np.float32)
    got = np.array([L.og_atan2f(float(y), float(x)) for y, x in zip(ys, xs)], np.float32)
    want = np.arctan2(ys.astype(np.float64), xs.astype(np.float64)).astype(np.float32)
    assert _ulp_diff(got, want).max() <= 1
    for y, x in [(0.0, 1.0), (0.0, -1.0), (-0.0, -1.0), (1.0, 0.0), (-1.0, 0.0), (0.0, 0.0), (-0.0, 0.0),
                 (0.0, -0.0)]:
        assert np.float32(L.og_atan2f(y, x)) == np.float32(np.arctan2(y, x))


def test_truncation_device_semantics():
    assert L.og_trunc_d(float("nan")) == 0
    assert L.og_trunc_d(-0.999) == 0 and L.og_trunc_d(-1.0) == -1 and L.og_trunc_d(127.9999) == 127
    assert L.og_trunc_d(1e20) == 2**31 - 1 and L.og_trunc_d(-1e20) == -2**31


def test_philox4x32_10_known_answers():
    """Random123 known-answer vectors for Philox4x32-10 (Salmon et al. SC'11)."""
    def ph(c, k0, k1):
        ci = (ctypes.c_uint32 * 4)(*c)
        o = (ctypes.c_uint32 * 4)()
        L.og_philox(ci, k0, k1, o)
        return list(o)
    assert ph([0, 0, 0, 0], 0, 0) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert ph([0xffffffff] * 4, 0xffffffff, 0xffffffff) == [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert ph([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], 0xa4093822, 0x299f31d0) == \
        [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_expansion_draws_are_in_half_open_unit_interval():
    out = (ctypes.c_float * 2)()
    vals = []
    for cell in range(4000):
        L.og_expansion_draws(99, 7, 3, cell, out)
        vals += [out[0], out[1]]
    v = np.array(vals)
    assert v.min() > 0.0 and v.max() <= 1.0
    assert abs(v.mean() - 0.5) < 0.02


# ----------------------------------------------------------------------------
# grid basics (unittests/src/test_graph.cpp)
# ----------------------------------------------------------------------------

def test_search_params_halving():
    """test_graph.cpp:38-50"""
    g = O.Graph(100, 100)
    g.search(4, 6, 4, 5, 6, 7)
    s = g.g.contents
    assert (s.min_dist_x, s.min_dist_z, s.lb_x, s.lb_z, s.ub_x, s.ub_z) == (2, 3, 4, 5, 6, 7)


def test_physical_params():
    """test_graph.cpp:23-36"""
    g = O.Graph(100, 100)
    L.og_set_physical_params(g.g, 11.1, 11.2, math.radians(21), 11.3, -1.0)
    s = g.g.contents
    assert s.rate_w == pytest.approx(100 / 11.1, rel=1e-6)
    assert s.lr == pytest.approx(11.3 / 2, rel=1e-7)
    assert s.max_steer_rad == pytest.approx(math.radians(21))


def test_basic_features():
    """test_graph.cpp:132-205"""
    g = O.Graph(100, 100)
    assert L.og_count(g.g, O.NODE) == 0
    g.add(50, 50, 0.12, -1, -1, 0)
    g.add(50, 45, 0.12, 50, 50, 10)
    g.add(55, 40, 0.12, 50, 45, 10)
    g.add(20, 35, 0.12, 50, 45, 10)
    g.add(20, 15, 0.12, 20, 35, 10)
    g.add(20, 5, 0.12, 20, 15, 10)
    g.add(20, 0, 0.12, 20, 5, 10)
    assert L.og_count(g.g, O.NODE) == 7
    L.og_remove(g.g, 20, 0)
    assert L.og_count(g.g, O.NODE) == 6
    L.og_remove(g.g, 0, 0)
    L.og_remove(g.g, 999, -999)
    assert L.og_count(g.g, O.NODE) == 6
    assert g.parent(20, 0) == (-1, -1) and g.parent(2000, -1000) == (-1, -1)
    assert g.parent(20, 5) == (20, 15)
    assert L.og_get_cost(g.g, 55, 40) == 10.0 and L.og_get_cost(g.g, 99, 99) == 0.0
    assert L.og_get_cost(g.g, 999, 999) == -1.0


def test_dump_matches_reference_test_dat():
    """test_graph_dump.cpp:8-33 wrote unittests/test.dat: the same three adds
    must dump byte-identically (format of graph.cu:496-520)."""
    g = O.Graph(100, 100)
    g.add(50, 50, np.float32(0.0), -1, -1, 0)
    g.add(40, 40, np.float32(math.radians(-11)), 50, 50, 10)
    g.add_temporary(30, 30, np.float32(math.radians(4)), 40, 40, -20)
    assert L.og_count_all(g.g) == 3
    path = "/tmp/_og_dump_test.dat"
    L.og_dump_graph(g.g, path.encode())
    with open(path, "rb") as f:
        mine = f.read()
    with gzip.open(os.path.join(GOLDEN, "test.dat.gz"), "rb") as f:
        ref = f.read()
    assert mine == ref
    g2 = O.Graph(100, 100)
    L.og_read_dump(g2.g, path.encode())
    assert L.og_count_all(g2.g) == 3
    assert g2.parent(30, 30) == (40, 40) and g2.type_at(30, 30) == O.TEMP
    assert L.og_get_cost(g2.g, 30, 30) == -20.0
    os.remove(path)


def test_dump_round_trip_load_test():
    """test_graph.cpp:325-378: a 100x100 chain survives dump -> clear -> read."""
    g = O.Graph(100, 100)
    px, pz, cost = -1, -1, 0.0
    for z in range(100):
        for x in range(100):
            g.add(x, z, 2.0, px, pz, cost)
            cost += 1
            px, pz = x, z
    path = "/tmp/_og_chain.dat"
    L.og_dump_graph(g.g, path.encode())
    L.og_clear(g.g)
    assert L.og_count_all(g.g) == 0
    L.og_read_dump(g.g, path.encode())
    assert (g.plane("type") == O.NODE).all()
    assert (g.plane("cost").ravel() == np.arange(10000, dtype=np.float32)).all()
    os.remove(path)


# ----------------------------------------------------------------------------
# kinematics (unittests/src/test_graph_node_derivation.cpp, test_kinematics.cpp)
# ----------------------------------------------------------------------------

def _derive(g, frame, steer, psize, v, x, z):
    out = (ctypes.c_int * 2)()
    L.og_derivate_node(g.g, O.fptr(frame), steer, psize, v, x, z, out)
    return (out[0], out[1])


def test_straight_derivation_known_answers():
    """StraightDerivationInXcoord (:13-53) and ...WithDelocatedLocalPosition (:89-111)."""
    f = O.empty_frame(256, 256)
    g = test_graph()
    g.add(128, 128, 0.0, -1, -1, 0)
    assert _derive(g, f, 0.0, 49, 1.0, 128, 128) == (128, 78)
    L.og_accept_derived_nodes(g.g)
    assert g.type_at(128, 78) == O.NODE and g.heading_at(128, 78) == 0.0
    g = test_graph()
    g.add(50, 200, 0.0, -1, -1, 0)
    assert _derive(g, f, 0.0, 49, 1.0, 50, 200) == (50, 150)


def _reference_derive(start_x, start_z, h0, psize, steer, v, lr, W=256, H=256):
    """test_graph_node_derivation.cpp:113-191 (the reference test's own
    independent rollout, FP32 throughout, coordinate origin (128, 128))."""
    f32 = np.float32
    t = f32(L.og_tanf(float(steer)))
    dt = f32(0.1)
    ds = f32(f32(v) * dt)
    beta = f32(L.og_atanf(float(f32(t / f32(2)))))
    hif = f32(np.float64(f32(f32(ds * f32(L.og_cosf(float(beta)))) * t)) / (2.0 * lr))
    x = f32(128.0 - start_z)
    y = f32(start_x - 128.0)
    max_size = int(psize) + 1
    size, lx, lz = 0, start_x, start_z
    heading = f32(h0)
    while size < max_size:
        a = f32(heading + beta)
        x = f32(x + f32(ds * f32(L.og_cosf(float(a)))))
        y = f32(y + f32(ds * f32(L.og_sinf(float(a)))))
        heading = f32(heading + hif)
        cx, cz = int(128.0 + float(y)), int(128.0 - float(x))
        if (cx, cz) == (lx, lz):
            continue
        if not (0 <= cx < W and 0 <= cz < H):
            break
        size += 1
        lx, lz = cx, cz
    return lx, lz, heading


def test_curvy_derivation_matches_reference_test_rollout():
    """CurvyNodeDerivation (:193-224): 361 headings x 5 steering angles, P = 15."""
    f = O.empty_frame(256, 256)
    g = test_graph()
    lr = L_VEH / 2  # the test passes 0.5 * 5.412658773 as lr
    lr32 = float(np.float32(np.float32(L_VEH) / np.float32(2)))
    bad = 0
    for a in range(-180, 181, 3):
        for i in range(-40, 41, 18):
            h = np.float32(math.radians(a))
            s = np.float32(math.radians(i))
            g.add(100, 200, float(h), -1, -1, 0)
            end = _derive(g, f, float(s), 15, 1.0, 100, 200)
            ex, ez, eh = _reference_derive(100, 200, h, 15, s, 1.0, lr32)
            if end != (ex, ez) or np.float32(g.heading_at(*end)).view(np.int32) != np.float32(eh).view(np.int32):
                bad += 1
            L.og_clear(g.g)
    assert bad == 0
    assert abs(lr - lr32) < 1e-6


def test_straight_line_closed_form_any_heading():
    """delta = 0: the heading never changes and the end cell lies on the ray
    (truncation of origin + (P+1)-th distinct cell along (sin h, -cos h))."""
    f = O.empty_frame(256, 256)
    for hdeg in (0, 90, -90, 180, 30, -45, 135):
        g = test_graph()
        h = np.float32(math.radians(hdeg))
        g.add(128, 128, float(h), -1, -1, 0)
        ex, ez = _derive(g, f, 0.0, 20, 1.0, 128, 128)
        L.og_accept_derived_nodes(g.g)
        assert g.heading_at(ex, ez) == float(h)
        d = math.hypot(ex - 128, ez - 128)
        # 21 distinct cells: between 21 / sqrt(2) (crossing one axis per
        # cell... at most two) and 21 * sqrt(2) (diagonal corner steps)
        assert 21 / math.sqrt(2) - 1 <= d <= 21 * math.sqrt(2) + 1
        ang = math.atan2(ex - 128, -(ez - 128))
        assert abs(math.remainder(ang - float(h), 2 * math.pi)) < 0.12


def test_heading_change_bound():
    """|dh| per 0.1-px step is ds cos(beta) tan(delta) / (2 lr) <= ds tan(dmax) / (2 lr)."""
    f = O.empty_frame(256, 256)
    g = test_graph()
    g.add(128, 128, 0.0, -1, -1, 0)
    out = (ctypes.c_float * 4)()
    L.og_kinematic_new_path(g.g, O.fptr(f), 128, 128, float(np.radians(40)), 30.0, 1.0, out)
    steps_max = 10 * 31 * 2
    assert abs(out[3]) <= steps_max * 0.1 * math.tan(math.radians(40)) / L_VEH + 1e-4
    assert out[3] > 0  # positive steering turns clockwise in the map frame


def test_kinematic_path_check():
    """TestCheckKinematicPath (test_kinematics.cpp:51-72)."""
    f = O.empty_frame(256, 256)
    g = test_graph(costs=(0, 1, 2, 3, 4, 5), perc=PHYS)
    g.add(128, 128, 0.0, -1, -1, 0)
    end = _derive(g, f, float(np.float32(math.radians(20))), 50, 2.0, 128, 128)
    assert end[0] > 128 and end[1] < 128
    assert L.og_check_feasible_connection(g.g, O.fptr(f), 128, 128, end[0], end[1], 2.0) == 1


def test_rollout_rejects_on_obstacle_and_respects_bounds():
    f = O.empty_frame(256, 256, cls=0.0)
    f[60:70, :, 0] = 5.0  # a wall of class 5 (cost -1)
    g = test_graph(mdx=4, mdz=4)
    g.add(128, 128, 0.0, -1, -1, 0)
    assert _derive(g, f, 0.0, 49, 1.0, 128, 128) == (128, 78)  # stops before the wall
    g.add(128, 100, 0.0, -1, -1, 0)
    assert _derive(g, f, 0.0, 49, 1.0, 128, 100) == (-1, -1)   # drives into it
    g.add(128, 3, 0.0, -1, -1, 0)
    ex, ez = _derive(g, f, 0.0, 49, 1.0, 128, 3)               # leaves the grid: stops at row 0
    assert (ex, ez) == (128, 0)


# ----------------------------------------------------------------------------
# footprint feasibility (src/cuda/path_feasible.cu)
# ----------------------------------------------------------------------------

def test_feasibility_at_zero_heading_is_a_box_query():
    """At heading 0, c = 1 and s = 0: the lattice is the axis-aligned
    (2mz+1)x(2mx+1) box, so feasibility == no obstacle in that box (2-D
    prefix-sum query)."""
    rng = np.random.default_rng(5)
    f = O.empty_frame(64, 64, cls=1.0)
    obst = rng.random((64, 64)) < 0.02
    f[:, :, 0] = np.where(obst, 5.0, 1.0)
    g = test_graph(64, 64, mdx=6, mdz=10)
    S = np.zeros((65, 65), np.int64)
    S[1:, 1:] = obst.cumsum(0).cumsum(1)
    for z in range(64):
        for x in range(64):
            z0, z1, x0, x1 = max(0, z - 5), min(64, z + 6), max(0, x - 3), min(64, x + 4)
            n = S[z1, x1] - S[z0, x1] - S[z1, x0] + S[z0, x0]
            assert L.og_feasible(g.g, O.fptr(f), x, z, 0.0) == (1 if n == 0 else 0)


def test_feasibility_rotated_rectangle_geometry():
    """For any heading, an obstacle pixel well inside the rotated footprint
    rectangle makes the cell infeasible and one well outside cannot.  An
    asymmetric footprint (mx != mz) catches a transposed rotation."""
    rng = np.random.default_rng(11)
    mx, mz = 4, 9
    g = test_graph(64, 64, mdx=2 * mx, mdz=2 * mz)
    for _ in range(400):
        h = float(np.float32(rng.uniform(-math.pi, math.pi)))
        f = O.empty_frame(64, 64, cls=1.0)
        ox, oz = int(rng.integers(10, 54)), int(rng.integers(10, 54))
        f[oz, ox, 0] = 5.0
        f[oz, min(63, ox + 1), 0] = 5.0
        f[min(63, oz + 1), ox, 0] = 5.0
        f[min(63, oz + 1), min(63, ox + 1), 0] = 5.0
        # obstacle block centre (ox+1, oz+1) in the footprint frame of cell (32, 32)
        c, s = math.cos(h), math.sin(h)
        dx, dz = (ox + 1.0) - 32.0, (oz + 1.0) - 32.0
        j = dx * c + dz * s       # inverse rotation of (j c - i s, j s + i c)
        i = -dx * s + dz * c
        feas = L.og_feasible(g.g, O.fptr(f), 32, 32, h)
        if abs(j) <= mx - 1.5 and abs(i) <= mz - 1.5:
            assert feas == 0
        if abs(j) >= mx + 3.0 or abs(i) >= mz + 3.0:
            assert feas == 1


def test_all_angle_bit_uses_bit_zero():
    """path_feasible.cu:15: 'TO_INT(z) & 256 > 0' parses as '& 1'."""
    f = O.empty_frame(32, 32, cls=5.0)  # everything is an obstacle
    g = test_graph(32, 32, mdx=2, mdz=2)
    assert L.og_feasible(g.g, O.fptr(f), 10, 10, 0.0) == 0
    f[10, 10, 2] = 256.0
    assert L.og_feasible(g.g, O.fptr(f), 10, 10, 0.0) == 0
    f[10, 10, 2] = 1.0
    assert L.og_feasible(g.g, O.fptr(f), 10, 10, 0.0) == 1


def test_ego_box_is_ignored():
    f = O.empty_frame(64, 64, cls=1.0)
    f[30, 30, 0] = 5.0
    g = test_graph(64, 64, mdx=4, mdz=4)
    assert L.og_feasible(g.g, O.fptr(f), 30, 30, 0.0) == 0
    L.og_set_search_params(g.g, 4, 4, 28, 32, 32, 28)  # lb (28, 32), ub (32, 28)
    assert L.og_feasible(g.g, O.fptr(f), 30, 30, 0.0) == 1


# ----------------------------------------------------------------------------
# goal test and best node (unittests/src/test_graph_best_node.cpp)
# ----------------------------------------------------------------------------

def test_goal_reached_known_answer():
    g = test_graph(costs=(0, 1, 2, 3, 4, 5), perc=PHYS)
    h10 = float(np.float32(math.radians(10)))
    assert L.og_check_goal_reached(g.g, 128, 0, h10, 5.0, 1.0) == 0
    g.add(128, 128, 0.0, -1, -1, 0)
    g.add(128, 3, h10, 128, 128, 10)
    assert L.og_check_goal_reached(g.g, 128, 0, h10, 5.0, 1.0) == 1


def test_best_node_known_answer():
    f = O.empty_frame(256, 256)
    g = test_graph(costs=(0, 1, 2, 3, 4, 5), perc=PHYS)
    g.add(128, 128, 0.0, -1, -1, 0)
    for x in (128, 108, 158):
        g.add(x, 3, 0.0, 128, 128, 10)
    out = (ctypes.c_int * 2)()
    L.og_find_best_node(g.g, O.fptr(f), 0.0, 50.0, 128, 0, 1.0, out)
    assert (out[0], out[1]) == (128, 3)


def test_best_node_tie_breaks_on_lowest_cell():
    f = O.empty_frame(256, 256)
    g = test_graph(costs=(0, 1, 2, 3, 4, 5), perc=PHYS)
    g.add(128, 128, 0.0, -1, -1, 0)
    g.add(130, 3, 0.0, 128, 128, 10)
    g.add(126, 3, 0.0, 128, 128, 10)  # same distance and cost
    out = (ctypes.c_int * 2)()
    L.og_find_best_node(g.g, O.fptr(f), 0.0, 50.0, 128, 0, 1.0, out)
    assert (out[0], out[1]) == (126, 3)


# ----------------------------------------------------------------------------
# DAG consistency (unittests/src/test_graph_dag_consistency.cpp)
# ----------------------------------------------------------------------------

@pytest.mark.parametrize("adds,want", [
    ([], True),
    ([(128, 128, -1, -1)], True),
    ([(128, 128, -1, -1), (128, 78, 128, 128)], True),
    ([(128, 128, -1, -1), (128, 48, 78, 90), (78, 90, 128, 78), (128, 78, 128, 48)], False),
    ([(128, 128, -1, -1), (78, 90, 128, 78), (128, 78, 78, 90)], False),
    ([(128, 128, -1, -1), (78, 90, 78, 90)], False),
    ([(128, 128, -1, -1), (100, 100, 128, 128), (135, 100, 128, 128), (110, 70, 135, 100), (145, 70, 135, 100)],
     True),
])
def test_dag_consistency_cases(adds, want):
    g = test_graph()
    for x, z, px, pz in adds:
        g.add(x, z, 0.0, px, pz, 0)
    assert bool(L.og_check_graph_consistent(g.g)) == want


# ----------------------------------------------------------------------------
# collision erase (unittests/src/test_graph_solve_collisions.cpp + fixtures)
# ----------------------------------------------------------------------------

def _tree(g, extra=()):
    g.add(128, 128, 0.0, -1, -1, 0)
    g.add(100, 100, 0.0, 128, 128, 0)
    g.add(135, 100, 0.0, 128, 128, 0)
    g.add(110, 70, 0.0, 135, 100, 0)
    for x, z, px, pz in extra:
        g.add(x, z, 0.0, px, pz, 0)


def test_single_collision_solve():
    g = test_graph()
    _tree(g, [(145, 70, 135, 100)])
    assert [g.child_count(*c) for c in ((128, 128), (100, 100), (135, 100))] == [2, 0, 2]
    L.og_set_collision(g.g, 110, 70, 100, 100, 0.0, 0)
    L.og_solve_collisions(g.g)
    assert [g.child_count(*c) for c in ((128, 128), (100, 100), (135, 100))] == [2, 1, 1]
    assert L.og_check_graph_consistent(g.g)
    assert g.parent(110, 70) == (100, 100) and g.parent(145, 70) == (135, 100)
    assert g.type_at(110, 70) == O.NODE


def test_collision_solve_graph_erase():
    g = test_graph()
    _tree(g, [(90, 50, 110, 70), (40, 35, 110, 70), (145, 70, 135, 100)])
    L.og_set_collision(g.g, 110, 70, 100, 100, 0.0, 0)
    L.og_solve_collisions(g.g)
    assert L.og_check_graph_consistent(g.g)
    assert g.type_at(90, 50) == O.NULL and g.type_at(40, 35) == O.NULL and g.type_at(110, 70) == O.NODE
    assert [g.child_count(*c) for c in ((128, 128), (100, 100), (135, 100), (110, 70), (145, 70))] == \
        [2, 1, 1, 0, 0]


def test_dual_collision():
    g = test_graph()
    _tree(g, [(90, 50, 110, 70), (40, 35, 110, 70), (145, 70, 135, 100), (185, 60, 135, 100)])
    L.og_set_collision(g.g, 110, 70, 100, 100, 0.0, 0)
    L.og_set_collision(g.g, 90, 50, 185, 60, 0.0, 0)
    L.og_solve_collisions(g.g)
    assert [g.child_count(*c) for c in ((128, 128), (100, 100), (135, 100), (110, 70), (145, 70), (90, 50),
                                        (185, 60))] == [2, 1, 2, 0, 0, 0, 1]
    assert L.og_check_graph_consistent(g.g)
    assert g.type_at(90, 50) == O.NODE and g.parent(90, 50) == (185, 60)
    assert g.type_at(40, 35) == O.NULL


def _load_node_list(g, text):
    """assertGraphRemainsDAGAfterCollision (test_graph_solve_collisions.cpp:231-263)"""
    for line in text.strip().splitlines():
        p = line.split()
        x, z, h, t, px, pz, c = int(p[0]), int(p[1]), float(p[2]), int(p[3]), int(p[4]), int(p[5]), float(p[7])
        if t == 1:
            g.add(x, z, h, px, pz, c)
        elif t == 2:
            g.add_temporary(x, z, h, px, pz, c)
        elif t == 4:
            g.add(x, z, h, px, pz, c)
            L.og_set_collision(g.g, x, z, px, pz, h, c)


def test_log_collision_bug_investigation():
    g = test_graph()
    _load_node_list(g, "148 116 1.62867 2 133 120 0 31\n140 115 0.97043 1 133 120 0 23\n"
                       "129 123 0.14363 4 128 128 0 6\n128 128 0 1 -1 -1 0 0\n130 117 0.186008 2 129 123 0 12\n"
                       "133 120 0.690726 1 128 128 0 12\n")
    assert L.og_check_graph_consistent(g.g)
    L.og_solve_collisions(g.g)
    assert L.og_check_graph_consistent(g.g)


def test_before_collision_fixture():
    """unittests/src/test_data/before_collision.txt (184 nodes, 4 COLLISION):
    the current erase rule leaves 165 nodes and a consistent DAG."""
    g = test_graph()
    _load_node_list(g, open(os.path.join(GOLDEN, "before_collision.txt")).read())
    assert L.og_count_all(g.g) == 184 and L.og_count(g.g, O.COLLISION) == 4
    assert L.og_check_graph_consistent(g.g)
    L.og_solve_collisions(g.g)
    assert L.og_count_all(g.g) == 165
    assert L.og_count(g.g, O.COLLISION) == 0
    assert L.og_check_graph_consistent(g.g)


def test_after_collision_fixture_is_flagged():
    """after_collision.txt was written by an older NODE-only erase and holds
    dangling TEMP nodes: the DAG oracle must reject it."""
    g = test_graph()
    _load_node_list(g, open(os.path.join(GOLDEN, "after_collision.txt")).read())
    assert not L.og_check_graph_consistent(g.g)


def test_erase_matches_brute_force_on_random_trees():
    """Erase == 'some ancestor within n+4 steps is COLLISION or NULL', by a
    brute-force reachability check written on the parent arrays."""
    rng = np.random.default_rng(17)
    for trial in range(30):
        g = test_graph(32, 32)
        cells = rng.permutation(32 * 32)[:60]
        g.add(int(cells[0] % 32), int(cells[0] // 32), 0.0, -1, -1, 0)
        for k in range(1, 60):
            p = cells[rng.integers(0, k)]
            g.add(int(cells[k] % 32), int(cells[k] // 32), 0.0, int(p % 32), int(p // 32), 0)
        for k in rng.choice(np.arange(1, 60), 3, replace=False):
            np_ = cells[rng.integers(0, 60)]
            if np_ != cells[k]:
                L.og_set_collision(g.g, int(cells[k] % 32), int(cells[k] // 32), int(np_ % 32), int(np_ // 32),
                                   0.0, 0)
        t0 = g.plane("type").ravel().copy()
        px, pz = g.plane("px").ravel().copy(), g.plane("pz").ravel().copy()
        n = int((t0 == O.NODE).sum())
        L.og_solve_collisions(g.g)
        t1 = g.plane("type").ravel()
        if n <= 2:
            continue
        for pos in range(32 * 32):
            if t0[pos] == O.NULL:
                continue
            erased, cur = False, pos
            for _ in range(n + 4):
                if px[cur] == -1 and pz[cur] == -1:
                    break
                nxt = pz[cur] * 32 + px[cur]
                if t0[nxt] in (O.COLLISION, O.NULL):
                    erased = True
                    break
                cur = nxt
            assert (t1[pos] == O.NULL) == erased


# ----------------------------------------------------------------------------
# goal direct connection (unittests/src/test_graph_direct_goal_connection.cpp)
# ----------------------------------------------------------------------------

def test_narrow_space():
    """NarrowSpace (:39-132): a wall at row 100 with a 3-px gap at x=127..129;
    with a 3x3 footprint only x = 128 on row 100 connects to (128, 0)."""
    f = O.empty_frame(256, 256)
    f[100, :, 0] = 5.0
    f[100, 127:130, 0] = 0.0
    g = test_graph(mdx=2, mdz=2)
    g.add(128, 128, 0.0, -1, -1, 0)
    L.og_process_direct_goal_connection(g.g, O.fptr(f), 128, 0, 0.0, -1.0)
    row = g.plane("gc_flag")[100]
    assert [i for i in range(256) if row[i] > 0] == [128]


def test_goal_is_unfeasible():
    """GoalIsUnfeasible (:134-208): the goal pixel is an obstacle."""
    f = O.empty_frame(256, 256, cls=1.0)
    f[0, 128, 0] = 0.0
    g = test_graph(mdx=2, mdz=2, costs=(-1, 1, 2, 3, 4, 5))
    g.add(128, 128, 0.0, -1, -1, 0)
    L.og_process_direct_goal_connection(g.g, O.fptr(f), 128, 0, 0.0, -1.0)
    assert not (g.plane("gc_flag")[100] > 0).any()
    # Cells 1 <= d < 2 px from the goal have numPoints = 1, so t = 0/0 = NaN
    # (graph_direct_goal_connection.cu:57); device float->int gives 0, the
    # walk samples pixel (0, 0) and the '&&' endpoint test (:155) passes on
    # z == goal_z.  Those five cells are the only ones flagged.
    flagged = {tuple(v) for v in np.argwhere(g.plane("gc_flag") > 0).tolist()}
    assert flagged == {(0, 127), (0, 129), (1, 127), (1, 128), (1, 129)}


def test_straight_hermite_connection():
    """StraightConnectionSuccess/Fail (test_graph_can_connect_hermite.cpp:31-90):
    (128,128) heading 0 -> (128,0) heading 0 connects on a free frame and
    fails when every pixel is an obstacle."""
    f = O.empty_frame(256, 256)
    g = test_graph(costs=(0, 1, 2, 3, 4, 5))
    g.add(128, 128, 0.0, -1, -1, 0)
    c = L.og_check_direct_connection(g.g, O.fptr(f), 0.8, 128, 128, 0.0, 128, 0, 0.0, 0, 0)
    assert c > 0
    # closed form: a near-vertical curve visits one cell per row, 128 rows,
    # and the start cell's cost (0) is the base: cost = number of new cells
    assert 126.0 <= c <= 129.0
    g2 = test_graph(costs=(-1, 1, 2, 3, 4, 5))
    g2.add(128, 128, 0.0, -1, -1, 0)
    assert L.og_check_direct_connection(g2.g, O.fptr(f), 0.8, 128, 128, 0.0, 128, 0, 0.0, 0, 0) < 0


def _hermite_max_curvature(x0, z0, h0, x1, z1, h1):
    """Max curvature over the sample points of the cubic Hermite curve with
    tangents d*(cos, sin)(h - pi/2), from the derivatives of the Hermite
    basis (numpy, FP64)."""
    d = math.hypot(x1 - x0, z1 - z0)
    n = int(d)
    p0, p1 = np.array([x0, z0], float), np.array([x1, z1], float)
    m0 = d * np.array([math.cos(h0 - math.pi / 2), math.sin(h0 - math.pi / 2)])
    m1 = d * np.array([math.cos(h1 - math.pi / 2), math.sin(h1 - math.pi / 2)])
    ks = []
    for t in np.linspace(0.0, 1.0, n):
        d1 = (6 * t * t - 6 * t) * p0 + (3 * t * t - 4 * t + 1) * m0 + (-6 * t * t + 6 * t) * p1 + (3 * t * t - 2 * t) * m1
        d2 = (12 * t - 6) * p0 + (6 * t - 4) * m0 + (-12 * t + 6) * p1 + (6 * t - 2) * m1
        ks.append(abs(d1[0] * d2[1] - d1[1] * d2[0]) / (d1 @ d1) ** 1.5)
    return max(ks)


@pytest.mark.parametrize("gx,gz,gh", [(140, 128, math.pi), (150, 110, math.pi / 2), (128, 60, 0.0),
                                      (160, 100, -math.pi / 2), (110, 90, math.pi * 0.75)])
def test_hermite_curvature_bound(gx, gz, gh):
    """A connection fails exactly when the sampled Hermite curvature exceeds
    the bound (free frame, 1-px footprint); with the bound disabled it
    succeeds."""
    f = O.empty_frame(256, 256)
    g = test_graph(costs=(0, 1, 2, 3, 4, 5))
    g.add(128, 128, 0.0, -1, -1, 0)
    gh = float(np.float32(gh))
    kmax = _hermite_max_curvature(128, 128, 0.0, gx, gz, gh)
    c = L.og_check_direct_connection(g.g, O.fptr(f), 0.8, 128, 128, 0.0, gx, gz, gh, 0, 0)
    if abs(kmax - 0.8) > 0.05:
        assert (c < 0) == (kmax > 0.8), (kmax, c)
    assert L.og_check_direct_connection(g.g, O.fptr(f), -1.0, 128, 128, 0.0, gx, gz, gh, 0, 0) > 0


# ----------------------------------------------------------------------------
# output interpolation (src/interpolation.cpp)
# ----------------------------------------------------------------------------

def test_interpolation_closed_forms():
    buf = np.zeros(3 * 512, np.float32)
    n = L.og_interpolate_hermite(256, 256, 128, 128, 0.0, 128, 28, 0.0, O.fptr(buf), 512)
    pts = buf[:3 * n].reshape(n, 3)
    # 100 samples from t = 0 to t = 1 inclusive: both endpoints, rounded cells
    assert n == 100 and tuple(pts[0, :2]) == (128, 128) and tuple(pts[-1, :2]) == (128, 28)
    assert (pts[:, 0] == 128).all() and set(np.diff(pts[:, 1]).tolist()) <= {-1.0, -2.0}
    assert np.allclose(pts[:, 2], 0.0, atol=1e-6)
    # a 90-degree turn ends on the target cell row/column neighbourhood
    n = L.og_interpolate_hermite(256, 256, 100, 200, 0.0, 160, 140, math.pi / 2, O.fptr(buf), 512)
    pts = buf[:3 * n].reshape(n, 3)
    assert abs(pts[-1, 0] - 160) <= 1 and abs(pts[-1, 1] - 140) <= 1
    assert abs(pts[-1, 2] - math.pi / 2) < 0.05


# ----------------------------------------------------------------------------
# expansion rules and whole-loop invariants
# ----------------------------------------------------------------------------

def test_lowest_source_claims_a_shared_end_cell():
    """Two nodes whose straight rollouts end on the same cell: the lower cell
    index wins the claim (the rule that replaces the reference's CAS race)."""
    f = O.empty_frame(256, 256)
    g = test_graph()
    g.add(128, 128, 0.0, -1, -1, 0)
    g.add(128, 100, 0.0, -1, -1, 0)
    # force both to roll straight ahead with the same length through draws:
    # find a seed/iteration where both steer within a hair of 0 is unlikely, so
    # instead use derivate semantics: sequential host expansion claims first
    a = _derive(g, f, 0.0, 10, 1.0, 128, 100)
    assert a == (128, 89)
    b = _derive(g, f, 0.0, 38, 1.0, 128, 128)
    assert b == (128, 89) and g.type_at(128, 89) == O.TEMP and g.parent(128, 89) == (128, 100)


def _run_plan(sc, max_iter, smart=False, seed=1, stream=0, min_dist=(10, 16), max_path=30.0, tol=12.0):
    r = L.og_rrt_new(sc.width, sc.height, sc.perception_m, sc.perception_m, 40.0, L_VEH, -1, -1, -1, -1, -1.0)
    L.og_rrt_set_seed(r, seed, stream)
    frame = np.ascontiguousarray(sc.frame)
    costs = np.ascontiguousarray(sc.class_costs)
    L.og_rrt_set_plan_data(r, O.fptr(frame), O.fptr(costs), len(costs), sc.start[0], sc.start[1], sc.start[2],
                           sc.goal[0], sc.goal[1], sc.goal[2], 1.0, min_dist[0], min_dist[1], -1, max_path, tol,
                           np.float32(math.radians(10)))
    L.og_rrt_search_init(r)
    return r, frame


def test_loop_invariants_small_scenes():
    """After every loop(): the DAG check passes; at the end the returned path
    starts at the start, ends within the goal tolerance, every waypoint is
    footprint-feasible at its heading, and consecutive headings differ by
    less than pi/2."""
    from fastrrt_amd import scenes
    solved, long_solves = 0, 0
    for seed in range(10):
        sc = scenes.scene_sized(seed, 128, 0.10, min_dist=(10, 16))
        r, frame = _run_plan(sc, 60, seed=seed, stream=seed, min_dist=(10, 16), max_path=30.0, tol=12.0)
        g = L.og_rrt_graph(r)
        it = 0
        while not L.og_rrt_goal_reached(r) and L.og_rrt_loop(r, 0):
            it += 1
            assert L.og_check_graph_consistent(g), f"seed {seed} iteration {it}"
            if it >= 60:
                break
        if L.og_rrt_goal_reached(r):
            solved += 1
            long_solves += it > 10
            buf = np.zeros(3 * 4096, np.float32)
            n = L.og_rrt_get_planned_path(r, O.fptr(buf), 4096)
            p = buf[:3 * n].reshape(n, 3)
            assert n >= 2 and tuple(p[0, :2]) == (sc.start[0], sc.start[1])
            assert math.hypot(p[-1, 0] - sc.goal[0], p[-1, 1] - sc.goal[1]) <= 12.0
            for x, z, h in p:
                assert L.og_feasible(g, O.fptr(frame), int(x), int(z), float(h)) == 1
            assert (np.abs(np.diff(p[:, 2])) < math.pi / 2).all()
            while L.og_rrt_path_optimize(r):
                assert L.og_check_graph_consistent(g)
        L.og_rrt_free(r)
    assert solved >= 4 and long_solves >= 1


def test_density_counts_match_recount():
    from fastrrt_amd import scenes
    sc = scenes.scene_sized(4, 96, 0.05, min_dist=(8, 12))
    r, frame = _run_plan(sc, 10, smart=True, min_dist=(8, 12), max_path=24.0, tol=10.0)
    g = L.og_rrt_graph(r)
    L.og_set_region_side(g, 32)
    for _ in range(6):
        L.og_rrt_loop(r, 1)
    s = g.contents
    types = np.ctypeslib.as_array(s.type, shape=(96, 96))
    L.og_count_regions(g)
    reg = np.ctypeslib.as_array(s.region, shape=(s.dens_h, s.dens_w)).copy()
    want = np.zeros_like(reg)
    for z in range(96):
        for x in range(96):
            if types[z, x] not in (O.NULL, O.PROCESSING):
                want[z // 32, x // 32] += 1
    assert (reg == want).all() and want.sum() > 0
    L.og_rrt_free(r)


def test_reseed_when_start_blocked():
    """fastrrt.cpp:132-139: a start whose every expansion fails leaves the
    graph empty after erase only if it collides; with an obstacle-locked
    start the tree keeps the start and produces no TEMP nodes."""
    from fastrrt_amd import scenes
    sc = scenes.scene_sized(2, 64, 0.0, min_dist=(8, 12))
    sc.frame[:, :, 0] = 0.0  # every pixel is an obstacle (class 0)
    r, frame = _run_plan(sc, 3, min_dist=(8, 12), max_path=24.0, tol=4.0)
    g = L.og_rrt_graph(r)
    assert L.og_rrt_loop(r, 0) == 1
    assert L.og_count(g, O.NODE) == 1 and L.og_count(g, O.TEMP) == 0
    L.og_rrt_free(r)
