"""compute_apf_repulsion_noalloc: the potential-field repulsion with its
scratch in the library's static device array.  Through the C ABI, its
intrinsic-cost plane equals compute_apf_repulsion's (and the oracle's) bit
for bit, on two calls in a row (the second on a smaller grid and radius, so
the array still holds the first call's larger contents), on sums over a
previous call, and a grid/radius beyond the array's capacity is refused on
the host before any kernel runs."""
import ctypes

import numpy as np
import pytest

import _oracle as O
from fastrrt_amd import _lib
from fastrrt_amd import pyfastrrt as P
from fastrrt_amd import scenes

pytestmark = pytest.mark.gpu


def _graph(W, H, costs, ego):
    lb, ub = ego
    return P.CudaGraph(W, H, 34.641016151377535, 34.641016151377535, 40, 5.412658773, 22, 40, lb[0], lb[1], ub[0],
                       ub[1], np.asarray(costs, np.float32))


def _oracle_plane(frame, costs, ego, calls):
    W, H = frame.shape[1], frame.shape[0]
    o = O.Graph(W, H)
    o.physical(34.641016151377535, 34.641016151377535, 40.0, 5.412658773)
    o.search(22, 40, ego[0][0], ego[0][1], ego[1][0], ego[1][1])
    o.costs(costs)
    for kr, radius in calls:
        o.apf_repulsion(frame, kr, radius)
    return o.plane("intrinsic").copy()  # the view dies with the oracle graph


def _check_case(frame_px, costs, ego, calls, what):
    H, W = frame_px.shape[:2]
    L = _lib.load()
    frame = P.SearchFrame(frame_px, costs)
    g_ref = _graph(W, H, costs, ego)
    g_na = _graph(W, H, costs, ego)
    for kr, radius in calls:
        print(f"[no-alloc] {what} {W}x{H} kr={kr} radius={radius}", flush=True)
        L.compute_apf_repulsion(g_ref._ptr, frame.get_cuda_ptr(), ctypes.c_float(kr), radius)
        _lib.check("compute_apf_repulsion")
        ok = L.compute_apf_repulsion_noalloc(g_na._ptr, frame.get_cuda_ptr(), ctypes.c_float(kr), radius)
        assert ok, _lib.last_error()
    a, b = g_ref.get_intrinsic_costs(), g_na.get_intrinsic_costs()
    bad = np.argwhere(a.view(np.int32) != b.view(np.int32))
    assert bad.shape[0] == 0, f"{what}: {bad.shape[0]} cells differ from compute_apf_repulsion, first {bad[:3]}"
    c = _oracle_plane(np.ascontiguousarray(frame_px, np.float32), costs, ego, calls)
    bad = np.argwhere(b.view(np.int32) != c.view(np.int32))
    assert bad.shape[0] == 0, f"{what}: {bad.shape[0]} cells differ from the oracle"
    assert b.max() > 0


def test_noalloc_matches_allocating_entry_point_two_calls():
    sc = scenes.scene_256(91, coverage=0.15)
    _check_case(sc.frame, sc.class_costs, (sc.ego_lb, sc.ego_ub), [(2.0, 40)], "first call")
    small = scenes.scene_sized(92, 128, 0.2, min_dist=(10, 16))
    f = np.ascontiguousarray(small.frame[:96, :])
    _check_case(f, small.class_costs, ((50, 60), (70, 40)), [(1.5, 12)], "second call, smaller")


def test_noalloc_sums_over_previous_calls():
    sc = scenes.scene_sized(93, 96, 0.2, min_dist=(10, 16))
    f = np.ascontiguousarray(sc.frame[:45, :70])
    _check_case(f, sc.class_costs, ((30, 40), (40, 30)), [(2.0, 300), (0.5, -3), (1.5, 9)], "repeated")


def test_noalloc_refuses_beyond_capacity():
    L = _lib.load()
    cap = L.fastrrt_apf_noalloc_scratch_capacity()
    assert cap == 4 << 20
    assert L.fastrrt_apf_noalloc_scratch_bytes(1024, 1024, 1006) <= cap
    assert L.fastrrt_apf_noalloc_scratch_bytes(1024, 1024, 1007) > cap
    sc = scenes.scene_1024(5, coverage=0.05, corridors=0)
    frame = P.SearchFrame(sc.frame, sc.class_costs)
    g = _graph(1024, 1024, sc.class_costs, ((-1, -1), (-1, -1)))
    print("[no-alloc] 1024x1024 radius=1007 (refused on the host)", flush=True)
    assert not L.compute_apf_repulsion_noalloc(g._ptr, frame.get_cuda_ptr(), ctypes.c_float(1.0), 1007)
    assert "scratch" in _lib.last_error()
    assert float(np.abs(g.get_intrinsic_costs()).max()) == 0.0
