"""The batched engine: a plan's result is the same alone, in a mixed batch,
under continuous admission, and through the plain FastRRT handle; and it is
the oracle's result."""
import numpy as np
import pytest

from _parity import OraclePlanner, Params

from fastrrt_amd import scenes
from fastrrt_amd.batch import Batch, PlanSpec
from fastrrt_amd import pyfastrrt as P

pytestmark = pytest.mark.gpu

SIZES = (96, 128, 160, 128, 96, 192)


def _scene(i):
    return scenes.scene_sized(300 + i, SIZES[i % len(SIZES)], 0.06 + 0.02 * (i % 3), min_dist=(10, 16))


def _spec(sc, frame, i, max_iter=50, rounds=4, deadline=0.0):
    return PlanSpec(frame, sc.start, sc.goal, min_dist=(10, 16), max_path_size=30.0, dist_to_goal_tolerance=12.0,
                    max_iterations=max_iter, optimize_rounds=rounds, seed=777, stream=i, tag=i,
                    deadline_ms=deadline)


def _key(r):
    return (r.status, r.iterations, r.path.shape, r.path.tobytes())


def test_alone_equals_in_the_mix():
    scs = [_scene(i) for i in range(10)]
    frames = [P.SearchFrame(sc.frame, sc.class_costs) for sc in scs]
    alone = {}
    for i, sc in enumerate(scs):
        print(f"[batch] alone {i} size={sc.width}", flush=True)
        b = Batch(max_plans=1, max_cells=sc.width * sc.height)
        b.submit(_spec(sc, frames[i], i))
        (r,) = b.drain()
        alone[i] = _key(r)
    print("[batch] mixed", flush=True)
    b = Batch(max_plans=len(scs), max_cells=sum(s.width * s.height for s in scs))
    for i, sc in enumerate(scs):
        b.submit(_spec(sc, frames[i], i))
    mixed = {r.tag: _key(r) for r in b.drain()}
    assert mixed == alone
    print("[batch] continuous (3 slots, staggered)", flush=True)
    b = Batch(max_plans=3, max_cells=3 * 192 * 192)
    got = {}
    for i, sc in enumerate(scs):
        b.submit(_spec(sc, frames[i], i))
        b.step(2)
        for r in b.poll():
            got[r.tag] = _key(r)
    for r in b.drain():
        got[r.tag] = _key(r)
    assert got == alone
    assert sum(1 for k in alone.values() if k[0] == 2) >= 3


def test_batch_equals_plain_handle_and_oracle():
    prm = Params(min_dist=(10, 16), max_path=30.0, dist_tol=12.0)
    for i in range(4):
        sc = _scene(i)
        frame = P.SearchFrame(sc.frame, sc.class_costs)
        b = Batch(max_plans=2, max_cells=2 * sc.width * sc.height)
        b.submit(_spec(sc, frame, i, max_iter=40, rounds=3))
        (r,) = b.drain()
        o = OraclePlanner(sc, prm, 777, i)
        o.search_init()
        it = 0
        while not o.goal_reached() and o.loop():
            it += 1
            if it >= 40:
                break
        if o.goal_reached():
            k = 0
            while k < 3 and o.path_optimize():
                k += 1
            po = o.path()
            assert r.status == (2 if po.shape[0] else 4)
            assert r.path.shape == po.shape and (r.path.view(np.int32) == po.view(np.int32)).all(), i
        else:
            assert r.status == 3


def test_deadline_retires_plans():
    sc = scenes.scene_sized(901, 256, 0.30, min_dist=(22, 40))
    frame = P.SearchFrame(sc.frame, sc.class_costs)
    b = Batch(max_plans=4, max_cells=4 * 256 * 256)
    for i in range(4):
        b.submit(PlanSpec(frame, sc.start, sc.goal, max_iterations=0, optimize_rounds=0, deadline_ms=30.0,
                          seed=1, stream=i, tag=i))
    res = b.drain()
    assert len(res) == 4
    for r in res:
        assert r.status in (2, 3)
        assert r.latency_ms < 30.0 + 200.0  # retired at the first poll past the deadline


def test_deadline_latency_at_the_benchmark_workload():
    """20 ms deadlines, no iteration cap, 20 optimisation rounds, 256 px
    scenes with the 22 x 40 footprint, one iteration per poll (the
    benchmark's protocol): every plan, found or failed, returns within its
    deadline plus 2 ms, several steps' worth of margin less than the loose
    bound above."""
    pool = scenes.scenes_256_sweep(8, base_seed=6060)
    frames = [P.SearchFrame(sc.frame, sc.class_costs) for sc in pool]
    b = Batch(max_plans=len(pool), max_cells=len(pool) * 65536)
    for k, sc in enumerate(pool):
        b.submit(PlanSpec(frames[k], sc.start, sc.goal, min_dist=sc.min_dist, ego_lb=sc.ego_lb, ego_ub=sc.ego_ub,
                          max_iterations=0, optimize_rounds=20, deadline_ms=20.0, seed=1234, stream=8800 + k, tag=k))
    res = []
    while b.step(1) > 0:
        res += b.poll()
    res += b.poll()
    assert sorted(r.tag for r in res) == list(range(len(pool)))
    found = [r.latency_ms for r in res if r.ok]
    failed = [r.latency_ms for r in res if not r.ok]
    print(f"[batch] deadline 20 ms: {len(found)} found, max {max(found, default=0):.2f} ms; {len(failed)} failed, "
          f"max {max(failed, default=0):.2f} ms", flush=True)
    assert failed and all(r.status in (2, 3, 4) for r in res)
    assert max(r.latency_ms for r in res) <= 20.0 + 2.0


def test_batch_smart_regions_equals_oracle():
    """Smart expansion with 32 px density regions in the batch engine (the
    request's region side) gives the oracle's search and path."""
    prm = Params(min_dist=(10, 16), max_path=30.0, dist_tol=12.0)
    for i in range(3):
        sc = _scene(10 + i)
        frame = P.SearchFrame(sc.frame, sc.class_costs)
        b = Batch(max_plans=1, max_cells=sc.width * sc.height)
        spec = PlanSpec(frame, sc.start, sc.goal, min_dist=(10, 16), max_path_size=30.0, dist_to_goal_tolerance=12.0,
                        max_iterations=40, optimize_rounds=2, seed=778, stream=i, tag=i, smart=True, region_side=32)
        b.submit(spec)
        (r,) = b.drain()
        o = OraclePlanner(sc, prm, 778, i, region_side=32)
        o.search_init()
        it = 0
        while not o.goal_reached() and o.loop(True):
            it += 1
            if it >= 40:
                break
        if o.goal_reached():
            k = 0
            while k < 2 and o.path_optimize():
                k += 1
            po = o.path()
            assert r.status == (2 if po.shape[0] else 4)
            assert r.path.shape == po.shape and (r.path.view(np.int32) == po.view(np.int32)).all(), i
        else:
            assert r.status == 3


def test_odd_sizes_alone_equals_in_the_mix():
    """Grids whose sides are not multiples of 8, 32 or 64, non-square, in one
    batch: each plan's result is its result alone."""
    from test_gpu_parity import _odd_scene
    shapes = [(101, 83), (77, 130), (200, 47), (65, 65), (129, 97)]
    scs = [_odd_scene(w, h, 500 + i) for i, (w, h) in enumerate(shapes)]
    frames = [P.SearchFrame(sc.frame, sc.class_costs) for sc in scs]

    def spec(i, sc):
        return PlanSpec(frames[i], sc.start, sc.goal, min_dist=(10, 16), max_path_size=30.0,
                        dist_to_goal_tolerance=12.0, max_iterations=40, optimize_rounds=3, seed=31, stream=i, tag=i,
                        smart=(i % 2 == 1), region_side=16 if i % 2 else 0)
    alone = {}
    for i, sc in enumerate(scs):
        b = Batch(max_plans=1, max_cells=sc.width * sc.height)
        b.submit(spec(i, sc))
        (r,) = b.drain()
        alone[i] = _key(r)
    b = Batch(max_plans=len(scs), max_cells=sum(sc.width * sc.height for sc in scs))
    for i, sc in enumerate(scs):
        b.submit(spec(i, sc))
    mixed = {r.tag: _key(r) for r in b.drain()}
    assert mixed == alone


def test_launch_shapes_leave_results_unchanged():
    """Every iteration kernel strides over its work: the launch shapes the
    autotuner picks among change the time of an iteration, never a result."""
    scs = [_scene(i) for i in range(8)]
    frames = [P.SearchFrame(sc.frame, sc.class_costs) for sc in scs]
    b = Batch(max_plans=len(scs), max_cells=sum(s.width * s.height for s in scs))
    default = b.launch()
    assert default[0] == 4096 and 1 <= default[1] <= 2048 and 1 <= default[2] <= 2048 and b.block() == 256
    runs = []
    for shape in [default + (256,), (8, 1, 1, 256), (1024, 512, 256, 64), (8192, 4096, 1024, 128),
                  (2048, 1024, 512, 192), (8, 1, 1, 64)]:
        print(f"[batch] launch shape {shape}", flush=True)
        b.set_launch(*shape)
        assert b.launch() + (b.block(),) == shape
        for i, sc in enumerate(scs):
            b.submit(_spec(sc, frames[i], i))
        runs.append({r.tag: _key(r) for r in b.drain()})
    assert all(r == runs[0] for r in runs[1:])
    assert sum(1 for k in runs[0].values() if k[0] == 2) >= 2
    for bad in [(12, 2048, 2048), (0, 2048, 2048), (4096, 0, 2048), (4096, 2048, 70000), (4096, 2048, 2048, 96),
                (4096, 2048, 2048, 512)]:
        with pytest.raises(RuntimeError):
            b.set_launch(*bad)
    assert b.launch() == (8, 1, 1) and b.block() == 64


def test_profiled_capture_equals_the_graph():
    """Kernel timing, with and without the three-part capture of the
    iteration (fastrrt_profile_split), gives every plan the result of the
    unprofiled iteration graph; the timers record the iteration (and, split,
    k_expand) once per loop iteration."""
    import ctypes
    from fastrrt_amd import _lib
    L = _lib.load()
    scs = [_scene(i) for i in range(6)]
    frames = [P.SearchFrame(sc.frame, sc.class_costs) for sc in scs]
    b = Batch(max_plans=len(scs), max_cells=sum(s.width * s.height for s in scs))
    out = {}
    for mode in ("off", "whole", "split"):
        print(f"[batch] profiling {mode}", flush=True)
        L.fastrrt_profile_reset()
        L.fastrrt_profile_split(mode == "split")
        L.fastrrt_profile_enable(mode != "off")
        for i, sc in enumerate(scs):
            b.submit(_spec(sc, frames[i], i, max_iter=30, rounds=2))
        res = b.drain()
        L.fastrrt_profile_enable(False)
        L.fastrrt_profile_split(False)
        out[mode] = {r.tag: _key(r) for r in res}
        n_it, n_ex, ms = ctypes.c_long(), ctypes.c_long(), ctypes.c_double()
        L.fastrrt_profile_read(3, ctypes.byref(n_it), ctypes.byref(ms))
        L.fastrrt_profile_read(1, ctypes.byref(n_ex), ctypes.byref(ms))
        if mode == "off":
            assert n_it.value == 0 and n_ex.value == 0
        else:
            assert n_it.value >= max(k[1] for k in out[mode].values()) - 1 > 0
            assert n_ex.value == (n_it.value if mode == "split" else 0)
    assert out["whole"] == out["off"] and out["split"] == out["off"]
