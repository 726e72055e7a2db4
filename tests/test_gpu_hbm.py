"""Admission by measured device memory (fastrrt_batch_memory,
fastrrt_batch_set_hbm_reserve): the cell budget from hipMemGetInfo at
creation, and a plan whose new planes do not fit beside the reserve waiting
in the queue until a running plan gives its planes back."""
import pytest

from fastrrt_amd import scenes
from fastrrt_amd.batch import Batch, PlanSpec
from fastrrt_amd.pyfastrrt import SearchFrame

pytestmark = pytest.mark.gpu


def _spec(sc, frame, tag, iters=24):
    return PlanSpec(frame, sc.start, sc.goal, min_dist=sc.min_dist, ego_lb=sc.ego_lb, ego_ub=sc.ego_ub,
                    max_iterations=iters, optimize_rounds=2, seed=5, stream=31000 + tag, tag=tag)


def test_budget_measured_at_creation():
    b = Batch(max_plans=4, max_cells=4 * 65536)
    m = b.memory()
    print(f"[hbm] {m}", flush=True)
    assert 0 < m["free_bytes_at_create"] <= m["total_bytes"]
    assert m["reserve_bytes"] == min(2 << 30, m["total_bytes"] // 16)
    assert m["requested_cells"] == m["cell_budget"] == 4 * 65536
    assert m["engine_bytes"] > 0 and m["slab_bytes"] == 0
    assert m["memory_stalls"] == 0 and m["running"] == 0 and m["queued"] == 0


def test_admission_stalls_at_the_measured_limit_and_resumes():
    scs = [scenes.scene_256(7100 + k, coverage=0.10) for k in range(5)]
    frames = [SearchFrame(sc.frame, sc.class_costs) for sc in scs]
    b = Batch(max_plans=8, max_cells=8 * 65536)
    # one plan alone: its planes are allocated and then pooled
    b.submit(_spec(scs[0], frames[0], 100))
    (r0,) = b.drain()
    slab = b.memory()["slab_bytes"]
    assert slab > 65536 * 50, slab
    print(f"[hbm] one 256 x 256 plan's planes: {slab} bytes", flush=True)
    # leave room for one new slab beside the pooled one
    free = b.memory()["free_bytes_last"]
    b.set_hbm_reserve(int(free - 1.5 * slab))
    for k, (sc, fr) in enumerate(zip(scs, frames)):
        b.submit(_spec(sc, fr, k))
    b.step(1)
    m = b.memory()
    print(f"[hbm] after the first step: {m}", flush=True)
    assert m["running"] == 2 and m["queued"] == 3, m
    assert m["memory_stalls"] >= 1 and m["slab_bytes"] == 2 * slab
    res, most = [], m["running"]
    while b.step(1) > 0:
        most = max(most, b.memory()["running"])
        res += b.poll()
    res += b.poll()
    m = b.memory()
    print(f"[hbm] drained: {m}", flush=True)
    assert sorted(r.tag for r in res) == list(range(5))
    assert most <= 2 and m["slab_bytes"] == 2 * slab and m["queued"] == 0
    # each plan's result is its result alone
    for k in (0, 4):
        b1 = Batch(max_plans=1, max_cells=65536)
        b1.submit(_spec(scs[k], frames[k], k))
        (ra,) = b1.drain()
        rb = next(r for r in res if r.tag == k)
        assert (ra.status, ra.iterations, ra.path.tobytes()) == (rb.status, rb.iterations, rb.path.tobytes())


def test_plan_that_never_fits_fails_loudly():
    sc = scenes.scene_256(7200, coverage=0.10)
    fr = SearchFrame(sc.frame, sc.class_costs)
    b = Batch(max_plans=2, max_cells=2 * 65536)
    b.set_hbm_reserve(b.memory()["free_bytes_last"])  # no new planes fit, none pooled, none running
    b.submit(_spec(sc, fr, 0))
    with pytest.raises(RuntimeError, match="device memory"):
        b.step(1)
    assert b.memory()["queued"] == 0 and b.memory()["memory_stalls"] == 1
