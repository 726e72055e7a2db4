"""The large, mixed and served configurations at full size, every returned
path validated (fastrrt_amd.validate, itself pinned against the oracle in
test_validate.py: the chain from the start, the goal tolerances, each
edge's footprints or Hermite connection, heading continuity).

* 1024x1024 over 136 m, 20 % rectangles plus carved random-walk corridors
  20-60 px wide, smart expansion with 32 px density regions, 300 ms per plan,
  22x40 footprint: plain loop and batched form.
* 512 plans over {128, 256, 384, 512}^2, coverage U[5, 30] %, start and goal
  at least 60 % of the frame height apart, 22x40 footprint, one batch; four
  of them also run alone and return the same result.
* a Poisson stream of 2 000 requests/s for 2 s over grid sides 128-1024
  (log-spaced), deadlines 10/50/100 ms, 22x40 footprint, continuous
  admission: every request returns once."""
import math

import numpy as np
import pytest

from fastrrt_amd import serve
from fastrrt_amd.batch import Batch, PlanSpec

pytestmark = pytest.mark.gpu


def _all_valid(v, what):
    print(f"[{what}] validation {v}", flush=True)
    assert v["valid"] == v["paths"], f"{what}: invalid paths {v['failed']}"


def test_1024_large_frame_300ms():
    out = serve.run_1024(n_plans=24, seed=13, timeout_ms=300.0, opt_rounds=20)
    print(f"[1024] {out}", flush=True)
    for form in ("plain", "batched"):
        f = out[form]
        assert f["mean_iterations"] > 0 and 0.0 <= f["success"] <= 1.0
        _all_valid(f["validation"], f"1024 {form}")
    assert out["plain"]["validation"]["paths"] + out["batched"]["validation"]["paths"] >= 1
    assert out["batched"]["p99_ms"] <= 300.0 + 5.0
    assert out["plain"]["p99_ms"] <= 300.0 + 60.0  # set_plan_data and path_optimize fall outside the timeout


def test_512_mixed_size_batch():
    out = serve.run_batched(n_plans=512, seed=12, max_iter=200, opt_rounds=20)
    print(f"[mixed] {dict((k, v) for k, v in out.items() if k != 'validation')}", flush=True)
    assert out["returned"] == 512 and set(out["statuses"]) <= {2, 3, 4}
    assert set(out["by_size"]) == {"128", "256", "384", "512"}
    assert out["success_rate"] > 0.0
    _all_valid(out["validation"], "mixed 512")


def test_512_mixed_size_batch_plans_alone_equal_in_the_mix():
    """Four plans of the mixed workload alone and among the others (one per
    grid size): the same status, iterations and path."""
    from fastrrt_amd import scenes
    from fastrrt_amd.pyfastrrt import SearchFrame
    rng = np.random.default_rng(12)
    plans = []
    for i in range(64):
        size = serve.BATCHED_SIZES[int(rng.integers(0, 4))]
        cov = float(rng.uniform(0.05, 0.30))
        sc = scenes.scene_sized(int(12 * 100000 + i), size, cov)
        plans.append((size, sc, SearchFrame(sc.frame, sc.class_costs)))

    def spec(i):
        size, sc, frame = plans[i]
        return PlanSpec(frame, sc.start, sc.goal, min_dist=sc.min_dist, max_iterations=120, optimize_rounds=10,
                        seed=12, stream=700000 + i, tag=i)

    b = Batch(max_plans=64, max_cells=sum(s * s for s, _, _ in plans))
    for i in range(64):
        b.submit(spec(i))
    mixed = {r.tag: (r.status, r.iterations, r.path.tobytes()) for r in b.drain()}
    picks = {}
    for i, (size, _, _) in enumerate(plans):
        picks.setdefault(size, i)
    for size, i in sorted(picks.items()):
        b1 = Batch(max_plans=1, max_cells=size * size)
        b1.submit(spec(i))
        (r,) = b1.drain()
        assert (r.status, r.iterations, r.path.tobytes()) == mixed[i], (size, i)


def test_served_poisson_stream_full_sizes():
    pool = serve.ScenePool(serve.SERVED_SIZES, 3, seed=11)
    out = serve.run_served(duration_s=2.0, rate=2000.0, sizes=serve.SERVED_SIZES, seed=11, per_size=3,
                           max_plans=256, max_cells=64 << 20, pool=pool)
    print(f"[served] {dict((k, v) for k, v in out.items() if k != 'validation')}", flush=True)
    assert out["returned"] == out["requests"] > 3000
    assert set(out["by_size"]) == {str(s) for s in serve.SERVED_SIZES}
    for k in ("p50_ms", "p99_ms", "completed_per_s"):
        assert math.isfinite(out[k]) and out[k] > 0
    assert out["p99_ms"] <= 100.0 + 5.0
    _all_valid(out["validation"], "served")
