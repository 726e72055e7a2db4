"""Served and mixed-size batched modes: every request comes back exactly once
with a valid status, requests of several grid sizes run together, and the
reported figures are finite."""
import math

import numpy as np
import pytest

from fastrrt_amd import serve

pytestmark = pytest.mark.gpu


def test_served_poisson_stream_mixed_sizes():
    pool = serve.ScenePool((128, 192, 256), 3, seed=5, coverage=(0.02, 0.05), min_dist=(10, 16))
    out = serve.run_served(duration_s=1.0, rate=300.0, sizes=(128, 192, 256), per_size=3, seed=5, max_plans=64,
                           max_cells=64 * 256 * 256, max_iter=60, opt_rounds=5, pool=pool)
    print(f"[serve] {out}", flush=True)
    assert out["requests"] > 100
    assert set(out["by_size"]) == {"128", "192", "256"}
    assert sum(v["requests"] for v in out["by_size"].values()) == out["requests"]
    for k in ("p50_ms", "p99_ms", "completed_per_s"):
        assert math.isfinite(out[k]) and out[k] > 0
    assert 0.0 < out["success_rate"] <= 1.0
    assert 0.0 <= out["deadline_miss_rate"] <= 1.0


def test_batched_mixed_sizes_all_return():
    out = serve.run_batched(n_plans=24, sizes=(128, 256, 384), seed=9, max_iter=80, coverage=(0.02, 0.05),
                            min_dist=(10, 16))
    print(f"[batched-mixed] {out}", flush=True)
    assert sum(v["plans"] for v in out["by_size"].values()) == 24
    assert out["success_rate"] > 0.0
    assert all(v["mean_iterations"] <= 80 for v in out["by_size"].values())


def test_1024_configuration_small():
    """The 1024x1024 runner (smart expansion, 32 px density regions) on two
    plans, both forms; every result is a valid status and the figures are
    finite."""
    out = serve.run_1024(n_plans=2, seed=3, timeout_ms=60.0, opt_rounds=2)
    print(f"[1024] {out}", flush=True)
    for form in ("plain", "batched"):
        f = out[form]
        assert math.isfinite(f["plans_per_s"]) and f["plans_per_s"] > 0
        assert 0.0 <= f["success"] <= 1.0
        assert f["mean_iterations"] > 0
    assert out["batched"]["p99_ms"] <= 60.0 + 5.0
