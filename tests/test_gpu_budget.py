"""The optional expansion budget (PlanSpec(expansion_budget=...), off by
default): with it the search thins its eligible nodes, so it is not the
reference's search and has no oracle; what holds is that budget 0 is the
default search, that a budgeted plan is the same alone and in a mixed batch
(its thinning draws come from its own Philox key), that it runs fewer rollouts
per iteration, and that every path it returns is valid."""
import math

import pytest

from fastrrt_amd import scenes, validate
from fastrrt_amd.batch import Batch, PlanSpec
from fastrrt_amd.pyfastrrt import SearchFrame

pytestmark = pytest.mark.gpu


def _plans(n=12):
    pool = scenes.scenes_256_sweep(n // 3, base_seed=5150)
    return [(sc, SearchFrame(sc.frame, sc.class_costs)) for sc in pool]


def _spec(sc, fr, k, budget):
    return PlanSpec(fr, sc.start, sc.goal, min_dist=sc.min_dist, ego_lb=sc.ego_lb, ego_ub=sc.ego_ub,
                    max_iterations=60, optimize_rounds=4, seed=99, stream=4000 + k, tag=k, expansion_budget=budget)


def _run(plans, budget, only=None):
    idx = range(len(plans)) if only is None else [only]
    b = Batch(max_plans=len(idx), max_cells=len(idx) * 65536)
    for k in idx:
        b.submit(_spec(plans[k][0], plans[k][1], k, budget))
    return {r.tag: r for r in b.drain()}


def _key(r):
    return (r.status, r.iterations, r.expansions, r.path.tobytes())


def test_budget_zero_is_the_default_search():
    plans = _plans()
    b = Batch(max_plans=len(plans), max_cells=len(plans) * 65536)
    for k, (sc, fr) in enumerate(plans):
        b.submit(PlanSpec(fr, sc.start, sc.goal, min_dist=sc.min_dist, ego_lb=sc.ego_lb, ego_ub=sc.ego_ub,
                          max_iterations=60, optimize_rounds=4, seed=99, stream=4000 + k, tag=k))
    default = {r.tag: _key(r) for r in b.drain()}
    zero = {t: _key(r) for t, r in _run(plans, 0).items()}
    assert default == zero


def test_budgeted_plans_alone_equal_in_the_mix_and_return_valid_paths():
    plans = _plans()
    full = _run(plans, 0)
    mixed = _run(plans, 4)
    for k in (0, 5, 11):
        alone = _run(plans, 4, only=k)[k]
        assert _key(alone) == _key(mixed[k]), k
    rate = lambda res: sum(r.expansions for r in res.values()) / max(1, sum(r.iterations for r in res.values()))
    print(f"[budget] rollouts per iteration: {rate(full):.2f} without, {rate(mixed):.2f} with a budget of 4",
          flush=True)
    assert rate(mixed) < rate(full)
    found = [(k, r) for k, r in mixed.items() if r.ok]
    for k, r in found:
        sc = plans[k][0]
        fails, _ = validate.check_path(sc.frame, sc.class_costs, r.path, sc.start, sc.goal, sc.min_dist,
                                       (sc.ego_lb, sc.ego_ub), 20.0, math.radians(10))
        assert not fails, (k, fails)
    print(f"[budget] {len(found)} of {len(mixed)} budgeted plans found a path, all valid", flush=True)
