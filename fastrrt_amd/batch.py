"""Batched / served planning over the fastrrt_batch_* C ABI.

A Batch holds up to `max_plans` concurrent plans (total grid cells bounded by
`max_cells`, i.e. by HBM).  submit() queues a request; step() admits queued
requests as slots free up, runs loop iterations of every running plan (one
hipGraph replay each) and retires finished plans; poll() returns results.
"""
import ctypes
import math

import numpy as np

from . import _lib
from .pyfastrrt import SearchFrame


class PlanSpec:
    """One plan request (the parameters set_plan_data takes, plus batch-only
    fields: iteration cap, optimisation rounds, deadline, RNG key, and the
    optional expansion budget: 0 expands every eligible node each iteration,
    as the reference; > 0 keeps each eligible node with probability budget /
    (eligible nodes of the previous iteration), a different search)."""

    def __init__(self, frame: SearchFrame, start, goal, min_dist=(22, 40), ego_lb=(-1, -1), ego_ub=(-1, -1),
                 velocity=1.0, max_steering_deg=40.0, vehicle_length=5.412658773, max_curvature=-1.0,
                 max_path_size=40.0, dist_to_goal_tolerance=20.0, heading_error_tolerance=math.radians(10),
                 max_iterations=200, optimize_rounds=20, smart=False, deadline_ms=0.0, seed=0, stream=0, tag=0,
                 region_side=0, expansion_budget=0):
        self.frame = frame
        r = _lib.PlanRequest()
        r.frame = frame.get_cuda_ptr()
        r.start_x, r.start_z, r.start_heading_rad = int(start[0]), int(start[1]), float(start[2])
        r.goal_x, r.goal_z, r.goal_heading_rad = int(goal[0]), int(goal[1]), float(goal[2])
        r.velocity_m_s = velocity
        r.min_dist_x, r.min_dist_z = int(min_dist[0]), int(min_dist[1])
        r.lower_bound_x, r.lower_bound_z = int(ego_lb[0]), int(ego_lb[1])
        r.upper_bound_x, r.upper_bound_z = int(ego_ub[0]), int(ego_ub[1])
        r.max_steering_angle_deg = max_steering_deg
        r.vehicle_length = vehicle_length
        r.max_curvature = max_curvature
        r.max_path_size = max_path_size
        r.dist_to_goal_tolerance = dist_to_goal_tolerance
        r.heading_error_tolerance_rad = float(np.float32(heading_error_tolerance))
        r.max_iterations = int(max_iterations)
        r.optimize_rounds = int(optimize_rounds)
        r.smart = 1 if smart else 0
        r.region_side = int(region_side)
        r.deadline_ms = float(deadline_ms)
        r.seed = int(seed) & ((1 << 64) - 1)
        r.stream = int(stream) & ((1 << 64) - 1)
        r.user_tag = int(tag)
        r.expansion_budget = int(expansion_budget)
        self.req = r


class Result:
    __slots__ = ("tag", "status", "iterations", "expansions", "path", "submit_ms", "done_ms", "path_cost",
                 "work")

    @property
    def ok(self):
        return self.status == 2

    @property
    def latency_ms(self):
        return self.done_ms - self.submit_ms


class Batch:
    """launch: None keeps the engine's launch shapes; "auto" applies the
    tuned shapes (fastrrt_amd.autotune) of the grid side max_cells/max_plans
    corresponds to; a (grid_waves, grid_tiles, grid_items[, threads per
    block]) tuple sets them.  Results do not depend on the choice."""

    def __init__(self, max_plans=256, max_cells=256 * 256 * 256, path_cap=65536, launch=None):
        L = _lib.load()
        self._ptr = L.fastrrt_batch_create(int(max_plans), int(max_cells))
        if not self._ptr:
            _lib.check("batch_create")
        self._keep = {}
        self._buf = np.zeros(3 * path_cap, np.float32)
        self._cap = path_cap
        if launch == "auto":
            from . import autotune
            launch = autotune.lookup(math.sqrt(max(1, int(max_cells)) / max(1, int(max_plans))))
        if launch is not None:
            self.set_launch(*launch)

    def __del__(self):
        try:
            if getattr(self, "_ptr", None):
                _lib.load().fastrrt_batch_destroy(self._ptr)
                self._ptr = None
        except Exception:
            pass

    def submit(self, spec: PlanSpec) -> int:
        t = _lib.load().fastrrt_batch_submit(self._ptr, ctypes.byref(spec.req))
        if t < 0:
            _lib.check("batch_submit")
        self._keep[spec.req.user_tag] = spec  # frames stay alive until the plan is polled
        return t

    def step(self, iterations=1) -> int:
        n = _lib.load().fastrrt_batch_step(self._ptr, int(iterations))
        if n < 0:
            _lib.check("batch_step")
        return n

    def poll(self):
        L = _lib.load()
        res = _lib.PlanResult()
        out = []
        while L.fastrrt_batch_poll(self._ptr, ctypes.byref(res), self._buf.ctypes.data_as(_lib.fp), self._cap):
            r = Result()
            r.tag, r.status, r.iterations, r.expansions = res.user_tag, res.status, res.iterations, res.expansions
            n = res.path_len
            r.path = self._buf[:3 * n].reshape(n, 3).copy() if n > 0 else np.zeros((0, 3), np.float32)
            r.submit_ms, r.done_ms = res.submit_ms, res.done_ms
            r.path_cost = float(res.path_cost)
            r.work = (res.work_rollout_cells, res.work_goal_points, res.work_goal_foot_points)
            self._keep.pop(r.tag, None)
            out.append(r)
        _lib.check("batch_poll")
        return out

    def drain(self):
        _lib.load().fastrrt_batch_drain(self._ptr)
        _lib.check("batch_drain")
        return self.poll()

    def clock_ms(self) -> float:
        return _lib.load().fastrrt_batch_clock_ms(self._ptr)

    def set_launch(self, grid_waves, grid_tiles, grid_items, block=None):
        """Launch shapes of the iteration graph, in blocks, and optionally the
        threads per block (64, 128, 192 or 256; the results do not depend on
        them; fastrrt_amd.autotune measures them)."""
        L = _lib.load()
        if block is not None and not L.fastrrt_batch_set_block_threads(self._ptr, int(block)):
            _lib.check("batch_set_block_threads")
        if not L.fastrrt_batch_set_launch(self._ptr, int(grid_waves), int(grid_tiles), int(grid_items)):
            _lib.check("batch_set_launch")

    def block(self):
        """threads per block of the iteration kernels"""
        return int(_lib.load().fastrrt_batch_get_block_threads(self._ptr))

    def memory(self):
        """The batch's device memory as measured (fastrrt_batch_memory): free
        and total bytes, the reserve, what the batch holds, the cell budget
        and the admissions held back by memory."""
        h = _lib.BatchHbm()
        if not _lib.load().fastrrt_batch_memory(self._ptr, ctypes.byref(h)):
            _lib.check("batch_memory")
        return {name: int(getattr(h, name)) for name, _ in h._fields_}

    def set_hbm_reserve(self, nbytes):
        """Bytes left free beside the batch when new plans' planes are
        allocated (default 2 GiB, at most a 16th of the device)."""
        if not _lib.load().fastrrt_batch_set_hbm_reserve(self._ptr, int(nbytes)):
            _lib.check("batch_set_hbm_reserve")

    def launch(self):
        """(grid_waves, grid_tiles, grid_items) of the iteration graph"""
        out = (ctypes.c_int * 3)()
        if not _lib.load().fastrrt_batch_get_launch(self._ptr, out):
            _lib.check("batch_get_launch")
        return tuple(int(v) for v in out)
