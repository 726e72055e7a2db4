"""Seeded synthetic planning scenes.

A scene is a bird's-eye frame of float3 pixels in the reference's layout
(channel 0 = segmentation class id, channel 1 = distance-to-goal, channel 2 =
traversability bits; reference src/cuda/path_feasible.cu:7-16), plus a start
and a goal pose in grid pixels.

Class ids follow the 29-entry segmentation table the reference's Python tests
use (unittests/python/test_utils.py:56-86): classes 1, 6, 7, 22, 24, 25, 26,
27 are drivable (cost 0), every other class is an obstacle (cost -1).  Free
pixels are class 1 and obstacle pixels class 0.  Channels 1 and 2 stay 0, so
no pixel is marked "feasible for all headings".

Obstacles are axis-aligned rectangles with integer sides drawn uniformly from
[4, 30] px at uniform positions until the obstacle coverage reaches the target
fraction; optional random-walk corridors (used on the 1024x1024 frame) carve
or add strips of a given width.  The start and goal footprints (a
min_dist-sized box around each) are cleared so that every scene is feasible at
its endpoints.  The random stream is numpy's Philox keyed by the scene seed.
"""
import math

import numpy as np

SEGMENTATION_CLASS_COST = np.array(
    [-1, 0, -1, -1, -1, -1, 0, 0, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, 0, -1, 0, 0, 0, 0, -1],
    dtype=np.float32)
FREE_CLASS = 1.0
OBSTACLE_CLASS = 0.0

# Vehicle/search constants of the 256x256 workload (unittests/python/test_utils.py:16-23,
# unittests/src/test_rrt.cpp:170-171,203).
EGO_LOWER_BOUND = (119, 148)
EGO_UPPER_BOUND = (137, 108)


class Scene:
    __slots__ = ("frame", "class_costs", "start", "goal", "width", "height", "seed", "coverage",
                 "perception_m", "min_dist", "ego_lb", "ego_ub")

    def __init__(self, frame, start, goal, seed, coverage, perception_m, min_dist, ego_lb, ego_ub):
        self.frame = frame
        self.class_costs = SEGMENTATION_CLASS_COST
        self.start = start
        self.goal = goal
        self.height, self.width = frame.shape[:2]
        self.seed = seed
        self.coverage = coverage
        self.perception_m = perception_m
        self.min_dist = min_dist
        self.ego_lb = ego_lb
        self.ego_ub = ego_ub

    def obstacle_mask(self):
        cls = self.frame[:, :, 0].astype(np.int64)
        return SEGMENTATION_CLASS_COST[np.clip(cls, 0, len(SEGMENTATION_CLASS_COST) - 1)] < 0


def _rng(seed):
    return np.random.Generator(np.random.Philox(key=int(seed) & ((1 << 64) - 1)))


def _clear_box(obst, x, z, hx, hz):
    H, W = obst.shape
    obst[max(0, z - hz):min(H, z + hz + 1), max(0, x - hx):min(W, x + hx + 1)] = False


def make_frame(width, height, coverage, seed, keep_free=(), corridors=0, corridor_width=(20, 60), carve=False):
    """Obstacle mask -> float3 frame.  keep_free: [(x, z, half_x, half_z), ...].
    corridors: random walks that add obstacle strips, or with `carve` clear
    free corridors of the full width."""
    rng = _rng(seed)
    obst = np.zeros((height, width), dtype=bool)
    target = coverage * width * height
    # rectangles until coverage (counting overlap once)
    count = 0
    guard = 0
    while count < target and guard < 1_000_000:
        guard += 1
        w = int(rng.integers(4, 31))
        h = int(rng.integers(4, 31))
        x0 = int(rng.integers(0, max(1, width - w + 1)))
        z0 = int(rng.integers(0, max(1, height - h + 1)))
        before = int(obst[z0:z0 + h, x0:x0 + w].sum())
        obst[z0:z0 + h, x0:x0 + w] = True
        count += w * h - before
    # random-walk corridors: obstacle strips of the given width
    for _ in range(corridors):
        cw = int(rng.integers(corridor_width[0], corridor_width[1] + 1))
        x, z = float(rng.integers(0, width)), float(rng.integers(0, height))
        ang = float(rng.uniform(0, 2 * math.pi))
        steps = int(rng.integers(width // 4, width))
        for _s in range(steps):
            ang += float(rng.normal(0.0, 0.15))
            x += math.cos(ang)
            z += math.sin(ang)
            if not (0 <= x < width and 0 <= z < height):
                break
            r = cw // 2
            xi, zi = int(x), int(z)
            if carve:
                obst[max(0, zi - r):min(height, zi + r + 1), max(0, xi - r):min(width, xi + r + 1)] = False
            else:
                obst[max(0, zi - r // 4):min(height, zi + r // 4 + 1),
                     max(0, xi - r // 4):min(width, xi + r // 4 + 1)] = True
    for (x, z, hx, hz) in keep_free:
        _clear_box(obst, x, z, hx, hz)
    frame = np.zeros((height, width, 3), dtype=np.float32)
    frame[:, :, 0] = np.where(obst, OBSTACLE_CLASS, FREE_CLASS)
    return frame


def scene_256(seed, coverage=0.10, goal_x=128, goal_heading=0.0, min_dist=(22, 40)):
    """The 256x256 plain workload: start (128,128,0), goal on row 0."""
    hx, hz = min_dist[0] // 2 + 1, min_dist[1] // 2 + 1
    start = (128, 128, 0.0)
    goal = (int(goal_x), 0, float(goal_heading))
    frame = make_frame(256, 256, coverage, seed, keep_free=[(start[0], start[1], hx, hz), (goal[0], goal[1], hx, hz)])
    return Scene(frame, start, goal, seed, coverage, 34.641016151377535, min_dist, EGO_LOWER_BOUND, EGO_UPPER_BOUND)


def scenes_256_sweep(n, base_seed=1000, coverages=(0.05, 0.15, 0.30)):
    """n seeded scenes per coverage: goal x ~ U[64,192] on row 0, goal heading ~ U[-pi/4, pi/4]."""
    out = []
    for cov in coverages:
        for i in range(n):
            seed = base_seed + i + int(cov * 1e6)
            rng = _rng(seed ^ 0x9E3779B97F4A7C15)
            gx = int(rng.integers(64, 193))
            gh = float(rng.uniform(-math.pi / 4, math.pi / 4))
            out.append(scene_256(seed, cov, gx, gh))
    return out


def scene_1024(seed, coverage=0.20, corridors=12, min_dist=(22, 40)):
    """The 1024x1024 workload (136 m): rectangles to `coverage`, then free
    random-walk corridors 20-60 px wide carved through them, start
    (512,512,0), goal x ~ U[256,768] on row 0, goal heading ~ U[-pi/4, pi/4]."""
    rng = _rng(seed ^ 0x51ED270B)
    gx = int(rng.integers(256, 769))
    gh = float(rng.uniform(-math.pi / 4, math.pi / 4))
    hx, hz = min_dist[0] // 2 + 1, min_dist[1] // 2 + 1
    start = (512, 512, 0.0)
    goal = (gx, 0, gh)
    frame = make_frame(1024, 1024, coverage, seed, keep_free=[(start[0], start[1], hx, hz), (goal[0], goal[1], hx, hz)],
                       corridors=corridors, carve=True)
    return Scene(frame, start, goal, seed, coverage, 136.0, min_dist, (-1, -1), (-1, -1))


def scene_sized(seed, size, coverage, min_dist=(22, 40)):
    """Square scene of side `size` with a start/goal pair at least 60% of the
    frame height apart (batched workload)."""
    rng = _rng(seed ^ 0xA5A5A5A5)
    hx, hz = min_dist[0] // 2 + 1, min_dist[1] // 2 + 1
    sx = int(rng.integers(size // 4, 3 * size // 4))
    sz = int(rng.integers(int(0.8 * size), size - 1))
    gx = int(rng.integers(size // 4, 3 * size // 4))
    gz = int(rng.integers(0, max(1, int(0.2 * size) - 1)))
    if sz - gz < int(0.6 * size):
        gz = max(0, sz - int(0.6 * size))
    gh = float(rng.uniform(-math.pi / 4, math.pi / 4))
    frame = make_frame(size, size, coverage, seed, keep_free=[(sx, sz, hx, hz), (gx, gz, hx, hz)])
    return Scene(frame, (sx, sz, 0.0), (gx, gz, gh), seed, coverage, size * 0.135316469, min_dist, (-1, -1),
                 (-1, -1))
