"""The safe-zone bit 0x100 of channel 2 (fastrrt_frame_process_safe_zone).

The reference reads the bit (graph_direct_goal_connection.cu:124-135, with
isSafeZoneChecked = true at fastrrt.cpp:145) but takes it from libdriveless,
which is not in the reference tree: parity of the producer is unpinned.  The
producer is checked against its definition, brute force (no obstacle pixel
within Chebyshev distance ceil(hypot(min_dist_x // 2, min_dist_z // 2)) + 1,
obstacles = class cost < 0 or a class outside the table, outside the frame
free; every other bit kept), and the planner's use of the bit against the
oracle in lock step on a frame the producer processed.
"""
import math

import numpy as np
import pytest

from _parity import Params, run_side_by_side
from fastrrt_amd import scenes

pytestmark = pytest.mark.gpu


def _expected(frame, costs, min_dist):
    H, W = frame.shape[:2]
    cls = frame[:, :, 0].astype(np.int64)
    inside = (cls >= 0) & (cls < len(costs))
    obst = ~inside | (np.asarray(costs, np.float32)[np.clip(cls, 0, max(len(costs) - 1, 0))] < 0)
    if len(costs) == 0:
        obst = np.ones_like(obst)
    r = int(math.ceil(math.hypot(min_dist[0] // 2, min_dist[1] // 2))) + 1
    pad = np.pad(obst, r, constant_values=False)
    near = np.zeros((H, W), bool)
    for dz in range(-r, r + 1):
        for dx in range(-r, r + 1):
            near |= pad[r + dz:r + dz + H, r + dx:r + dx + W]
    tv = frame[:, :, 2].astype(np.int64)
    return np.where(near, tv & ~0x100, tv | 0x100).astype(np.float32)


@pytest.mark.parametrize("W,H,cov,md,seed", [(64, 48, 0.10, (22, 40), 1), (130, 77, 0.05, (4, 6), 2),
                                             (256, 256, 0.15, (22, 40), 3), (97, 211, 0.02, (0, 0), 4),
                                             (200, 120, 0.30, (10, 16), 5)])
def test_safe_zone_matches_definition(W, H, cov, md, seed):
    from fastrrt_amd import pyfastrrt as P
    rng = np.random.default_rng(seed)
    frame = scenes.make_frame(W, H, cov, seed)
    # other classes, one outside the table, and other channel-2 bits (kept)
    frame[:, :, 0] = np.where(rng.random((H, W)) < 0.01, 40.0, frame[:, :, 0])
    frame[:, :, 2] = rng.integers(0, 4, (H, W)).astype(np.float32) + 0x100 * (rng.random((H, W)) < 0.5)
    costs = scenes.SEGMENTATION_CLASS_COST
    f = P.SearchFrame(frame, costs)
    f.process_safe_zone(md[0], md[1])
    got = f.read()
    exp = _expected(frame, costs, md)
    assert np.array_equal(got[:, :, 2], exp), f"{np.count_nonzero(got[:, :, 2] != exp)} pixels differ"
    assert np.array_equal(got[:, :, :2], frame[:, :, :2])
    # a second pass over its own output gives the same bits
    f.process_safe_zone(md[0], md[1])
    assert np.array_equal(f.read(), got)


def test_planner_reads_the_bit_as_the_oracle():
    """A sparse 256x256 scene whose safe zone covers much of the free space:
    the goal connection skips the footprints of curves from safe cells on
    both sides, grids equal after every loop()."""
    from fastrrt_amd import pyfastrrt as P
    sc = scenes.scene_256(31, 0.05)
    f = P.SearchFrame(sc.frame, sc.class_costs)
    f.process_safe_zone(*sc.min_dist)
    processed = f.read()
    assert np.count_nonzero(processed[:, :, 2].astype(np.int64) & 0x100) > 1000
    sc2 = scenes.Scene(processed, sc.start, sc.goal, sc.seed, sc.coverage, sc.perception_m, sc.min_dist,
                       sc.ego_lb, sc.ego_ub)
    run_side_by_side(sc2, Params(min_dist=sc.min_dist), 91, 3, 40)
