"""The oracle's potential-field costs (src/cuda/apf.cu) against what the
mathematics fixes, independently of the oracle's float32 operation order:
the closed form around one isolated obstacle, brute-force float64 sums over
several obstacles (source selection: edge pixels only, the grid border as an
obstacle, the ego box skipped), the attraction's closed form, and the
reference's own TestAPF sequence (unittests/src/test_apf.cpp:13-33)."""
import numpy as np

import _oracle as O

COSTS = [-1.0, 0.0, 2.0, -1.0]  # class 0 and 3 obstacles, 1 and 2 free


def _graph(W, H, ego=(-1, -1, -1, -1), costs=COSTS):
    g = O.Graph(W, H)
    g.physical(34.641016151377535, 34.641016151377535, 40.0, 5.412658773)
    g.search(0, 0, *ego)
    g.costs(costs)
    return g


def _frame(W, H, obstacles):
    f = O.empty_frame(W, H, cls=1.0)
    for (x, z) in obstacles:
        f[z, x, 0] = 0.0
    return f


def _brute(W, H, frame, kr, radius, prior, ego=(-1, -1, -1, -1)):
    """float64 restatement from the definition: edge obstacles push every
    pixel within radius, obstacles are set to 100*kr/2."""
    lbx, lbz, ubx, ubz = ego
    in_ego = lambda x, z: lbx <= x <= ubx and ubz <= z <= lbz
    obst = np.array([[COSTS[int(frame[z, x, 0])] < 0 for x in range(W)] for z in range(H)])
    ob = lambda x, z: not (0 <= x < W and 0 <= z < H) or obst[z, x]
    out = prior.astype(np.float64).copy()
    for z in range(H):
        for x in range(W):
            if obst[z, x] and not in_ego(x, z):
                out[z, x] = 100 * kr / 2
    for z in range(H):
        for x in range(W):
            if not obst[z, x] or in_ego(x, z):
                continue
            if ob(x - 1, z) and ob(x + 1, z) and ob(x, z - 1) and ob(x, z + 1):
                continue
            for h in range(max(0, z - radius), min(H, z + radius + 1)):
                for w in range(max(0, x - radius), min(W, x + radius + 1)):
                    if (w, h) == (x, z) or in_ego(w, h):
                        continue
                    p = np.hypot(w - x, h - z)
                    if p <= radius:
                        out[h, w] += kr / 2 * (1 / p - 1 / radius) ** 2
    return out


def test_single_obstacle_closed_form():
    W = H = 64
    g = _graph(W, H)
    g.apf_repulsion(_frame(W, H, [(30, 20)]), 2.0, 5)
    got = g.plane("intrinsic").astype(np.float64)
    zz, xx = np.mgrid[0:H, 0:W]
    p = np.hypot(xx - 30, zz - 20)
    want = np.where((p <= 5) & (p > 0), 1.0 * (1 / np.maximum(p, 1e-9) - 0.2) ** 2, 0.0)
    want[20, 30] = 100.0
    assert np.allclose(got, want, rtol=2e-6, atol=1e-7)
    # the disc edge: p = 5 exactly adds zero, p just beyond adds nothing
    assert got[20, 35] == 0.0 and got[23, 34] == 0.0 and got[20, 36] == 0.0
    assert np.count_nonzero(got) == 1 + np.count_nonzero((p <= 5) & (p > 0) & (np.abs(p - 5) > 1e-12))


def test_many_obstacles_match_float64_sums():
    rng = np.random.default_rng(5)
    W, H = 48, 40
    obs = set()
    for _ in range(6):  # rectangles: interior pixels are not sources
        x0, z0 = rng.integers(0, W - 6), rng.integers(0, H - 6)
        for z in range(z0, z0 + rng.integers(2, 7)):
            for x in range(x0, x0 + rng.integers(2, 7)):
                obs.add((int(x), int(z)))
    obs |= {(0, z) for z in range(10, 20)}  # touches the border: still an edge (free neighbour at x=1)
    obs |= {(x, 0) for x in range(W)}  # a full border row: edge through its free neighbours below
    ego = (20, 30, 26, 24)  # lb=(20,30), ub=(26,24): x in [20,26], z in [24,30]
    frame = _frame(W, H, obs)
    for kr, radius in ((2.0, 7), (0.5, 13), (3.0, 1)):
        g = _graph(W, H, ego)
        prior = np.linspace(-1, 1, W * H, dtype=np.float32).reshape(H, W)
        g.plane("intrinsic")[:] = prior
        g.apf_repulsion(frame, kr, radius)
        got = g.plane("intrinsic").astype(np.float64)
        want = _brute(W, H, frame, kr, radius, prior, ego)
        assert np.allclose(got, want, rtol=1e-5, atol=1e-5), (kr, radius, np.abs(got - want).max())
        lbx, lbz, ubx, ubz = ego
        # the ego box keeps its prior values
        assert (got[ubz:lbz + 1, lbx:ubx + 1] == prior[ubz:lbz + 1, lbx:ubx + 1]).all()


def test_enclosed_obstacles_do_not_push():
    """A 5x5 block: its 3x3 core is enclosed and pushes nothing; only its
    rim does.  Removing the core from the frame must not change any free
    pixel's cost (the rim is the same edge set)."""
    W = H = 32
    block = {(x, z) for x in range(10, 15) for z in range(10, 15)}
    rim = {(x, z) for (x, z) in block if x in (10, 14) or z in (10, 14)}
    a = _graph(W, H)
    a.apf_repulsion(_frame(W, H, block), 2.0, 6)
    b = _graph(W, H)
    b.apf_repulsion(_frame(W, H, rim), 2.0, 6)
    free = np.ones((H, W), bool)
    for (x, z) in block:
        free[z, x] = False
    assert np.array_equal(a.plane("intrinsic")[free], b.plane("intrinsic")[free])


def test_attraction_closed_form():
    W, H = 40, 30
    obs = [(5, 5), (6, 5), (20, 12)]
    g = _graph(W, H, (10, 20, 15, 15))  # the attraction does not skip the ego box
    prior = np.full((H, W), 0.25, np.float32)
    g.plane("intrinsic")[:] = prior
    g.apf_attraction(_frame(W, H, obs), 3.0, 17, 0)
    got = g.plane("intrinsic").astype(np.float64)
    zz, xx = np.mgrid[0:H, 0:W]
    want = 0.25 - 1.5 * ((17 - xx) ** 2 + zz ** 2)
    for (x, z) in obs:
        want[z, x] = 0.25
    assert np.allclose(got, want, rtol=1e-6, atol=0)


def test_reference_apf_sequence():
    """test_apf.cpp:13-33: class costs {0..5} (none negative), one pixel of
    class 2 at (128, 128): the repulsion finds no obstacle and the attraction
    (Ka = 3 towards (128, 0)) sets -1.5 * squared distance everywhere."""
    W = H = 256
    g = _graph(W, H, costs=[0, 1, 2, 3, 4, 5])
    f = O.empty_frame(W, H)
    f[128, 128, 0] = 2
    g.apf_repulsion(f, 2.0, 5)
    assert not g.plane("intrinsic").any()
    g.apf_attraction(f, 3.0, 128, 0)
    zz, xx = np.mgrid[0:H, 0:W]
    assert np.array_equal(g.plane("intrinsic"), (-1.5 * ((128 - xx) ** 2 + zz ** 2)).astype(np.float32))
