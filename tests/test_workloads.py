"""Seeded workload generators (CPU): heightfield tiles equal slices of the
full grid, the multi-view and serving camera streams are reproducible and
have the stated mix."""
import math

import numpy as np

import crx
import workloads as wl


def test_heightfield_tiles_equal_full_grid():
    n = 97
    pos, idx, col = wl.heightfield_mesh(n)
    g = crx.gen()
    for (r0, c0, rows, cols) in [(0, 0, 33, 33), (32, 64, 33, 33), (64, 32, 33, 33), (90, 0, 7, 97)]:
        p = np.empty((rows * cols, 3), np.float32)
        c = np.empty(rows * cols, np.uint32)
        i = np.empty((2 * (rows - 1) * (cols - 1), 3), np.uint32)
        g.crx_gen_heightfield_tile(n, r0, c0, rows, cols, 1.0, 120.0, 2048.0, 6, wl.SEED_HEIGHTFIELD,
                                   wl._p(p), wl._p(i), wl._p(c), 0)
        assert np.array_equal(p, pos.reshape(n, n, 3)[r0:r0 + rows, c0:c0 + cols].reshape(-1, 3))
        assert np.array_equal(c, col.reshape(n, n)[r0:r0 + rows, c0:c0 + cols].reshape(-1))
        # tile-local triangles are the full mesh's triangles of the same quads
        lr, lc = np.divmod(i.astype(np.int64), cols)
        assert np.array_equal((lr + r0) * n + (lc + c0),
                              wl.heightfield_indices(n, r0, r0 + rows - 1, c0, c0 + cols - 1).astype(np.int64))


def test_multiview_views_mix():
    pos, _, _ = wl.heightfield_mesh(257, colors=False, indices=False)
    a = wl.multiview_views(257, pos, count=64)
    b = wl.multiview_views(257, pos, count=64)
    assert sum(c for *_, c in a) == 16
    assert {(W, H) for (_, _, W, H, _) in a} <= set(wl.MULTIVIEW_SIZES)
    assert all(bytes(x[0]) == bytes(y[0]) for x, y in zip(a, b))


def test_serving_requests_poisson_and_reproducible():
    ground = lambda x, z: 0.0
    r1 = wl.serving_requests(1001, ground, duration=20.0, rate=200.0)
    r2 = wl.serving_requests(1001, ground, duration=20.0, rate=200.0)
    assert len(r1) == len(r2) and all(bytes(a[1]) == bytes(b[1]) for a, b in zip(r1, r2))
    # Poisson: ~4000 arrivals (sd ~63), exponential gaps with mean 5 ms
    assert 3700 < len(r1) < 4300
    gaps = np.diff([t for t, *_ in r1])
    assert abs(gaps.mean() - 0.005) < 0.0005 and abs(np.median(gaps) - 0.005 * math.log(2)) < 0.0006
    # cameras stay over the terrain
    for (_, view, _, _, _, _) in r1[::97]:
        eye = np.linalg.inv(view.to_numpy().astype(np.float64)) @ np.array([0, 0, 0, 1.0])
        assert -1 <= eye[0] <= 1001 and -1 <= eye[2] <= 1001
