"""Seeded synthetic workloads standing in for CuRast's datasets.

Each function returns numpy arrays (or crx structs) that are bit-for-bit
reproducible from the seed; the generators themselves are native
(include/crx_gen.h) so a 400M-triangle heightfield takes seconds.

    sphere_mesh        displaced lat-long sphere (512 x 1024 grid = 1,048,576 triangles)
    heightfield_mesh   n x n value-noise terrain (n = 14142 gives 399,935,762 triangles)
    instance_transforms  random T*R*S world matrices in a cube
    *_view             the camera each configuration is rendered from
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

import crx

SEED_SPHERE = 0xC0FFEE01
SEED_HEIGHTFIELD = 0xC0FFEE02
SEED_INSTANCES = 0xC0FFEE03
SEED_VIEWS = 0xC0FFEE04

FOVY = math.radians(60.0)
NEAR = 0.01


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def hash64(a: np.ndarray) -> int:
    a = np.ascontiguousarray(a)
    return int(crx.gen().crx_gen_hash64(_p(a), a.nbytes))


def sphere_mesh(rings: int = 512, segments: int = 1024, radius: float = 1.0, noise: float = 0.02,
                seed: int = SEED_SPHERE, colors: bool = True):
    V = (rings + 1) * segments
    T = 2 * rings * segments
    pos = np.empty((V, 3), dtype=np.float32)
    idx = np.empty((T, 3), dtype=np.uint32)
    col = np.empty(V, dtype=np.uint32) if colors else None
    crx.gen().crx_gen_sphere(rings, segments, radius, noise, seed, _p(pos), _p(idx), _p(col))
    return pos, idx, col


def heightfield_mesh(n: int, spacing: float = 1.0, amplitude: float = 120.0, scale: float = 2048.0,
                     octaves: int = 6, seed: int = SEED_HEIGHTFIELD, colors: bool = True,
                     indices: bool = True, threads: int = 0):
    V = n * n
    T = 2 * (n - 1) * (n - 1)
    pos = np.empty((V, 3), dtype=np.float32)
    idx = np.empty((T, 3), dtype=np.uint32) if indices else None
    col = np.empty(V, dtype=np.uint32) if colors else None
    crx.gen().crx_gen_heightfield(n, spacing, amplitude, scale, octaves, seed, _p(pos), _p(idx),
                                  _p(col), threads)
    return pos, idx, col


def heightfield_indices(n: int, r0: int, r1: int, c0: int, c1: int) -> np.ndarray:
    """Triangles of the quads [r0, r1) x [c0, c1) of an n x n heightfield."""
    T = 2 * (r1 - r0) * (c1 - c0)
    idx = np.empty((T, 3), dtype=np.uint32)
    got = crx.gen().crx_gen_heightfield_indices(n, r0, r1, c0, c1, _p(idx))
    assert got == T
    return idx


def instance_transforms(count: int = 3000, extent: float = 200.0, smin: float = 0.5, smax: float = 2.0,
                        seed: int = SEED_INSTANCES):
    arr = (crx.Mat4 * count)()
    crx.gen().crx_gen_instances(count, extent, smin, smax, seed, arr)
    return arr


def projection(width: int, height: int, fovy: float = FOVY, near: float = NEAR) -> crx.Mat4:
    aspect = float(np.float32(width) / np.float32(height))
    return crx.create_projection_matrix(near, fovy, aspect)


def sphere_view(width: int = 1920, height: int = 1080, radius: float = 1.0):
    """Camera at 3 radii on +z looking at the origin."""
    view = crx.look_at((0.0, 0.0, 3.0 * radius), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0))
    return view, projection(width, height)


def heightfield_view(n: int, width: int = 1920, height: int = 1080, altitude: float = 300.0,
                     pitch_deg: float = 40.0, setback: float = 200.0):
    """Oblique view from beyond the z = 0 edge of the terrain, looking along +z
    and down by `pitch_deg`: cells near the camera project to ~3 px, the far
    edge of the view to ~0.1-0.5 px."""
    x = 0.5 * (n - 1)
    eye = (x, altitude, -setback)
    pitch = math.radians(pitch_deg)
    center = (x, altitude - math.sin(pitch), -setback + math.cos(pitch))
    view = crx.look_at(eye, center, (0.0, 1.0, 0.0))
    return view, projection(width, height)


def instances_view(extent: float = 200.0, width: int = 1920, height: int = 1080, distance: float = 350.0):
    """Camera outside the instance cube looking at its centre."""
    c = 0.5 * extent
    d = distance / math.sqrt(3.0)
    view = crx.look_at((c + d, c + d, c + d), (c, c, c), (0.0, 1.0, 0.0))
    return view, projection(width, height)


MULTIVIEW_SIZES = ((256, 256), (512, 512), (1280, 720), (1920, 1080), (3840, 2160))


def _look_dir(yaw: float, pitch: float):
    """Unit direction with heading `yaw` (about +y) tilted down by `pitch`."""
    cp = math.cos(pitch)
    return (cp * math.sin(yaw), -math.sin(pitch), cp * math.cos(yaw))


def multiview_views(n: int, positions: np.ndarray, count: int = 64, closeup_fraction: float = 0.25,
                    sizes=MULTIVIEW_SIZES, seed: int = SEED_VIEWS, spacing: float = 1.0):
    """Seeded camera poses over an n x n heightfield (positions = its vertex
    array, used for the ground height under each camera).

    Resolutions are drawn uniformly from `sizes`.  round(count*closeup_fraction)
    views are close-ups: the eye 1.5-4 units above the ground looking down
    30-70 degrees, so one cell spans several hundred pixels and its triangles
    go through stages 2 and 3.  The others hover 100-600 units above the ground
    at a random point, heading and 20-60 degree pitch.
    Returns a list of (view, proj, width, height, is_closeup)."""
    rng = np.random.default_rng(seed)
    pos = positions.reshape(-1, 3)
    n_close = int(round(count * closeup_fraction))
    close = np.zeros(count, dtype=bool)
    close[rng.permutation(count)[:n_close]] = True
    out = []
    for k in range(count):
        W, H = sizes[int(rng.integers(len(sizes)))]
        yaw = float(rng.uniform(0.0, 2.0 * math.pi))
        if close[k]:
            c = int(rng.integers(8, n - 9))
            r = int(rng.integers(8, n - 9))
            ground = float(pos[r * n + c, 1])
            eye = ((c + 0.5) * spacing, ground + float(rng.uniform(1.5, 4.0)), (r + 0.5) * spacing)
            pitch = math.radians(float(rng.uniform(30.0, 70.0)))
        else:
            c = int(rng.integers(0, n))
            r = int(rng.integers(0, n))
            ground = float(pos[r * n + c, 1])
            eye = (c * spacing, ground + float(rng.uniform(100.0, 600.0)), r * spacing)
            pitch = math.radians(float(rng.uniform(20.0, 60.0)))
        d = _look_dir(yaw, pitch)
        view = crx.look_at(eye, (eye[0] + d[0], eye[1] + d[1], eye[2] + d[2]), (0.0, 1.0, 0.0))
        out.append((view, projection(W, H), W, H, bool(close[k])))
    return out


SEED_SERVING = 0xC0FFEE05


def tiled_heightfield(ctx, n: int, tile: int, amplitude: float = 120.0, scale: float = 2048.0, octaves: int = 6,
                      seed: int = SEED_HEIGHTFIELD, chunk_bytes: int = 1 << 30, threads: int = 0,
                      colors: bool = True):
    """The n x n heightfield of heightfield_mesh cut into square nodes of
    `tile` x `tile` cells ((tile+1)^2 vertices; neighbours share their border
    vertices), generated directly into page-locked host memory.

    Returns (nodes, buffers): nodes is a list of dicts with the pinned
    pointers (pos, idx, col), counts (nv, nt), the object-space aabb and the
    tile origin (r0, c0); buffers are the crx.PinnedBuffer chunks that own
    the memory."""
    al = lambda x: (x + 255) // 256 * 256
    starts = list(range(0, n - 1, tile))
    specs = []
    for r0 in starts:
        for c0 in starts:
            rows = min(tile, n - 1 - r0) + 1
            cols = min(tile, n - 1 - c0) + 1
            nv = rows * cols
            nt = 2 * (rows - 1) * (cols - 1)
            size = al(12 * nv) + al(12 * nt) + (al(4 * nv) if colors else 0)
            specs.append((r0, c0, rows, cols, nv, nt, size))
    buffers, nodes = [], []
    cur, used = None, 0
    for (r0, c0, rows, cols, nv, nt, size) in specs:
        if cur is None or used + size > cur.nbytes:
            cur = crx.PinnedBuffer(ctx, max(chunk_bytes, size))
            buffers.append(cur)
            used = 0
        pos = cur.array(np.float32, 3 * nv, used)
        idx = cur.array(np.uint32, 3 * nt, used + al(12 * nv))
        col = cur.array(np.uint32, nv, used + al(12 * nv) + al(12 * nt)) if colors else None
        crx.gen().crx_gen_heightfield_tile(n, r0, c0, rows, cols, 1.0, amplitude, scale, octaves, seed,
                                           _p(pos), _p(idx), _p(col), threads)
        p3 = pos.reshape(-1, 3)
        lo, hi = p3.min(axis=0), p3.max(axis=0)
        box = crx.Box3(crx.Vec3(*map(float, lo)), crx.Vec3(*map(float, hi)))
        nodes.append({"pos": pos.ctypes.data, "idx": idx.ctypes.data, "col": col.ctypes.data if colors else None,
                      "nv": nv, "nt": nt, "aabb": box, "r0": r0, "c0": c0, "rows": rows, "cols": cols,
                      "heights": p3[:, 1].reshape(rows, cols)})
        used += size
    return nodes, buffers


def ground_lookup(nodes, tile: int):
    """Height of the tiled terrain at the grid vertex nearest to (x, z)."""
    by_origin = {(nd["r0"], nd["c0"]): nd for nd in nodes}
    last = max(nd["r0"] for nd in nodes)

    def ground(x: float, z: float) -> float:
        r, c = int(round(z)), int(round(x))
        r0 = min((max(r, 0) // tile) * tile, last)
        c0 = min((max(c, 0) // tile) * tile, last)
        nd = by_origin[(r0, c0)]
        lr = min(max(r - r0, 0), nd["rows"] - 1)
        lc = min(max(c - c0, 0), nd["cols"] - 1)
        return float(nd["heights"][lr, lc])
    return ground


def serving_requests(n: int, ground, duration: float = 10.0, rate: float = 200.0, sessions: int = 8,
                     sizes=MULTIVIEW_SIZES, seed: int = SEED_SERVING, speed: float = 40.0,
                     altitude=(80.0, 250.0)):
    """Seeded request stream for the serving workload: Poisson arrivals at
    `rate` per second for `duration` seconds, each from one of `sessions`
    cameras flying over the terrain (straight paths at `speed` units/s that
    reflect at the terrain border, `altitude` (80-250) units above the ground,
    looking down 35-70 degrees), resolutions uniform over `sizes`.
    Returns a list of (arrival_seconds, view, proj, width, height, session)."""
    rng = np.random.default_rng(seed)
    span = float(n - 1)
    sess = []
    for _ in range(sessions):
        sess.append({
            "x": float(rng.uniform(0.1, 0.9) * span), "z": float(rng.uniform(0.1, 0.9) * span),
            "heading": float(rng.uniform(0.0, 2.0 * math.pi)), "alt": float(rng.uniform(*altitude)),
            "pitch": math.radians(float(rng.uniform(35.0, 70.0)))})

    def reflect(p: float) -> float:
        p = p % (2.0 * span)
        return p if p <= span else 2.0 * span - p

    out, t = [], 0.0
    while True:
        t += float(rng.exponential(1.0 / rate))
        if t >= duration:
            break
        k = int(rng.integers(sessions))
        s = sess[k]
        x = reflect(s["x"] + speed * t * math.sin(s["heading"]))
        z = reflect(s["z"] + speed * t * math.cos(s["heading"]))
        eye = (x, ground(x, z) + s["alt"], z)
        d = _look_dir(s["heading"], s["pitch"])
        view = crx.look_at(eye, (eye[0] + d[0], eye[1] + d[1], eye[2] + d[2]), (0.0, 1.0, 0.0))
        W, H = sizes[int(rng.integers(len(sizes)))]
        out.append((t, view, projection(W, H), W, H, k))
    return out
