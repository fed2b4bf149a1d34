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
