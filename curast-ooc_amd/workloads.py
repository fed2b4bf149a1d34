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
