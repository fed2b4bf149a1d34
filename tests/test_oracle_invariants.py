"""Invariants and brute-force checks of the CPU oracle.

* brute force: on tiny random scenes every pixel's winner is recomputed by
  exhaustive evaluation with exact rational coverage and double-precision
  depths; pixels whose decision is within a small margin of an edge or of a
  depth tie are skipped (the oracle's float arithmetic may legitimately fall
  either way there), all others must agree;
* order independence: shuffling triangles, thread counts and instancing vs
  duplicated meshes leave the buffer bit-identical;
* frame preparation: CuRast::draw's sort / group / prefix structure.
"""
import ctypes as C
from fractions import Fraction

import numpy as np
import pytest

import crx
import oracle as orc
import workloads as wl
from crxtest import SceneSpec, oracle_render

L = orc.lib(crx)


def _random_small_scene(rng, n, W, H):
    """n small front-facing triangles in front of an identity camera."""
    tris = []
    for _ in range(n):
        d = rng.uniform(2.0, 6.0)
        cx, cy = rng.uniform(-0.8, 0.8, 2) * d
        size = rng.uniform(0.02, 0.12) * d
        pts = []
        for k in range(3):
            ang = rng.uniform(0, 2 * np.pi)
            pts.append((cx + size * np.cos(ang), cy + size * np.sin(ang), -d + rng.uniform(-0.3, 0.3)))
        # make it counter-clockwise seen from the camera (front-facing)
        (ax, ay, _), (bx, by, _), (qx, qy, _) = pts
        if (bx - ax) * (qy - ay) - (by - ay) * (qx - ax) < 0:
            pts[1], pts[2] = pts[2], pts[1]
        tris.append(pts)
    pos = np.array(tris, dtype=np.float32).reshape(-1, 3)
    idx = np.arange(3 * n, dtype=np.uint32).reshape(-1, 3)
    return pos, idx


def _proj(W, H, fovy=np.radians(60)):
    return wl.projection(W, H, fovy=fovy)


def test_brute_force_tiny_scene_winners():
    rng = np.random.default_rng(42)
    W, H = 48, 40
    pos, idx = _random_small_scene(rng, 60, W, H)
    view = crx.identity()
    proj = _proj(W, H)
    mesh = orc.HostMesh(crx, pos, idx)
    inst = [mesh.cmesh(crx)]
    fb, st = orc.draw(crx, inst, inst, [crx.identity()], W, H, view, proj, True, threads=1)
    f = float(proj.m[5])
    aspect = float(np.float32(W) / np.float32(H))
    # exact screen coordinates of the vertices (same float pipeline up to the
    # projection; coverage and depth are then evaluated exactly / in double)
    P = pos.astype(np.float64)
    d = -P[:, 2]
    sx = (P[:, 0] * (f / aspect) / d * 0.5 + 0.5) * W - 0.5
    sy = (P[:, 1] * f / d * 0.5 + 0.5) * H - 0.5
    checked = 0
    for y in range(H - 1):          # stage 1 never samples the last row / column
        for x in range(W - 1):
            cands = []
            ambiguous = False
            for t in range(len(idx)):
                a, b, c = idx[t]
                e = []
                for (i, j) in ((a, b), (b, c), (c, a)):
                    e.append((sx[j] - sx[i]) * (y - sy[i]) - (sy[j] - sy[i]) * (x - sx[i]))
                area = (sx[b] - sx[a]) * (sy[c] - sy[a]) - (sy[b] - sy[a]) * (sx[c] - sx[a])
                bary = [v / area for v in e]
                if min(bary) > 1e-4:
                    # perspective-correct depth at the sample: 1 / sum(w_i / d_i); the
                    # edge function of (b, c) is the weight of a, (c, a) of b, (a, b) of c
                    wa, wb, wc = bary[1], bary[2], bary[0]
                    inv = wa / d[a] + wb / d[b] + wc / d[c]
                    cands.append((1.0 / inv, t))
                elif min(bary) > -1e-4:
                    ambiguous = True
            if ambiguous:
                continue
            got = int(fb[y * W + x])
            if not cands:
                assert got == crx.CLEAR_PIXEL, (x, y)
                continue
            cands.sort()
            if len(cands) > 1 and cands[1][0] - cands[0][0] < 1e-4 * cands[0][0]:
                continue
            assert got & 0xFFFFFFFFF == cands[0][1], (x, y, cands[:2], got & 0xFFFFFFFFF)
            checked += 1
    assert checked > 100


def test_order_and_thread_independence():
    pos, idx, col = wl.sphere_mesh(rings=64, segments=128, seed=5)
    W, H = 320, 200
    view, proj = wl.sphere_view(W, H)
    spec = SceneSpec().add(orc.HostMesh(crx, pos, idx, col), [crx.identity()])
    fb1, _, st1 = oracle_render(spec, W, H, view, proj, threads=1)
    fb8, _, _ = oracle_render(spec, W, H, view, proj, threads=8)
    assert np.array_equal(fb1, fb8)
    # permuting triangles permutes IDs: compare the per-pixel depth and the
    # winning triangle mapped back through the permutation
    perm = np.random.default_rng(1).permutation(len(idx))
    spec_p = SceneSpec().add(orc.HostMesh(crx, pos, idx[perm], col), [crx.identity()])
    fbp, _, _ = oracle_render(spec_p, W, H, view, proj, threads=4)
    covered = fb1 != crx.CLEAR_PIXEL
    assert np.array_equal(covered, fbp != crx.CLEAR_PIXEL)
    depth1 = fb1 >> np.uint64(36)
    depthp = fbp >> np.uint64(36)
    assert np.array_equal(depth1, depthp)
    ids1 = (fb1 & np.uint64(0xFFFFFFFFF))[covered]
    idsp = perm[(fbp & np.uint64(0xFFFFFFFFF))[covered].astype(np.int64)]
    # equal depths may break ties differently after relabelling; everywhere else the same triangle wins
    same = ids1.astype(np.int64) == idsp
    assert same.mean() > 0.999


def test_instancing_equals_duplicated_meshes():
    pos, idx, col = wl.sphere_mesh(rings=24, segments=48, seed=8)
    worlds = list(wl.instance_transforms(9, extent=6.0, seed=4))
    W, H = 256, 192
    view, proj = wl.instances_view(extent=6.0, width=W, height=H, distance=12.0)
    inst_spec = SceneSpec().add(orc.HostMesh(crx, pos, idx, col), worlds)
    dup_spec = SceneSpec()
    for w in worlds:
        dup_spec.add(orc.HostMesh(crx, pos, idx, col), [w])
    a, _, _ = oracle_render(inst_spec, W, H, view, proj, instanced=True)
    b, _, _ = oracle_render(inst_spec, W, H, view, proj, instanced=False)
    c, _, _ = oracle_render(dup_spec, W, H, view, proj, instanced=True)
    assert np.array_equal(a, b) and np.array_equal(a, c)
    assert (a != crx.CLEAR_PIXEL).sum() > 1000


def test_small_scene_independent_of_tile_size():
    """When every triangle is small, stages 2 and 3 are idle and the tile size
    cannot matter."""
    pos, idx, col = wl.sphere_mesh(rings=32, segments=64, seed=9)
    W, H = 200, 150
    view, proj = wl.sphere_view(W, H)
    spec = SceneSpec().add(orc.HostMesh(crx, pos, idx, col), [crx.identity()])
    outs = []
    for tile in (16, 64, 128):
        t = crx.default_tuning()
        t.tile_size = tile
        fb, _, st = oracle_render(spec, W, H, view, proj, tuning=t)
        outs.append(fb)
    assert all(np.array_equal(outs[0], o) for o in outs[1:])


def test_prepare_frame_sort_group_prefix():
    """Nodes of three meshes: visible instances sorted by triangle count (desc)
    then mesh, instances of one mesh grouped into one unique entry, prefix
    sums and offsets consecutive (CuRast_render.h:338-435)."""
    sizes = {1: 100, 2: 300, 3: 100}
    order = [1, 2, 3, 2, 1, 3, 3, 2, 1, 1]
    n = len(order)
    node_meshes = (crx.CMesh * n)()
    node_world = (crx.Mat4 * n)()
    for k, mid in enumerate(order):
        m = crx.CMesh()
        m.numTriangles = sizes[mid]
        m.address = mid
        m.isLoaded = 1
        m.aabb.min = crx.Vec3(-1, -1, -1)
        m.aabb.max = crx.Vec3(1, 1, 1)
        node_meshes[k] = m
        w = np.eye(4, dtype=np.float32)
        w[0, 3] = k * 0.01
        node_world[k] = crx.Mat4.from_numpy(w)
    view = crx.make_view(crx.look_at((0, 0, 10), (0, 0, 0)), wl.projection(640, 480), 640, 480)
    uniq = (crx.CMesh * n)()
    inst = (crx.CMesh * n)()
    xf = (crx.Mat4 * n)()
    pre = (C.c_uint64 * n)()
    nu = C.c_uint32()
    nv = L.crx_oracle_prepare_frame(node_meshes, node_world, n, C.byref(view), 1, 0, uniq, C.byref(nu), inst, xf, pre)
    assert nv == n
    tri = [inst[k].numTriangles for k in range(nv)]
    addr = [inst[k].address for k in range(nv)]
    assert tri == sorted(tri, reverse=True)
    assert addr == [2, 2, 2, 1, 1, 1, 1, 3, 3, 3]
    assert [pre[k] for k in range(nv)] == list(np.cumsum([0] + tri[:-1]))
    assert [inst[k].instances.offset for k in range(nv)] == list(range(nv))
    assert nu.value == 3
    assert [(uniq[k].address, uniq[k].instances.offset, uniq[k].instances.count) for k in range(3)] == \
        [(2, 0, 3), (1, 3, 4), (3, 7, 3)]
    # instances of one mesh keep node order (world translation grows with node index)
    xs = [inst[k].world.m[12] for k in range(3)]
    assert xs == sorted(xs)
    # transforms are view * world
    for k in range(nv):
        np.testing.assert_array_equal(xf[k].to_numpy(), crx.mat4_mul(view.view, inst[k].world).to_numpy())


def test_prepare_frame_culls_nodes_outside_frustum():
    n = 4
    node_meshes = (crx.CMesh * n)()
    node_world = (crx.Mat4 * n)()
    offsets = [(0, 0, 0), (100, 0, 0), (0, 0, 50), (0, 60, 0)]   # ahead, far right, behind, far up
    for k, o in enumerate(offsets):
        m = crx.CMesh()
        m.numTriangles = 10
        m.address = 1
        m.isLoaded = 1
        m.aabb.min = crx.Vec3(-1, -1, -1)
        m.aabb.max = crx.Vec3(1, 1, 1)
        node_meshes[k] = m
        w = np.eye(4, dtype=np.float32)
        w[:3, 3] = o
        node_world[k] = crx.Mat4.from_numpy(w)
    view = crx.make_view(crx.look_at((0, 0, 10), (0, 0, 0)), wl.projection(640, 480), 640, 480)
    uniq = (crx.CMesh * n)()
    inst = (crx.CMesh * n)()
    xf = (crx.Mat4 * n)()
    pre = (C.c_uint64 * n)()
    nu = C.c_uint32()
    nv = L.crx_oracle_prepare_frame(node_meshes, node_world, n, C.byref(view), 1, 0, uniq, C.byref(nu), inst, xf, pre)
    # the four side planes all pass through the camera: they cull the nodes far
    # to the right, far above and behind the camera (w_clip = -z_view < 0)
    assert nv == 1
    assert (inst[0].world.m[12], inst[0].world.m[13], inst[0].world.m[14]) == (0.0, 0.0, 0.0)
    nv2 = L.crx_oracle_prepare_frame(node_meshes, node_world, n, C.byref(view), 0, 0, uniq, C.byref(nu), inst, xf, pre)
    assert nv2 == 4
