"""Known-answer tests of the CPU oracle, pinned by closed forms.

Each expected value here comes from the mathematics of the case (integer
pixel sets of axis-aligned right triangles, exact dyadic barycentrics, plane
depths, textbook Moller-Trumbore, the bit layout of the packed pixel) and not
from re-typing the oracle's formula, so a dropped term, a swapped operand or a
wrong edge rule in the oracle fails one of them.
"""
import ctypes as C
import math
import struct

import numpy as np
import pytest

import crx
import oracle as orc

L = orc.lib(crx)


def f32(x):
    return struct.unpack("<f", struct.pack("<f", x))[0]


def fbits(x):
    return struct.unpack("<I", struct.pack("<f", x))[0]


def identity_proj(f=1.0):
    """proj with proj[1][1] = f (only f enters the stages) and an infinite far plane."""
    m = np.zeros(16, dtype=np.float32)
    m[0] = f
    m[5] = f
    m[11] = -1.0
    m[14] = 0.01
    return crx.Mat4.from_numpy(m)


def screen_triangle_scene(pts, W, H, depth=1.0):
    """A single triangle whose vertices land exactly on the given -0.5-shifted
    screen coordinates (stage-1 space) with camera = identity, f = aspect = 1
    (W == H), depth `depth`.  Screen s = (x/d * 0.5 + 0.5) * W - 0.5, so
    x = ((s + 0.5) / W - 0.5) * 2 * d, exact for dyadic s and power-of-two W."""
    assert W == H
    verts = []
    for sx, sy in pts:
        x = ((sx + 0.5) / W - 0.5) * 2.0 * depth
        y = ((sy + 0.5) / H - 0.5) * 2.0 * depth
        verts.append((x, y, -depth))
    pos = np.array(verts, dtype=np.float32)
    # exact round trip of the chosen coordinates through the oracle's transform
    idx = np.array([[0, 1, 2]], dtype=np.uint32)
    return pos, idx


def render_one(pos, idx, W, H, tuning=None, colors=None):
    mesh = orc.HostMesh(crx, pos, idx, colors)
    view = crx.identity()
    proj = identity_proj(1.0)
    inst = [mesh.cmesh(crx)]
    fb, st = orc.draw(crx, inst, inst, [crx.mat4_mul(view, crx.identity())], W, H, view, proj, True,
                      tuning=tuning, threads=1)
    return fb, st, mesh, view, proj


def covered(fb, W):
    ids = np.nonzero(fb != crx.CLEAR_PIXEL)[0]
    return {(int(i % W), int(i // W)) for i in ids}


# ---------------------------------------------------------------------------
# pixel packing (rasterization_helpers.cuh:35-66)
# ---------------------------------------------------------------------------

def test_pack_layout_and_clear_value():
    # 28-bit depth in bits 36..63 (sign and 3 low mantissa bits dropped), 36-bit ID below
    p = L.crx_oracle_pack_pixel(f32(1.0), 5)
    assert p == ((0x3F800000 & 0x7FFFFFF8) << 33) | 5
    assert L.crx_oracle_pack_pixel(float("inf"), 0) == 0xFFFFFFF000000000 == crx.CLEAR_PIXEL
    d, i = C.c_float(), C.c_uint64()
    L.crx_oracle_unpack_pixel(crx.CLEAR_PIXEL, C.byref(d), C.byref(i))
    assert math.isinf(d.value) and i.value == 0
    # the three low mantissa bits are dropped; negative depths pack by magnitude
    x = struct.unpack("<f", struct.pack("<I", 0x40490FDB))[0]        # pi
    L.crx_oracle_unpack_pixel(L.crx_oracle_pack_pixel(x, 0xFFFFFFFFF), C.byref(d), C.byref(i))
    assert fbits(d.value) == 0x40490FD8 and i.value == 0xFFFFFFFFF
    assert L.crx_oracle_pack_pixel(-2.0, 1) == L.crx_oracle_pack_pixel(2.0, 1)


def test_pack_is_monotone_in_depth():
    rng = np.random.default_rng(1)
    depths = np.sort(rng.uniform(1e-3, 1e4, 2000).astype(np.float32))
    packed = [L.crx_oracle_pack_pixel(float(d), 7) for d in depths]
    assert all(a <= b for a, b in zip(packed, packed[1:]))


# ---------------------------------------------------------------------------
# GLM evaluation order: closed forms
# ---------------------------------------------------------------------------

def _m(rows):
    return crx.Mat4.from_numpy(np.array(rows, dtype=np.float32))


def test_mat_vec_closed_forms():
    T = _m([[1, 0, 0, 5], [0, 1, 0, -3], [0, 0, 1, 2], [0, 0, 0, 1]])
    out = (C.c_float * 4)()
    L.crx_oracle_mat4_vec4(C.byref(T), (C.c_float * 4)(1, 2, 3, 1), out)
    assert list(out) == [6.0, -1.0, 5.0, 1.0]
    M = _m([[1, 1, 1, 1], [0, 0, 0, 0], [0, 0, 0, 0], [0, 0, 0, 0]])
    L.crx_oracle_mat4_vec4(C.byref(M), (C.c_float * 4)(2.0 ** 24, 1.0, 1.0, -(2.0 ** 24)), out)
    # GLM pairs (x+y) + (z+w): RN(2^24 + 1) + (1 - 2^24) = 2^24 - (2^24 - 1) = 1;
    # left-to-right evaluation would give ((2^24 + 1) + 1) - 2^24 = 0
    assert out[0] == 1.0


def test_mat_mul_and_inverse_closed_forms():
    R = _m([[0, -1, 0, 0], [1, 0, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1]])      # 90 degrees about z
    Tm = _m([[1, 0, 0, 3], [0, 1, 0, 4], [0, 0, 1, 5], [0, 0, 0, 1]])
    out = crx.Mat4()
    L.crx_oracle_mat4_mul(C.byref(Tm), C.byref(R), C.byref(out))
    np.testing.assert_array_equal(out.to_numpy(), Tm.to_numpy() @ R.to_numpy())
    inv = crx.Mat4()
    L.crx_oracle_mat4_inverse(C.byref(out), C.byref(inv))
    # inverse of a rigid transform is exact here (entries 0/+-1 and small integers)
    np.testing.assert_array_equal(inv.to_numpy(), np.linalg.inv(out.to_numpy().astype(np.float64)).astype(np.float32))
    # inverse of a general matrix agrees with double precision to a few ulp
    rng = np.random.default_rng(3)
    A = rng.normal(size=(4, 4)).astype(np.float32) + 4 * np.eye(4, dtype=np.float32)
    L.crx_oracle_mat4_inverse(C.byref(crx.Mat4.from_numpy(A)), C.byref(inv))
    np.testing.assert_allclose(inv.to_numpy(), np.linalg.inv(A.astype(np.float64)), rtol=2e-5, atol=2e-6)


def test_transform_within_one_ulp_of_double_when_exact_terms():
    """view = worldView * p and ndc = (f/a * x / -z, f * y / -z, -z): with
    matrices whose products are exact, every component equals the correctly
    rounded double-precision value (<= 0.5 ulp)."""
    rng = np.random.default_rng(5)
    f, aspect = f32(1.7320508), f32(16 / 9)
    wv = _m([[1, 0, 0, 0.5], [0, 1, 0, -0.25], [0, 0, 1, -10], [0, 0, 0, 1]])
    view, ndc = (C.c_float * 3)(), (C.c_float * 3)()
    for _ in range(200):
        p = rng.integers(-512, 512, 3).astype(np.float32) / 8
        L.crx_oracle_transform(C.byref(wv), (C.c_float * 3)(*p), f, aspect, view, ndc)
        vx, vy, vz = p[0] + 0.5, p[1] - 0.25, p[2] - 10
        assert tuple(view) == (vx, vy, vz)
        d = -vz
        fa = np.float32(f) / np.float32(aspect)
        want_x = np.float32(np.float64(np.float32(fa * np.float32(vx))) / d)
        want_y = np.float32(np.float64(np.float32(np.float32(f) * np.float32(vy))) / d)
        assert ndc[0] == want_x and ndc[1] == want_y and ndc[2] == d


# ---------------------------------------------------------------------------
# stage 1 / stage 2 edge rules on exact right triangles
# ---------------------------------------------------------------------------

def test_stage1_strict_edges_integer_right_triangle():
    """Vertices on sample positions (10,10), (18,10), (10,18): bbox 8x8 = 64
    fragments -> stage 1, which uses strict tests s>0, t>0, v>0: exactly the
    samples with x > 10, y > 10, x + y < 28 are covered."""
    W = H = 64
    pos, idx = screen_triangle_scene([(10, 10), (18, 10), (10, 18)], W, H)
    fb, st, *_ = render_one(pos, idx, W, H)
    assert st.num_nontrivial == 0
    want = {(x, y) for x in range(64) for y in range(64) if x > 10 and y > 10 and x + y < 28}
    assert covered(fb, W) == want and len(want) == 21


def test_stage1_dyadic_offsets_right_triangle():
    """Vertices off the sample grid at (10.25,10.25), (20.25,10.25), (10.25,20.25):
    covered samples are x >= 11, y >= 11, x + y <= 30 (45 samples)."""
    W = H = 64
    pos, idx = screen_triangle_scene([(10.25, 10.25), (20.25, 10.25), (10.25, 20.25)], W, H)
    fb, st, *_ = render_one(pos, idx, W, H)
    want = {(x, y) for x in range(64) for y in range(64) if x >= 11 and y >= 11 and x + y <= 30}
    assert covered(fb, W) == want and len(want) == 45


def test_stage2_inclusive_edges_integer_right_triangle():
    """Legs of 16 samples: bbox 16x16 = 256 > 128 fragments -> stage 2, whose
    tests are inclusive (s,t,v >= 0): x >= 10, y >= 10, x + y <= 36 (153)."""
    W = H = 64
    pos, idx = screen_triangle_scene([(10, 10), (26, 10), (10, 26)], W, H)
    fb, st, *_ = render_one(pos, idx, W, H)
    assert st.num_nontrivial == 1 and st.num_huge == 0
    want = {(x, y) for x in range(64) for y in range(64) if x >= 10 and y >= 10 and x + y <= 36}
    assert covered(fb, W) == want and len(want) == 153


def test_stage1_never_samples_last_column_or_row():
    """Reference quirk kept: stage 1 clips the box to W-1 and iterates up to
    ceil(max)-1, so column W-1 and row H-1 stay empty for small triangles."""
    W = H = 32
    pos, idx = screen_triangle_scene([(26.5, 26.5), (34.5, 26.5), (26.5, 34.5)], W, H)
    fb, st, *_ = render_one(pos, idx, W, H)
    cov = covered(fb, W)
    assert cov and all(x < W - 1 and y < H - 1 for x, y in cov)
    want = {(x, y) for x in range(27, W - 1) for y in range(27, H - 1) if x + y < 61}
    assert cov == want


def test_backface_is_culled_and_winding_matters():
    W = H = 64
    pos, idx = screen_triangle_scene([(10, 10), (18, 10), (10, 18)], W, H)
    fb, *_ = render_one(pos, idx[:, ::-1].copy(), W, H)
    assert not covered(fb, W)


def test_depth_of_fronto_parallel_triangle():
    W = H = 64
    for depth in (0.5, 3.0, 1000.0):
        pos, idx = screen_triangle_scene([(10.25, 10.25), (20.25, 10.25), (10.25, 20.25)], W, H, depth)
        fb, *_ = render_one(pos, idx, W, H)
        d = C.c_float()
        i = C.c_uint64()
        for px in np.nonzero(fb != crx.CLEAR_PIXEL)[0]:
            L.crx_oracle_unpack_pixel(int(fb[px]), C.byref(d), C.byref(i))
            assert abs(d.value - depth) <= depth * 2e-6 and i.value == 0


def test_stage3_huge_triangle_matches_exact_ray_cast():
    """A triangle spanning most of a 256x256 view goes to stage 3 (bbox >=
    64^2); away from its edges every pixel centre is covered, outside none is,
    and the depth is the plane's (fronto-parallel at depth 2)."""
    W = H = 256
    pts = [(20.0, 20.0), (230.0, 30.0), (40.0, 220.0)]
    pos, idx = screen_triangle_scene(pts, W, H, depth=2.0)
    fb, st, *_ = render_one(pos, idx, W, H)
    assert st.num_nontrivial == 1 and st.num_huge > 0
    cov = covered(fb, W)
    (ax, ay), (bx, by), (cx, cy) = pts

    def edge(px, py, x0, y0, x1, y1):
        return (x1 - x0) * (py - y0) - (y1 - y0) * (px - x0)

    area = edge(cx, cy, ax, ay, bx, by)
    for y in range(H):
        for x in range(W):
            # pixel (x, y) is sampled at shifted-space (x, y); e_i / area are its barycentrics
            b = [edge(x, y, ax, ay, bx, by) / area, edge(x, y, bx, by, cx, cy) / area,
                 edge(x, y, cx, cy, ax, ay) / area]
            if min(b) > 0.01:
                assert (x, y) in cov
            elif min(b) < -0.01:
                assert (x, y) not in cov
    d = C.c_float()
    i = C.c_uint64()
    for px in list(np.nonzero(fb != crx.CLEAR_PIXEL)[0])[::97]:
        L.crx_oracle_unpack_pixel(int(fb[px]), C.byref(d), C.byref(i))
        assert abs(d.value - 2.0) < 2e-5


def test_moller_trumbore_textbook_cases():
    tri = [(C.c_float * 3)(-1, -1, -5), (C.c_float * 3)(1, -1, -5), (C.c_float * 3)(0, 1, -5)]
    t = L.crx_oracle_intersect_mt((C.c_float * 3)(0, 0, -1), *tri)
    assert t == 5.0
    assert math.isinf(L.crx_oracle_intersect_mt((C.c_float * 3)(0, 0, 1), *tri))      # behind
    assert math.isinf(L.crx_oracle_intersect_mt((C.c_float * 3)(1, 0, 0), *tri))      # parallel
    d = np.array([0.9, 0.0, -5.0]) / np.linalg.norm([0.9, 0.0, -5.0])
    assert math.isinf(L.crx_oracle_intersect_mt((C.c_float * 3)(*d.astype(np.float32)), *tri))  # outside


def test_stage2_bbox_contains_the_clipped_triangle():
    """The Sutherland-Hodgman box of stage 2 against a dense sampling of the
    view-space triangle clipped to the near plane and the screen."""
    rng = np.random.default_rng(11)
    f, aspect, W, H = f32(1.0), f32(1.0), 200.0, 200.0
    box = (C.c_float * 4)()
    for _ in range(200):
        v = rng.uniform(-3, 3, (3, 3)).astype(np.float32)
        v[:, 2] = rng.uniform(-4, 1, 3)
        if (v[:, 2] >= -0.01).all():
            continue
        L.crx_oracle_stage2_bbox((C.c_float * 3)(*v[0]), (C.c_float * 3)(*v[1]), (C.c_float * 3)(*v[2]),
                                 f, aspect, W, H, box)
        mn_x, mx_x, mn_y, mx_y = box
        # sample the triangle, keep points in front of the near plane and on screen
        w = rng.dirichlet([1, 1, 1], 4000)
        p = w @ v.astype(np.float64)
        p = p[p[:, 2] < -0.0101]
        if len(p) == 0:
            continue
        d = -p[:, 2]
        sx = (p[:, 0] / d * 0.5 + 0.5) * W
        sy = (p[:, 1] / d * 0.5 + 0.5) * H
        on = (sx >= 0) & (sx <= W) & (sy >= 0) & (sy <= H)
        if not on.any():
            continue
        assert (sx[on] >= mn_x - 1e-3).all() and (sx[on] <= mx_x + 1e-3).all()
        assert (sy[on] >= mn_y - 1e-3).all() and (sy[on] <= mx_y + 1e-3).all()


# ---------------------------------------------------------------------------
# BitEdit (src/BitEdit.h:8-49)
# ---------------------------------------------------------------------------

def test_bitedit_round_trip_brute_force():
    rng = np.random.default_rng(2)
    for bits in range(1, 33):
        buf = np.zeros(8, dtype=np.uint32)
        vals = []
        for k in range((8 * 32) // bits - 1):
            v = int(rng.integers(0, 1 << bits))
            L.crx_oracle_bitedit_write(buf.ctypes.data, k * bits, bits, v)
            vals.append(v)
        for k, v in enumerate(vals):
            assert L.crx_oracle_bitedit_read(buf.ctypes.data, k * bits, bits) == v
        # the whole buffer equals the little-endian concatenation of the values
        acc = 0
        for k, v in enumerate(vals):
            acc |= v << (k * bits)
        assert int.from_bytes(buf.tobytes(), "little") & ((1 << (len(vals) * bits)) - 1) == acc
    assert L.crx_oracle_bitedit_read(np.zeros(2, np.uint32).ctypes.data, 0, 0) == 0
    assert L.crx_oracle_bitedit_read(np.zeros(2, np.uint32).ctypes.data, 0, 33) == 0


# ---------------------------------------------------------------------------
# resolve (resolve.cu:640-1141)
# ---------------------------------------------------------------------------

def test_resolve_constant_and_blended_vertex_colors():
    W = H = 64
    pts = [(10.0, 10.0), (42.0, 10.0), (10.0, 42.0)]   # 32 legs -> stage 2 (inclusive)
    pos, idx = screen_triangle_scene(pts, W, H, depth=2.0)
    red, green, blue = 0xFF0000FF, 0xFF00FF00, 0xFFFF0000
    cols = np.array([red, green, blue], dtype=np.uint32)
    fb, st, mesh, view, proj = render_one(pos, idx, W, H, colors=cols)
    inst = [mesh.cmesh(crx)]
    cb = orc.resolve(crx, inst, [0], fb, W, H, view, proj, crx.DISPLAY_VERTEX_COLORS, threads=1)
    rgba = (cb & 0xFFFFFFFF).astype(np.uint32).view(np.uint8).reshape(H, W, 4)
    # sample (x, y) sits at barycentrics s = (x-10)/32 (towards B), t = (y-10)/32 (towards C)
    for (x, y) in [(10, 10), (26, 10), (10, 26), (18, 18), (20, 14)]:
        s, t = (x - 10) / 32, (y - 10) / 32
        want = np.array([255 * (1 - s - t), 255 * s, 255 * t])
        got = rgba[y, x, :3].astype(float)
        assert np.all(np.abs(got - want) <= 1.0 + 1e-9), ((x, y), got, want)
        assert rgba[y, x, 3] == 255
        assert abs(struct.unpack("<f", struct.pack("<I", int(cb[y * W + x]) >> 32))[0] - 2.0) < 1e-5
    # background: depth +inf, colour 0
    assert int(cb[0]) == 0x7F80000000000000


def test_resolve_triangle_id_recovers_linear_scan():
    """TRIANGLE_ID colours encode tri*12345678: recover tri for each pixel and
    compare with a linear scan of the instance prefix on the visbuffer ID."""
    import workloads as wl
    pos, idx, col = wl.sphere_mesh(rings=16, segments=32)
    mesh = orc.HostMesh(crx, pos, idx, col)
    worlds = list(wl.instance_transforms(5, extent=4.0, seed=3))
    view, proj = wl.instances_view(extent=4.0, width=160, height=120, distance=9.0)
    inst = [mesh.cmesh(crx, cumm=k * mesh.num_triangles, offset=k, world=w) for k, w in enumerate(worlds)]
    uniq = [mesh.cmesh(crx, cumm=0, offset=0, count=len(worlds))]
    xf = [crx.mat4_mul(view, w) for w in worlds]
    fb, _ = orc.draw(crx, uniq, inst, xf, 160, 120, view, proj, True, threads=1)
    prefix = [k * mesh.num_triangles for k in range(len(worlds))]
    cb = orc.resolve(crx, inst, prefix, fb, 160, 120, view, proj, crx.DISPLAY_TRIANGLE_ID, threads=1)
    checked = 0
    for px in np.nonzero(fb != crx.CLEAR_PIXEL)[0]:
        tid = int(fb[px]) & 0xFFFFFFFFF
        k = max(j for j in range(len(prefix)) if prefix[j] <= tid)
        tri = tid - prefix[k]
        color = int(cb[px]) & 0xFFFFFFFF
        assert color == ((tri * 12345678) & 0xFFFFFFFF) | 0xFF000000
        checked += 1
    assert checked > 100


# ---------------------------------------------------------------------------
# TEXTURE (resolve.cu:148-195, 783-882)
# ---------------------------------------------------------------------------

def _textured_triangle(W, gx, gy, texture, size, jpeg=False, uvs=True, offset=(0.0, 0.0)):
    """The screen triangle A=(10,10), B=(42,10), C=(10,42) at depth 2 with
    uv_a = offset, uv_b = offset + (gx, 0), uv_c = offset + (0, gy): pixel
    (x, y) then has uv = offset + (gx (x-10)/32, gy (y-10)/32), and the next
    pixel right / down adds gx/32 / gy/32."""
    pos, idx = screen_triangle_scene([(10.0, 10.0), (42.0, 10.0), (10.0, 42.0)], W, W, depth=2.0)
    uv = np.array([[0.0, 0.0], [gx, 0.0], [0.0, gy]], dtype=np.float32) + np.float32(offset)
    mesh = orc.HostMesh(crx, pos, idx, None, uvs=uv if uvs else None, texture=texture, texture_size=size,
                        jpeg=jpeg)
    view, proj = crx.identity(), identity_proj(1.0)
    inst = [mesh.cmesh(crx)]
    fb, _ = orc.draw(crx, inst, inst, [crx.identity()], W, W, view, proj, True, threads=1)
    cb = orc.resolve(crx, inst, [0], fb, W, W, view, proj, crx.DISPLAY_TEXTURE, threads=1)
    return fb, cb


def _interior(W):
    """Covered pixels well inside the triangle (s, t >= 1/32, s + t <= 30/32)."""
    return [(x, y) for y in range(11, 40) for x in range(11, 40) if (x - 10) + (y - 10) <= 30]


def _rgba(c):
    return [(int(c) >> (8 * k)) & 0xFF for k in range(4)]


def test_texture_mip_level_of_the_uv_footprint():
    """Each level of a 100 x 60 chain is one colour.  The footprint of a pixel
    is gx/32 by gy/32 in uv, so m = (max(gx, gy)/32)^2 (100^2 + 60^2); g is
    chosen so that 0.5 log2 m = j + phase: level j, clamped to [0, 7].  The
    phases 0.04 and 0.96 put m within 6% of a level boundary, so a dropped or
    misplaced term of m moves the level.  Level k is ((w + 1) / 2) x
    ((h + 1) / 2) of level k - 1, so a wrong offset into the chain reads
    another level's colour."""
    import crxtest
    shapes = crxtest.mip_shapes(100, 60)
    assert shapes[3] == (8, 13) and shapes[7] == (1, 1)
    cols = [0xFF000000 | ((7 * k + 3) << 16) | ((200 - 20 * k) << 8) | (20 + 25 * k) for k in range(8)]
    chain = crxtest.mip_chain([np.full(s, cols[k], dtype=np.uint32) for k, s in enumerate(shapes)])
    W = 64
    for j, phase in [(j, p) for j in (-2, 0, 1, 2, 3, 5, 7, 9) for p in (0.5, 0.04, 0.96)]:
        g = 32.0 * math.sqrt(4.0 ** (j + phase) / (100 ** 2 + 60 ** 2))
        # the longer footprint side decides: along u for even j, along v for odd j
        gx, gy = (g, g / 1.5) if j % 2 == 0 else (g / 1.5, g)
        fb, cb = _textured_triangle(W, gx, gy, chain, (100, 60))
        want = _rgba(cols[min(max(j, 0), 7)])
        for (x, y) in _interior(W)[::7]:
            got = _rgba(cb[y * W + x])
            assert got[3] == 255
            assert all(abs(a - b) <= 1 for a, b in zip(got[:3], want[:3])), (j, phase, (x, y), got, want)


@pytest.mark.parametrize("level, gx, gy, offset", [(0, 0.194, 0.15, (0.0, 0.0)),
                                                   (2, 0.776, 0.6, (3.0, -2.0))])
def test_texture_bilinear_on_a_gradient(level, gx, gy, offset):
    """Level k holds red = texel column, green = texel row, blue = 40 k.  The
    four texels at ftx -+ 0.5, fty -+ 0.5 weighted by the fractions of
    ftx + 0.5 and fty + 0.5 give red = ftx - 0.5 and green = fty - 0.5 exactly
    (ftx = frac(u) * level width), then truncated to 8 bits.  gx != gy keeps
    the x and y weights apart, so swapped texels or weights fail; the
    offset (3, -2) checks the wrap of uv into [0, 1)."""
    import crxtest
    shapes = crxtest.mip_shapes(200, 120)
    levels = []
    for k, (h, w) in enumerate(shapes):
        ty, tx = np.mgrid[0:h, 0:w].astype(np.uint32)
        levels.append(0xFF000000 | ((40 * k) << 16) | (ty << 8) | tx)
    chain = crxtest.mip_chain(levels)
    W = 64
    # the footprint picks the level: m = (gx/32)^2 (200^2 + 120^2) in [4^k, 4^k * 4)
    m = (max(gx, gy) / 32.0) ** 2 * (200 ** 2 + 120 ** 2)
    assert int(0.5 * math.log2(m)) == level
    fb, cb = _textured_triangle(W, gx, gy, chain, (200, 120), offset=offset)
    lh, lw = shapes[level]
    checked = 0
    for (x, y) in _interior(W):
        ftx = gx * (x - 10) / 32.0 * lw
        fty = gy * (y - 10) / 32.0 * lh
        if ftx < 0.6 or fty < 0.6:
            continue
        got = _rgba(cb[y * W + x])
        for ch, f in ((0, ftx), (1, fty)):
            assert got[ch] in (math.floor(f - 0.5 - 1e-3), math.floor(f - 0.5 + 1e-3)), ((x, y), ch, got, f)
        assert abs(got[2] - 40 * level) <= 1 and got[3] == 255
        checked += 1
    assert checked > 200


@pytest.mark.parametrize("g", [0.2, 3.0])
def test_texture_jpeg_missing_data_and_missing_uvs(g):
    """A JPEG-coded texture leaves the pixel magenta (the decoder's pass is
    not part of the resolve); no texture data samples 0, shaded opaque, at
    level 0 (g = 0.2) and at level 3 (g = 3: m = (3/32)^2 2 64^2 = 72); a
    mesh without uvs takes the final branch of the attribute chain (blue)."""
    W = 64
    chain = np.full(64 * 64 * 2, 0xFF123456, dtype=np.uint32)
    cases = [(dict(texture=chain, jpeg=True), 0xFFFF00FF),
             (dict(texture=None), 0xFF000000),
             (dict(texture=chain, uvs=False), 0xFF0000FF)]
    for kw, want in cases:
        tex = kw.pop("texture")
        fb, cb = _textured_triangle(W, g, g, tex, (64, 64), **kw)
        cov = np.nonzero(fb != crx.CLEAR_PIXEL)[0]
        assert cov.size > 400
        assert np.all((cb[cov] & 0xFFFFFFFF) == want), (kw, hex(int(cb[cov[0]])))
        depth = (cb[cov] >> np.uint64(32)).astype(np.uint32).view(np.float32)
        assert np.all(np.abs(depth - 2.0) < 1e-5)
        assert int(cb[0]) == 0x7F80000000000000


# ---------------------------------------------------------------------------
# mip chain (textureTools.cu:41-133)
# ---------------------------------------------------------------------------

def _numpy_chain(level0, last):
    """Levels 0..last by whole-array 2x2 block sums: each level zero-padded
    to even sides, reshaped into blocks, summed per channel, // 4."""
    levels = [level0]
    cur = level0
    for k in range(1, last + 1):
        h, w = cur.shape
        pad = np.zeros((h + h % 2, w + w % 2), dtype=np.uint32)
        pad[:h, :w] = cur
        ch = [((pad >> np.uint32(8 * c)) & np.uint32(0xFF)).astype(np.int64) for c in range(3)]
        sums = [x.reshape(pad.shape[0] // 2, 2, pad.shape[1] // 2, 2).sum(axis=(1, 3)) // 4 for x in ch]
        alpha = 0 if k <= 9 else 0xFF
        cur = (np.uint32(alpha << 24) | (sums[2].astype(np.uint32) << np.uint32(16)) |
               (sums[1].astype(np.uint32) << np.uint32(8)) | sums[0].astype(np.uint32))
        levels.append(cur)
    return levels


def _oracle_chain(level0):
    h, w = level0.shape
    n = crx.texture_chain_texels(w, h)
    buf = np.full(n + 16, 0xDEADBEEF, dtype=np.uint32)
    buf[:w * h] = level0.reshape(-1)
    last = L.crx_oracle_compute_mipmap(buf.ctypes.data, w, h)
    assert np.all(buf[n:] == 0xDEADBEEF)      # nothing past the chain
    return buf[:n], last


@pytest.mark.parametrize("w, h, last", [(1, 1, 7), (1, 7, 7), (5, 3, 7), (64, 64, 7), (100, 60, 7), (255, 2, 7),
                                        (256, 9, 7), (1024, 3, 9), (4096, 3, 11), (37, 300, 7)])
def test_mip_chain_is_the_2x2_block_average(w, h, last):
    """Levels 1 .. max(floor(log2(max(w, h))) - 1, 7), each a 2x2 block sum
    // 4 of the level before with missing texels as 0; alpha 0 through level
    9, 255 beyond; the chain's size is the sum of its levels."""
    import crxtest
    rng = np.random.default_rng(w * 7919 + h)
    level0 = rng.integers(0, 2 ** 32, size=(h, w), dtype=np.uint32)
    chain, got_last = _oracle_chain(level0)
    assert got_last == last
    shapes = crxtest.mip_shapes(w, h, last + 1)
    assert chain.size == sum(a * b for a, b in shapes)
    want = np.concatenate([l.reshape(-1) for l in _numpy_chain(level0, last)])
    assert np.array_equal(chain, want)


def test_mip_chain_closed_forms():
    """A constant colour c keeps c on even sides and becomes c/2 on the odd
    edge and c/4 in the odd corner (missing texels add 0, the divisor stays
    4); a red ramp r = x gives r = 2x on level 1 ((2x + 2x + 1) * 2 // 4)."""
    c = 0x00C8A064            # r = 100, g = 160, b = 200
    chain, _ = _oracle_chain(np.full((3, 5), c, dtype=np.uint32))
    lvl1 = chain[15:15 + 6].reshape(2, 3)
    rgb = lambda t: [(int(t) >> s) & 0xFF for s in (0, 8, 16)]
    assert rgb(lvl1[0, 0]) == [100, 160, 200] and rgb(lvl1[0, 1]) == [100, 160, 200]
    assert rgb(lvl1[0, 2]) == [50, 80, 100] and rgb(lvl1[1, 0]) == [50, 80, 100]
    assert rgb(lvl1[1, 2]) == [25, 40, 50]
    assert all(int(t) >> 24 == 0 for t in lvl1.reshape(-1))
    ramp = np.tile(np.arange(256, dtype=np.uint32), (4, 1))
    chain, _ = _oracle_chain(ramp)
    lvl1 = chain[1024:1024 + 256].reshape(2, 128)
    assert np.array_equal(lvl1 & 0xFF, np.tile(2 * np.arange(128, dtype=np.uint32), (2, 1)))


def test_resolve_vertex_normals():
    """NORMAL with per-vertex normals (resolve.cu:927-943): the normals are
    turned by world (w = 0), blended with the pixel's barycentrics,
    normalised and written as clamp(255 N, 0, 255) per channel.  Against the
    same steps in float64 at exact dyadic barycentrics (+-1 LSB); negative
    components clamp to 0."""
    W = H = 64
    pos, idx = screen_triangle_scene([(10.0, 10.0), (42.0, 10.0), (10.0, 42.0)], W, H, depth=2.0)
    nrm = np.array([[0.6, 0.8, 0.0], [0.0, 0.6, 0.8], [-0.8, 0.0, 0.6]], dtype=np.float32)
    mesh = orc.HostMesh(crx, pos, idx, None, normals=nrm)
    view, proj = crx.identity(), identity_proj(1.0)
    inst = [mesh.cmesh(crx)]
    fb, _ = orc.draw(crx, inst, inst, [crx.identity()], W, H, view, proj, True, threads=1)
    cb = orc.resolve(crx, inst, [0], fb, W, H, view, proj, crx.DISPLAY_NORMAL, threads=1)
    rgba = (cb & 0xFFFFFFFF).astype(np.uint32).view(np.uint8).reshape(H, W, 4)
    n64 = nrm.astype(np.float64)
    for (x, y) in [(10, 10), (26, 10), (10, 26), (18, 18), (20, 14), (12, 30)]:
        s, t = (x - 10) / 32, (y - 10) / 32
        N = n64[0] * (1 - s - t) + n64[1] * s + n64[2] * t
        N = N / np.linalg.norm(N)
        want = np.clip(N * 255.0, 0.0, 255.0)
        got = rgba[y, x, :3].astype(float)
        assert np.all(np.abs(got - want) <= 1.0 + 1e-9), ((x, y), got, want)
        assert rgba[y, x, 3] == 255
    # a scaled, rotated world turns the normals before they are blended; the
    # positions are stored turned back (exactly: a quarter turn and a power
    # of two), so the pixels keep their barycentrics
    w = crx.identity()
    w.m[0], w.m[1], w.m[4], w.m[5] = 0.0, 2.0, -2.0, 0.0     # 90 degrees about z, scale 2
    pos2 = np.stack([pos[:, 1] * 0.5, -pos[:, 0] * 0.5, pos[:, 2]], axis=1).astype(np.float32)
    mesh2 = orc.HostMesh(crx, pos2, idx, None, normals=nrm)
    inst2 = [mesh2.cmesh(crx, world=w)]
    fb2, _ = orc.draw(crx, inst2, inst2, [w], W, H, view, proj, True, threads=1)
    assert np.array_equal(fb2, fb)
    cb2 = orc.resolve(crx, inst2, [0], fb2, W, H, view, proj, crx.DISPLAY_NORMAL, threads=1)
    rgba2 = (cb2 & 0xFFFFFFFF).astype(np.uint32).view(np.uint8).reshape(H, W, 4)
    R = np.array([[0.0, -2.0, 0.0], [2.0, 0.0, 0.0], [0.0, 0.0, 1.0]])
    for (x, y) in [(10, 10), (20, 14)]:
        s, t = (x - 10) / 32, (y - 10) / 32
        N = (R @ n64[0]) * (1 - s - t) + (R @ n64[1]) * s + (R @ n64[2]) * t
        N = N / np.linalg.norm(N)
        want = np.clip(N * 255.0, 0.0, 255.0)
        assert np.all(np.abs(rgba2[y, x, :3].astype(float) - want) <= 1.0 + 1e-9), ((x, y), rgba2[y, x], want)


def test_resolve_stage_by_box_size_and_near_plane():
    """STAGE (resolve.cu:974-1031): the triangle's screen box (ndcToScreen
    without the half-pixel shift, so a stage-1 screen point s lands on
    s + 0.5) clipped to the screen, ceil(max) - floor(min) per axis: at most
    128 fragments -> SCHEME_SPECTRAL[9], at most 4096 -> [7], more -> [3];
    a vertex at or behind the camera plane -> [3]."""
    cases = [  # (W, legs from (10, 10) in stage-1 screen units, expected box side, colour)
        (64, 4.0, 5, 0xFFBD8832),      # ceil(14.5) - floor(10.5) = 5: 25 fragments
        (64, 10.0, 11, 0xFFBD8832),    # 121
        (64, 11.0, 12, 0xFFA4DDAB),    # 144
        (128, 63.0, 64, 0xFFA4DDAB),   # 4096
        (128, 64.0, 65, 0xFF61AEFD),   # 4225
    ]
    for W, leg, side, want in cases:
        assert side * side <= 128 or want != 0xFFBD8832
        pos, idx = screen_triangle_scene([(10.0, 10.0), (10.0 + leg, 10.0), (10.0, 10.0 + leg)], W, W, depth=2.0)
        fb, st, mesh, view, proj = render_one(pos, idx, W, W)
        cb = orc.resolve(crx, [mesh.cmesh(crx)], [0], fb, W, W, view, proj, crx.DISPLAY_STAGE, threads=1)
        cov = np.nonzero(fb != crx.CLEAR_PIXEL)[0]
        assert cov.size > 0
        assert np.all((cb[cov] & 0xFFFFFFFF) == want), (W, leg, hex(int(cb[cov[0]]) & 0xFFFFFFFF))
    # a triangle reaching behind the camera: stage 2 whatever its box
    W = 64
    pos = np.array([[-0.5, -0.5, -2.0], [0.5, -0.5, -2.0], [0.0, 0.5, 1.0]], dtype=np.float32)
    idx = np.array([[0, 1, 2]], dtype=np.uint32)
    fb, st, mesh, view, proj = render_one(pos, idx, W, W)
    cb = orc.resolve(crx, [mesh.cmesh(crx)], [0], fb, W, W, view, proj, crx.DISPLAY_STAGE, threads=1)
    cov = np.nonzero(fb != crx.CLEAR_PIXEL)[0]
    assert cov.size > 0 and np.all((cb[cov] & 0xFFFFFFFF) == 0xFF61AEFD)
