"""Bit-exact parity of the IX_PU16 (bit-packed index, 3 x uint16 position)
path with the oracle: stage 1 (plain and instanced walks), stages 2/3 on
close-ups, resolve, multi-view and the scene API."""
import numpy as np
import pytest

import crx
import oracle as orc
import workloads as wl
from crxtest import CompressedMesh, GpuScene, SceneSpec, describe_mismatch, oracle_render

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = crx.Context(0)
    yield c
    c.close()


def _check(ctx, spec, W, H, view, proj, instanced=True, attribute=crx.DISPLAY_VERTEX_COLORS):
    print(f"[case] compressed W={W} H={H} instanced={instanced}", flush=True)
    want_fb, want_cb, st = oracle_render(spec, W, H, view, proj, instanced=instanced, attribute=attribute)
    fb, cb, gst = GpuScene(ctx, spec).render(W, H, view, proj, instanced=instanced, attribute=attribute)
    assert (gst.num_nontrivial, gst.num_huge) == (st.num_nontrivial, st.num_huge)
    assert np.array_equal(fb, want_fb), describe_mismatch(fb, want_fb, W)
    if attribute is not None:
        assert np.array_equal(cb, want_cb), describe_mismatch(cb, want_cb, W)
    return fb, st


def test_sphere_1m_compressed_1080p(ctx):
    pos, idx, col = wl.sphere_mesh()
    spec = SceneSpec().add(CompressedMesh(pos, idx, col), [crx.identity()])
    view, proj = wl.sphere_view(1920, 1080)
    fb, _ = _check(ctx, spec, 1920, 1080, view, proj)
    assert int((fb != crx.CLEAR_PIXEL).sum()) > 300_000


def test_instanced_mirrored_compressed(ctx):
    pos, idx, col = wl.sphere_mesh(rings=48, segments=96, seed=7)
    worlds = list(wl.instance_transforms(40, extent=20.0, smin=0.5, smax=2.0, seed=11))
    for k in range(0, 40, 5):
        for r in range(3):
            worlds[k].m[r] = -worlds[k].m[r]
    spec = SceneSpec().add(CompressedMesh(pos, idx, col), worlds)
    view, proj = wl.instances_view(extent=20.0, width=800, height=600, distance=30.0)
    for instanced in (True, False):
        _check(ctx, spec, 800, 600, view, proj, instanced=instanced)


def test_instanced_compressed_two_per_lane():
    """IX_PU16 stage 1 with each lane taking two triangles per batch."""
    t = crx.default_tuning()
    t.stage1_tris_per_lane = 2
    c = crx.Context(0, tuning=t)
    try:
        pos, idx, col = wl.sphere_mesh(rings=47, segments=95, seed=8)
        worlds = list(wl.instance_transforms(24, extent=20.0, smin=0.5, smax=2.0, seed=12))
        spec = SceneSpec().add(CompressedMesh(pos, idx, col), worlds)
        view, proj = wl.instances_view(extent=20.0, width=800, height=600, distance=30.0)
        for instanced in (True, False):
            _check(c, spec, 800, 600, view, proj, instanced=instanced)
    finally:
        c.close()


@pytest.mark.parametrize("per_lane", [1, 2, 4])
def test_small_compressed_meshes(ctx, per_lane):
    """Edge cases of the stage-1 IX_PU16 reads: meshes of one and two
    triangles (a stream shorter than four words, fewer than three vertices
    read directly), an odd vertex count whose even last vertex is read as
    three uint16, and 1-bit indices."""
    t = crx.default_tuning()
    t.stage1_tris_per_lane = per_lane
    c = crx.Context(0, tuning=t)
    try:
        tri = (np.array([[-1, -1, -3], [1, -1, -3], [0, 1, -3]], dtype=np.float32), np.array([[0, 1, 2]], dtype=np.uint32))
        quad = (np.array([[-2, -1, -4], [2, -1, -4], [2, 1, -4], [-2, 1, -4], [0, 2, -5]], dtype=np.float32),
                np.array([[0, 1, 2], [0, 2, 3], [3, 2, 4]], dtype=np.uint32))
        two = (np.array([[-1, -1, -2], [1, 1, -2]], dtype=np.float32), np.array([[0, 1, 1], [1, 0, 0]], dtype=np.uint32))
        hp, hi, hc = wl.heightfield_mesh(33, amplitude=2.0, scale=8.0)          # 1089 vertices: odd
        spec = SceneSpec()
        for p_, i_ in (tri, quad, two):
            col = np.arange(len(p_), dtype=np.uint32) * 0x01020304 + 0xFF000000
            spec.add(CompressedMesh(p_, i_, col), [crx.identity()])
        shift = np.array([[1, 0, 0, -16], [0, 1, 0, -3], [0, 0, 1, -40], [0, 0, 0, 1]], dtype=np.float32)
        spec.add(CompressedMesh(hp, hi, hc), [crx.Mat4.from_numpy(shift)])
        view = crx.look_at((0.0, 0.0, 2.0), (0.0, 0.0, -10.0))
        for instanced in (True, False):
            fb, _ = _check(c, spec, 640, 480, view, wl.projection(640, 480), instanced=instanced)
            assert int((fb != crx.CLEAR_PIXEL).sum()) > 1000
    finally:
        c.close()


def test_closeup_compressed_stage2_stage3(ctx):
    pos, idx, col = wl.heightfield_mesh(129, spacing=1.0, amplitude=6.0, scale=32.0)
    spec = SceneSpec().add(CompressedMesh(pos, idx, col), [crx.identity()])
    W, H = 1280, 720
    view = crx.look_at((64.0, 4.5, 20.0), (64.0, 2.0, 40.0), (0.0, 1.0, 0.0))
    _, st = _check(ctx, spec, W, H, view, wl.projection(W, H))
    assert st.num_nontrivial > 0 and st.num_huge > 0


@pytest.mark.parametrize("attribute", [crx.DISPLAY_TRIANGLE_ID, crx.DISPLAY_NORMAL])
def test_compressed_resolve_attributes(ctx, attribute):
    pos, idx, col = wl.sphere_mesh(rings=64, segments=128)
    spec = SceneSpec().add(CompressedMesh(pos, idx, col), list(wl.instance_transforms(5, extent=4.0, seed=3)))
    view, proj = wl.instances_view(extent=4.0, width=640, height=480, distance=9.0)
    _check(ctx, spec, 640, 480, view, proj, attribute=attribute)


def test_compressed_multiview(ctx):
    n = 257
    pos, idx, col = wl.heightfield_mesh(n, amplitude=20.0, scale=64.0)
    spec = SceneSpec().add(CompressedMesh(pos, idx, col), [crx.identity()])
    vs = wl.multiview_views(n, pos, count=6, sizes=((256, 256), (320, 180), (200, 200)), seed=17)
    views = [(W, H, v, p) for (v, p, W, H, _) in vs]
    g = GpuScene(ctx, spec)
    fbs, cbs, _ = g.render_views(views, attribute=crx.DISPLAY_VERTEX_COLORS)
    for k, (W, H, v, p) in enumerate(views):
        want_fb, want_cb, _ = oracle_render(spec, W, H, v, p, attribute=crx.DISPLAY_VERTEX_COLORS)
        assert np.array_equal(fbs[k], want_fb), f"view {k}\n" + describe_mismatch(fbs[k], want_fb, W)
        assert np.array_equal(cbs[k], want_cb), f"view {k} colours"


def test_scene_api_compressed_mesh(ctx):
    pos, idx, col = wl.sphere_mesh(rings=32, segments=64, seed=21)
    mesh = crx.Mesh(ctx, pos, idx, col, compressed=True)
    scene = crx.Scene(ctx)
    worlds = list(wl.instance_transforms(6, extent=10.0, seed=23))
    for w in worlds:
        scene.add(mesh, w)
    W, H = 640, 400
    view, proj = wl.instances_view(extent=10.0, width=W, height=H, distance=18.0)
    fb, cb = ctx.alloc(8 * W * H), ctx.alloc(8 * W * H)
    scene.render(crx.make_view(view, proj, W, H), fb, cb, crx.render_settings(frustum_culling=0))
    spec = SceneSpec().add(CompressedMesh(pos, idx, col), worlds)
    want_fb, want_cb, _ = oracle_render(spec, W, H, view, proj, attribute=crx.DISPLAY_VERTEX_COLORS)
    got_fb = fb.download(np.uint64, W * H)
    assert np.array_equal(got_fb, want_fb), describe_mismatch(got_fb, want_fb, W)
    assert np.array_equal(cb.download(np.uint64, W * H), want_cb)


def test_mixed_compression_draw_is_rejected(ctx):
    pos, idx, col = wl.sphere_mesh(rings=8, segments=16)
    spec = SceneSpec().add(CompressedMesh(pos, idx, col), [crx.identity()]).add(
        orc.HostMesh(crx, pos, idx, col), [crx.identity()])
    view, proj = wl.sphere_view(64, 64)
    with pytest.raises(crx.CrxError) as e:
        GpuScene(ctx, spec).render(64, 64, view, proj, instanced=False)
    assert e.value.status == 6
