"""Bit-exact parity of the batched multi-view draw.

A multi-view draw fetches every triangle batch once and draws it into each
view whose frustum its box may touch.  Every view's visibility buffer (and
colour buffer) must equal the oracle's single-view draw of the same instance
list into that view, for mixed resolutions, close-ups that route through
stages 2 and 3, more views than lanes in a wave, mirrored instances, cameras
that see nothing and cameras inside the terrain.
"""
import math

import numpy as np
import pytest

import crx
import oracle as orc
import workloads as wl
from crxtest import GpuScene, SceneSpec, describe_mismatch, oracle_render, prepare

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = crx.Context(0)
    yield c
    c.close()


def _check_views(ctx, spec, views, instanced=True, attribute=crx.DISPLAY_VERTEX_COLORS, oracle_every=1):
    """views = [(W, H, view, proj)]; compares every view with the GPU
    single-view draw and every `oracle_every`-th view with the oracle."""
    print(f"[case] {len(views)} views instanced={instanced} attribute={attribute}", flush=True)
    g = GpuScene(ctx, spec)
    fbs, cbs, st = g.render_views(views, instanced=instanced, attribute=attribute)
    total_nt = 0
    for k, (W, H, view, proj) in enumerate(views):
        one_fb, one_cb, one_st = g.render(W, H, view, proj, instanced=instanced, attribute=attribute)
        total_nt += one_st.num_nontrivial
        assert np.array_equal(fbs[k], one_fb), f"view {k} vs single-view GPU draw\n" + describe_mismatch(fbs[k], one_fb, W)
        if attribute is not None:
            assert np.array_equal(cbs[k], one_cb), f"view {k} colours\n" + describe_mismatch(cbs[k], one_cb, W)
        if k % oracle_every == 0:
            want_fb, want_cb, _ = oracle_render(spec, W, H, view, proj, instanced=instanced, attribute=attribute)
            assert np.array_equal(fbs[k], want_fb), f"view {k} vs oracle\n" + describe_mismatch(fbs[k], want_fb, W)
            if attribute is not None:
                assert np.array_equal(cbs[k], want_cb), f"view {k} colours vs oracle\n" + describe_mismatch(cbs[k], want_cb, W)
    # the batched cull never drops a stage-2 entry a view's own draw would make
    assert st.num_nontrivial <= total_nt
    return fbs, st


def _terrain(n=257, amplitude=20.0, scale=64.0):
    pos, idx, col = wl.heightfield_mesh(n, amplitude=amplitude, scale=scale)
    return pos, SceneSpec().add(orc.HostMesh(crx, pos, idx, col), [crx.identity()])


def test_mixed_resolutions_with_closeups(ctx):
    """Scaled-down batched config: 12 seeded views of a 257^2 terrain, mixed
    resolutions, a quarter of them close-ups that need stages 2 and 3."""
    n = 257
    pos, spec = _terrain(n)
    sizes = ((256, 256), (512, 512), (640, 360), (960, 540))
    vs = wl.multiview_views(n, pos, count=12, sizes=sizes, seed=0xC0FFEE04)
    assert sum(c for *_, c in vs) == 3
    views = [(W, H, v, p) for (v, p, W, H, _) in vs]
    fbs, st = _check_views(ctx, spec, views)
    assert st.num_nontrivial > 0 and st.num_huge > 0
    assert sum(int((fb != crx.CLEAR_PIXEL).sum()) for fb in fbs) > 100_000


def test_more_views_than_lanes(ctx):
    """70 views (two lane chunks of the per-view box test), tiny targets."""
    n = 129
    pos, spec = _terrain(n, amplitude=8.0, scale=32.0)
    sizes = ((64, 64), (96, 48), (128, 72), (40, 56))
    vs = wl.multiview_views(n, pos, count=70, sizes=sizes, seed=77)
    views = [(W, H, v, p) for (v, p, W, H, _) in vs]
    _check_views(ctx, spec, views, oracle_every=7)


def test_two_triangles_per_lane_views():
    """Batches of 2 x 256 triangles (each lane takes two): the per-wave boxes
    and view ballots run once per half-batch; 70 views, mixed sizes."""
    t = crx.default_tuning()
    t.stage1_tris_per_lane = 2
    c = crx.Context(0, tuning=t)
    try:
        n = 129
        pos, spec = _terrain(n, amplitude=8.0, scale=32.0)
        sizes = ((64, 64), (96, 48), (128, 72), (40, 56), (320, 180))
        vs = wl.multiview_views(n, pos, count=70, sizes=sizes, seed=78)
        views = [(W, H, v, p) for (v, p, W, H, _) in vs]
        _check_views(c, spec, views, oracle_every=5)
    finally:
        c.close()


def test_instanced_mirrored_views(ctx):
    """Instanced and non-instanced walks with mirrored instances, 5 views
    around the instance cube (one of them looking away from everything)."""
    pos, idx, col = wl.sphere_mesh(rings=32, segments=64, seed=31)
    worlds = list(wl.instance_transforms(40, extent=16.0, smin=0.5, smax=2.0, seed=32))
    for k in range(0, 40, 4):
        for r in range(3):
            worlds[k].m[r] = -worlds[k].m[r]
    spec = SceneSpec().add(orc.HostMesh(crx, pos, idx, col), worlds)
    views = []
    for k, (W, H) in enumerate(((320, 240), (256, 256), (400, 225), (200, 300))):
        a = 2.0 * math.pi * k / 4
        eye = (8.0 + 20.0 * math.cos(a), 8.0 + 3.0 * k, 8.0 + 20.0 * math.sin(a))
        views.append((W, H, crx.look_at(eye, (8.0, 8.0, 8.0)), wl.projection(W, H)))
    views.append((160, 120, crx.look_at((8.0, 8.0, 40.0), (8.0, 8.0, 80.0)), wl.projection(160, 120)))
    for instanced in (True, False):
        _check_views(ctx, spec, views, instanced=instanced)


def test_adversarial_cameras(ctx):
    """Cameras below the surface, grazing the frustum planes, far away, looking
    straight up or down, and a near-degenerate field of view."""
    n = 129
    pos, spec = _terrain(n, amplitude=8.0, scale=32.0)
    g = lambda c, r: float(pos.reshape(-1, 3)[r * n + c, 1])
    cams = [
        ((64.0, g(64, 64) - 0.5, 64.0), (70.0, g(64, 64) - 0.5, 80.0)),    # under the surface
        ((64.0, g(64, 64) + 0.02, 64.0), (64.0, g(64, 64) + 0.02, 90.0)),  # grazing, near plane crossings
        ((-400.0, 300.0, -400.0), (64.0, 0.0, 64.0)),                       # far away
        ((64.0, 50.0, 64.0), (64.0, 0.0, 64.0001)),                         # straight down
        ((64.0, 0.0, 64.0), (64.0, 50.0, 64.0001)),                         # straight up
        ((0.0, 10.0, 0.0), (128.0, 10.0, 0.0)),                             # along the x = edge
        ((-1.0, 5.0, 64.0), (-1.0, 5.0, 128.0)),                            # terrain just right of the frustum
    ]
    views = []
    for k, (eye, at) in enumerate(cams):
        W, H = ((160, 96), (128, 128), (200, 100))[k % 3]
        views.append((W, H, crx.look_at(eye, at), wl.projection(W, H)))
    W, H = 96, 64
    views.append((W, H, crx.look_at((64.0, 20.0, -10.0), (64.0, 0.0, 64.0)), wl.projection(W, H, fovy=math.radians(0.5))))
    _check_views(ctx, spec, views)


def test_render_views_scene_api(ctx):
    """crx_render_views: nodes visible in any view are kept; every view equals
    the oracle's single-view draw of that shared list."""
    pos, idx, col = wl.sphere_mesh(rings=32, segments=64, seed=41)
    pos2, idx2, col2 = wl.sphere_mesh(rings=16, segments=32, radius=2.0, seed=42)
    meshes = [crx.Mesh(ctx, pos, idx, col), crx.Mesh(ctx, pos2, idx2, col2)]
    hosts = [orc.HostMesh(crx, pos, idx, col), orc.HostMesh(crx, pos2, idx2, col2)]
    scene = crx.Scene(ctx)
    worlds = list(wl.instance_transforms(12, extent=20.0, seed=43))
    kinds = [k % 2 for k in range(12)]
    for w, kd in zip(worlds, kinds):
        scene.add(meshes[kd], w)
    # a node of the smaller mesh far above every camera: culled in all views
    far = crx.identity()
    far.m[12], far.m[13], far.m[14] = 10.0, 5000.0, 10.0
    scene.add(meshes[1], far)
    views = []
    for k, (W, H) in enumerate(((320, 200), (256, 256), (300, 300))):
        a = 2.0 * math.pi * k / 3
        eye = (10.0 + 30.0 * math.cos(a), 25.0, 10.0 + 30.0 * math.sin(a))
        views.append(crx.make_view(crx.look_at(eye, (10.0, 10.0, 10.0)), wl.projection(W, H), W, H))
    fbs = [ctx.alloc(8 * v.width * v.height) for v in views]
    cbs = [ctx.alloc(8 * v.width * v.height) for v in views]
    st = scene.render_views(views, fbs, cbs, crx.render_settings())
    # expected list: all 12 near nodes (the far one is culled everywhere),
    # sorted by triangle count (mesh 0 first), node order inside a mesh
    spec = SceneSpec()
    for kd in (0, 1):
        spec.add(hosts[kd], [w for w, k2 in zip(worlds, kinds) if k2 == kd])
    assert st.num_triangles == sum(h.num_triangles * kinds.count(k) for k, h in enumerate(hosts)) * len(views)
    for k, v in enumerate(views):
        W, H = v.width, v.height
        want_fb, want_cb, _ = oracle_render(spec, W, H, v.view, v.proj, attribute=crx.DISPLAY_VERTEX_COLORS)
        got_fb = fbs[k].download(np.uint64, W * H)
        got_cb = cbs[k].download(np.uint64, W * H)
        assert np.array_equal(got_fb, want_fb), describe_mismatch(got_fb, want_fb, W)
        # MESH_ID is not resolved here, so node ids do not enter the colours
        assert np.array_equal(got_cb, want_cb), describe_mismatch(got_cb, want_cb, W)


def test_multiview_rejects_bad_counts(ctx):
    fb = ctx.alloc(8 * 16)
    view, proj = wl.sphere_view(4, 4)
    t = crx.make_target(fb, None, 4, 4, view, proj)
    with pytest.raises(crx.CrxError) as e:
        ctx.draw_triangles_visbuffer_views([], [crx.CMesh()], [crx.identity()] * 257, [t] * 257, False)
    assert e.value.status == 1


def test_autotuner(ctx):
    """crx_autotune: schedule knobs never change a bit; routing knobs change
    the routing and the oracle, given the tuned values, still agrees."""
    n = 257
    pos, spec = _terrain(n)
    sizes = ((256, 256), (512, 512), (640, 360))
    vs = wl.multiview_views(n, pos, count=6, sizes=sizes, seed=99)
    fresh = crx.Context(0)
    try:
        scene_mesh = crx.Mesh(fresh, pos, spec.meshes[0].indices, spec.meshes[0].colors)
        scene = crx.Scene(fresh)
        scene.add(scene_mesh)
        views = [crx.make_view(v, p, W, H) for (v, p, W, H, _) in vs]
        fbs = [fresh.alloc(8 * v.width * v.height) for v in views]
        settings = crx.render_settings(resolve=0)

        def frame():
            scene.render_views(views, fbs, None, settings)
            return [b.download(np.uint64, v.width * v.height) for b, v in zip(fbs, views)]

        base = frame()
        res = fresh.autotune(scene, views, fbs, crx.TUNE_SCHEDULE | crx.TUNE_QUEUES, reps=2)
        print(f"[autotune] schedule: {res.start_ms:.3f} -> {res.best_ms:.3f} ms in {res.trials} trials", flush=True)
        assert res.trials >= 20 and res.best_ms <= res.start_ms
        assert bytes(fresh.tuning) == bytes(res.tuning)
        assert res.tuning.small_threshold == 128 and res.tuning.tile_size == 64
        after = frame()
        for k in range(len(views)):
            assert np.array_equal(after[k], base[k]), f"view {k}: schedule tuning changed the result"

        res2 = fresh.autotune(scene, views, fbs, crx.TUNE_ROUTING, reps=2)
        t = res2.tuning
        print(f"[autotune] routing: small_threshold={t.small_threshold} tile={t.tile_size}", flush=True)
        got = frame()
        for k, (v, p, W, H, _) in enumerate(vs):
            want, _, _ = oracle_render(spec, W, H, v, p, tuning=t)
            assert np.array_equal(got[k], want), f"view {k}\n" + describe_mismatch(got[k], want, W)
    finally:
        fresh.close()


def test_captured_frame_replays(ctx):
    """crx_frame: the frame recorded into a hipGraph, replayed with new
    cameras and with eager draws and synchronisations between the replays;
    every replay equals the eager crx_render_views of the same views."""
    pos, idx, col = wl.sphere_mesh(rings=32, segments=64, seed=51)
    pos2, idx2, col2 = wl.sphere_mesh(rings=16, segments=32, radius=2.0, seed=52)
    meshes = [crx.Mesh(ctx, pos, idx, col), crx.Mesh(ctx, pos2, idx2, col2)]
    scene = crx.Scene(ctx)
    for k, w in enumerate(wl.instance_transforms(10, extent=12.0, seed=53)):
        scene.add(meshes[k % 2], w)
    sizes = ((320, 200), (256, 256))

    def views_at(t):
        out = []
        for k, (W, H) in enumerate(sizes):
            a = 0.7 * t + 2.5 * k
            eye = (6.0 + 28.0 * math.cos(a), 14.0 + 2.0 * t, 6.0 + 28.0 * math.sin(a))
            out.append(crx.make_view(crx.look_at(eye, (6.0, 6.0, 6.0)), wl.projection(W, H), W, H))
        return out

    settings = crx.render_settings()
    fbs = [ctx.alloc(8 * W * H) for (W, H) in sizes]
    cbs = [ctx.alloc(8 * W * H) for (W, H) in sizes]
    efbs = [ctx.alloc(8 * W * H) for (W, H) in sizes]
    ecbs = [ctx.alloc(8 * W * H) for (W, H) in sizes]
    frame = crx.Frame(scene, views_at(0), fbs, cbs, settings)
    try:
        for t in range(4):
            views = views_at(t)
            if t:
                frame.set_views(views)
            frame.launch()
            st = frame.wait()
            est = scene.render_views(views, efbs, ecbs, settings)      # eager work between replays
            ctx.synchronize()
            assert st.num_nontrivial == est.num_nontrivial and st.num_huge == est.num_huge
            for k, (W, H) in enumerate(sizes):
                got = fbs[k].download(np.uint64, W * H)
                want = efbs[k].download(np.uint64, W * H)
                assert np.array_equal(got, want), f"replay {t} view {k}\n" + describe_mismatch(got, want, W)
                assert np.array_equal(cbs[k].download(np.uint64, W * H), ecbs[k].download(np.uint64, W * H))
                assert int((got != crx.CLEAR_PIXEL).sum()) > 1000
    finally:
        frame.close()


def test_captured_single_view_frame_replays(ctx):
    """A one-view crx_frame runs the single-view stage 1 (here with a mesh
    drawn over 40 instances: the per-wave instance lists), which reads its
    view from device memory: replays with new cameras equal eager
    crx_render of the same view."""
    pos, idx, col = wl.sphere_mesh(rings=48, segments=96, seed=54)
    pos2, idx2, col2 = wl.sphere_mesh(rings=12, segments=24, radius=1.5, seed=55)
    meshes = [crx.Mesh(ctx, pos, idx, col), crx.Mesh(ctx, pos2, idx2, col2)]
    scene = crx.Scene(ctx)
    for w in wl.instance_transforms(40, extent=20.0, smin=0.3, smax=1.2, seed=56):
        scene.add(meshes[0], w)
    for w in wl.instance_transforms(2, extent=20.0, seed=57):
        scene.add(meshes[1], w)
    W, H = 320, 200

    def view_at(t):
        a = 0.6 * t
        eye = (10.0 + 40.0 * math.cos(a), 12.0 + 3.0 * t, 10.0 + 40.0 * math.sin(a))
        return crx.make_view(crx.look_at(eye, (10.0, 10.0, 10.0)), wl.projection(W, H), W, H)

    settings = crx.render_settings()
    fb, cb, efb, ecb = (ctx.alloc(8 * W * H) for _ in range(4))
    frame = crx.Frame(scene, [view_at(0)], [fb], [cb], settings)
    try:
        for t in range(4):
            if t:
                frame.set_views([view_at(t)])
            frame.launch()
            st = frame.wait()
            est = scene.render(view_at(t), efb, ecb, settings)
            ctx.synchronize()
            assert st.num_nontrivial == est.num_nontrivial and st.num_huge == est.num_huge
            got, want = fb.download(np.uint64, W * H), efb.download(np.uint64, W * H)
            assert np.array_equal(got, want), f"replay {t}\n" + describe_mismatch(got, want, W)
            assert np.array_equal(cb.download(np.uint64, W * H), ecb.download(np.uint64, W * H))
            assert int((got != crx.CLEAR_PIXEL).sum()) > 500
    finally:
        frame.close()
