"""Bit-exact parity of the gfx950 kernels with the CPU oracle.

Every test renders the same host arrays through the C-ABI on the GPU and
through the oracle on the CPU and requires identical visibility buffers (and
colour buffers where resolved).  The configurations cover all three stages,
both stage-1 walk orders, mirrored instances, near-plane crossings, widths
that are not multiples of 64, non-default routing thresholds and queue
overflow.
"""
import math

import numpy as np
import pytest

import crx
import oracle as orc
import workloads as wl
from crxtest import GpuScene, SceneSpec, describe_mismatch, oracle_render

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = crx.Context(0)
    yield c
    c.close()


def _check(ctx, spec, W, H, view, proj, instanced=True, attribute=crx.DISPLAY_VERTEX_COLORS,
           tuning=None, expect_stage2=False, expect_stage3=False):
    print(f"[case] W={W} H={H} instanced={instanced} attribute={attribute}", flush=True)
    if tuning is not None:
        ctx.tuning = tuning
    try:
        want_fb, want_cb, st = oracle_render(spec, W, H, view, proj, instanced=instanced,
                                             tuning=tuning, attribute=attribute)
        g = GpuScene(ctx, spec)
        fb, cb, gst = g.render(W, H, view, proj, instanced=instanced, attribute=attribute)
    finally:
        ctx.tuning = crx.default_tuning()
    assert gst.num_nontrivial == st.num_nontrivial
    assert gst.num_huge == st.num_huge
    if expect_stage2:
        assert st.num_nontrivial > 0
    if expect_stage3:
        assert st.num_huge > 0
    assert np.array_equal(fb, want_fb), describe_mismatch(fb, want_fb, W)
    if attribute is not None:
        assert np.array_equal(cb, want_cb), describe_mismatch(cb, want_cb, W)
    return fb, st


def test_sphere_1m_1080p(ctx):
    """The 1,048,576-triangle displaced sphere, one 1920x1080 view, resolve to vertex colours."""
    pos, idx, col = wl.sphere_mesh()
    spec = SceneSpec().add(orc.HostMesh(crx, pos, idx, col), [crx.identity()])
    view, proj = wl.sphere_view(1920, 1080)
    fb, st = _check(ctx, spec, 1920, 1080, view, proj)
    assert int((fb != crx.CLEAR_PIXEL).sum()) > 300_000


def test_sphere_non_instanced_odd_size(ctx):
    pos, idx, col = wl.sphere_mesh(rings=128, segments=256)
    spec = SceneSpec().add(orc.HostMesh(crx, pos, idx, col), [crx.identity()])
    view, proj = wl.sphere_view(1000, 701)
    _check(ctx, spec, 1000, 701, view, proj, instanced=False)


def test_instanced_with_mirrors(ctx):
    """Instanced stage-1 walk (the _instanced kernels), including negative-scale instances."""
    pos, idx, col = wl.sphere_mesh(rings=48, segments=96, seed=7)
    worlds = list(wl.instance_transforms(60, extent=20.0, smin=0.5, smax=2.0, seed=11))
    for k in range(0, 60, 5):                       # mirror every fifth instance
        w = worlds[k]
        for r in range(3):
            w.m[r] = -w.m[r]
    pos2, idx2, col2 = wl.sphere_mesh(rings=16, segments=32, radius=3.0, seed=9)
    spec = (SceneSpec()
            .add(orc.HostMesh(crx, pos, idx, col), worlds)
            .add(orc.HostMesh(crx, pos2, idx2, col2), list(wl.instance_transforms(7, 20.0, seed=5))))
    view, proj = wl.instances_view(extent=20.0, width=800, height=600, distance=30.0)
    for instanced in (True, False):
        _check(ctx, spec, 800, 600, view, proj, instanced=instanced)


def _closeup_terrain():
    pos, idx, col = wl.heightfield_mesh(129, spacing=1.0, amplitude=6.0, scale=32.0)
    spec = SceneSpec().add(orc.HostMesh(crx, pos, idx, col), [crx.identity()])
    return spec


@pytest.mark.parametrize("attribute", [crx.DISPLAY_VERTEX_COLORS, crx.DISPLAY_STAGE])
def test_closeup_near_plane_stage2_stage3(ctx, attribute):
    """Camera 1.5 units above a terrain: triangles span hundreds of pixels and
    cross the camera plane, so stage 2 (wave per triangle) and stage 3 (tiles)
    do most of the work; STAGE shades each pixel by the stage its triangle
    is routed to."""
    spec = _closeup_terrain()
    W, H = 1280, 720
    view = crx.look_at((64.0, 4.5, 20.0), (64.0, 2.0, 40.0), (0.0, 1.0, 0.0))
    proj = wl.projection(W, H)
    fb, _ = _check(ctx, spec, W, H, view, proj, attribute=attribute, expect_stage2=True, expect_stage3=True)


@pytest.mark.parametrize("small,tile", [(32, 32), (8, 16), (256, 128), (128, 64)])
def test_routing_tuning(ctx, small, tile):
    """Non-default stage thresholds and tile sizes; the oracle routes identically."""
    spec = _closeup_terrain()
    W, H = 960, 540
    view = crx.look_at((30.0, 12.0, 10.0), (64.0, 0.0, 64.0), (0.0, 1.0, 0.0))
    proj = wl.projection(W, H)
    t = crx.default_tuning()
    t.small_threshold = small
    t.tile_size = tile
    _check(ctx, spec, W, H, view, proj, tuning=t)


def test_scene_render_redraw_resolves_again(ctx):
    """crx_render enqueues the resolve while the stages run; when the queues
    overflow and the frame is drawn again, the resolve must run again on the
    complete visibility buffer: a context with four-entry queues renders the
    close-up terrain to the same visibility and colour buffers as a context
    with default queues."""
    pos, idx, col = wl.heightfield_mesh(129, spacing=1.0, amplitude=6.0, scale=32.0)
    W, H = 1280, 720
    view = crx.make_view(crx.look_at((64.0, 4.5, 20.0), (64.0, 2.0, 40.0), (0.0, 1.0, 0.0)),
                         wl.projection(W, H), W, H)
    settings = crx.render_settings(instanced=1, frustum_culling=1, resolve=1, attribute=crx.DISPLAY_VERTEX_COLORS)
    out = []
    for tiny in (False, True):
        c = crx.Context(0)
        try:
            if tiny:
                t = crx.default_tuning()
                t.max_nontrivial = 4
                t.max_huge = 4
                c.tuning = t
            scene = crx.Scene(c)
            scene.add(crx.Mesh(c, pos, idx, col))
            fb, cb = c.alloc(8 * W * H), c.alloc(8 * W * H)
            st = scene.render(view, fb, cb, settings)
            c.synchronize()
            out.append((fb.download(np.uint64, W * H), cb.download(np.uint64, W * H), st))
            scene.close()
        finally:
            c.close()
    (fb0, cb0, st0), (fb1, cb1, st1) = out
    assert st0.redraws == 0 and st1.redraws >= 1
    assert st0.num_nontrivial > 8 and st0.num_huge > 8
    assert np.array_equal(fb1, fb0), describe_mismatch(fb1, fb0, W)
    assert np.array_equal(cb1, cb0), describe_mismatch(cb1, cb0, W)


def test_tuning_values_are_checked(ctx):
    """crx_context_set_tuning takes 0 (per draw), 1, 2 and 4 triangles per
    lane and cull orders 0-3, and rejects other values."""
    for per_lane in (0, 1, 2, 4):
        t = crx.default_tuning()
        t.stage1_tris_per_lane = per_lane
        ctx.tuning = t
    for field, bad in (("stage1_tris_per_lane", 3), ("stage1_tris_per_lane", 8), ("stage1_cull_order", 4),
                       ("stage1_block", 192)):
        t = crx.default_tuning()
        setattr(t, field, bad)
        with pytest.raises(crx.CrxError) as e:
            ctx.tuning = t
        assert e.value.status == 1, field
    ctx.tuning = crx.default_tuning()


def test_queue_overflow_regrows(ctx):
    spec = _closeup_terrain()
    W, H = 1280, 720
    view = crx.look_at((64.0, 4.5, 20.0), (64.0, 2.0, 40.0), (0.0, 1.0, 0.0))
    proj = wl.projection(W, H)
    want, _, st = oracle_render(spec, W, H, view, proj)
    assert st.num_nontrivial > 8 and st.num_huge > 8
    fresh = crx.Context(0)
    try:
        t = crx.default_tuning()
        t.max_nontrivial = 4
        t.max_huge = 4
        fresh.tuning = t
        fb, _, gst = GpuScene(fresh, spec).render(W, H, view, proj)
    finally:
        fresh.close()
    assert gst.redraws >= 1
    assert np.array_equal(fb, want), describe_mismatch(fb, want, W)


def test_heightfield_oblique(ctx):
    """The oblique heightfield scene scaled down: 2049^2-vertex terrain (8.4M triangles)."""
    n = 2049
    pos, idx, col = wl.heightfield_mesh(n)
    spec = SceneSpec().add(orc.HostMesh(crx, pos, idx, col), [crx.identity()])
    view, proj = wl.heightfield_view(n, 1920, 1080, altitude=150.0, setback=100.0)
    _check(ctx, spec, 1920, 1080, view, proj)


@pytest.mark.parametrize("attribute", [crx.DISPLAY_NONE, crx.DISPLAY_TRIANGLE_ID, crx.DISPLAY_MESH_ID,
                                       crx.DISPLAY_NORMAL, crx.DISPLAY_STAGE])
def test_resolve_attributes(ctx, attribute):
    pos, idx, col = wl.sphere_mesh(rings=64, segments=128)
    worlds = list(wl.instance_transforms(5, extent=4.0, seed=3))
    spec = SceneSpec().add(orc.HostMesh(crx, pos, idx, col), worlds)
    view, proj = wl.instances_view(extent=4.0, width=640, height=480, distance=9.0)
    _check(ctx, spec, 640, 480, view, proj, attribute=attribute)


def _two_mesh_scene(ctx, n=30):
    """A scene of n nodes over two sphere meshes (every third node the larger
    one), with the oracle's copies of the meshes."""
    pos, idx, col = wl.sphere_mesh(rings=32, segments=64, seed=21)
    pos2, idx2, col2 = wl.sphere_mesh(rings=24, segments=48, radius=2.0, seed=22)
    meshes = [crx.Mesh(ctx, pos, idx, col), crx.Mesh(ctx, pos2, idx2, col2)]
    hosts = [orc.HostMesh(crx, pos, idx, col), orc.HostMesh(crx, pos2, idx2, col2)]
    scene = crx.Scene(ctx)
    worlds = list(wl.instance_transforms(n, extent=30.0, seed=23))
    order = []
    for k, w in enumerate(worlds):
        mk = k % 3 == 0
        scene.add(meshes[1 if mk else 0], w)
        order.append(1 if mk else 0)
    return scene, meshes, hosts, worlds, order


def _oracle_scene_frame(meshes, hosts, worlds, order, v, view, proj, W, H, settings):
    """The oracle's CuRast::draw of the node list (node cull, sort/group,
    prefix), its three stages and its resolve."""
    import ctypes as C
    L = orc.lib(crx)
    n = len(worlds)
    node_meshes = (crx.CMesh * n)()
    node_world = (crx.Mat4 * n)()
    for k in range(n):
        cm = hosts[order[k]].cmesh(crx)
        cm.aabb = meshes[order[k]].aabb()
        cm.address = order[k] + 1
        node_meshes[k] = cm
        node_world[k] = worlds[k]
    uniq = (crx.CMesh * n)()
    nu = C.c_uint32()
    inst = (crx.CMesh * n)()
    xf = (crx.Mat4 * n)()
    pre = (C.c_uint64 * n)()
    nv = L.crx_oracle_prepare_frame(node_meshes, node_world, n, C.byref(v), int(settings.frustum_culling),
                                    0 if settings.instanced else 1, uniq, C.byref(nu), inst, xf, pre)
    for k in range(nv):
        inst[k].id = -1
    if nv:
        fbw, _ = orc.draw(crx, list(uniq)[:nu.value], list(inst)[:nv], list(xf)[:nv], W, H, view, proj,
                          bool(settings.instanced))
    else:
        fbw = np.empty(W * H, dtype=np.uint64)
        L.crx_oracle_clear(fbw.ctypes.data, fbw.size)
    cbw = None
    if settings.resolve:
        cbw = orc.resolve(crx, list(inst)[:nv], list(pre)[:nv], fbw, W, H, view, proj, settings.attribute)
    return fbw, cbw, nv


def test_scene_api_matches_oracle_frame(ctx):
    """crx_render (scene API: node cull, sort/group, prefix, clear, stages,
    resolve) against the oracle's restatement of CuRast::draw."""
    scene, meshes, hosts, worlds, order = _two_mesh_scene(ctx)
    W, H = 800, 450
    view, proj = wl.instances_view(extent=30.0, width=W, height=H, distance=25.0)
    v = crx.make_view(view, proj, W, H)
    fb = ctx.alloc(8 * W * H)
    cb = ctx.alloc(8 * W * H)
    settings = crx.render_settings()
    scene.render(v, fb, cb, settings)
    got_fb = fb.download(np.uint64, W * H)
    got_cb = cb.download(np.uint64, W * H)
    fbw, cbw, nv = _oracle_scene_frame(meshes, hosts, worlds, order, v, view, proj, W, H, settings)
    assert 0 < nv <= len(worlds)
    assert np.array_equal(got_fb, fbw), describe_mismatch(got_fb, fbw, W)
    assert np.array_equal(got_cb, cbw), describe_mismatch(got_cb, cbw, W)
    scene.close()


@pytest.mark.parametrize("nodes", [30, 400])
def test_scene_frames_keep_and_rebuild_lists(ctx, nodes):
    """A run of crx_render frames over one scene: the per-node lists and the
    uploaded instance records are kept while the node set, the instancing
    switch and the culled set stay the same, and rebuilt when any changes (a
    view that culls other nodes, a moved node, a frame that sees nothing, the
    other instancing mode, a frame without resolve, a C-ABI draw in between).
    Every frame equals the oracle's, colour buffer included.  400 nodes make
    the kept records large enough that a repeated frame copies only the
    rest of the metadata."""
    scene, meshes, hosts, worlds, order = _two_mesh_scene(ctx, nodes)
    W, H = 640, 360
    proj = wl.projection(W, H)
    view_a, _ = wl.instances_view(extent=30.0, width=W, height=H, distance=25.0)
    view_b = crx.look_at((0.0, 0.0, 0.0), (30.0, 0.0, 5.0), (0.0, 1.0, 0.0))
    view_e = crx.look_at((500.0, 0.0, 0.0), (1000.0, 0.0, 0.0), (0.0, 1.0, 0.0))
    fb = ctx.alloc(8 * W * H)
    cb = ctx.alloc(8 * W * H)
    plain = crx.render_settings()
    seen = set()

    def frame(name, view, settings=plain):
        print(f"[case] {nodes} nodes, frame {name}", flush=True)
        v = crx.make_view(view, proj, W, H)
        scene.render(v, fb, cb if settings.resolve else None, settings)
        got_fb = fb.download(np.uint64, W * H)
        fbw, cbw, nv = _oracle_scene_frame(meshes, hosts, worlds, order, v, view, proj, W, H, settings)
        assert np.array_equal(got_fb, fbw), f"{name}: " + describe_mismatch(got_fb, fbw, W)
        if settings.resolve:
            got_cb = cb.download(np.uint64, W * H)
            assert np.array_equal(got_cb, cbw), f"{name}: " + describe_mismatch(got_cb, cbw, W)
        seen.add(nv)

    frame("A", view_a)
    frame("A again", view_a)
    frame("B", view_b)
    frame("A after B", view_a)
    w3 = worlds[3]
    m = crx.Mat4()
    for k in range(16):
        m.m[k] = w3.m[k]
    m.m[12] += 4.0
    m.m[13] -= 3.0
    scene.set_transform(3, m)
    worlds[3] = m
    frame("A after a moved node", view_a)
    # a C-ABI draw of other lists on the same context between two frames
    pos, idx, col = wl.sphere_mesh(rings=16, segments=32, seed=5)
    other = SceneSpec().add(orc.HostMesh(crx, pos, idx, col), [crx.identity()])
    sv, sp = wl.sphere_view(W, H)
    GpuScene(ctx, other).render(W, H, sv, sp)
    frame("A after a C-ABI draw", view_a)
    frame("E (nothing visible)", view_e)
    frame("A after E", view_a)
    frame("A non-instanced", view_a, crx.render_settings(instanced=0))
    frame("A without resolve", view_a, crx.render_settings(resolve=0))
    frame("A with resolve", view_a)
    assert 0 in seen and len(seen) >= 3
    scene.close()


def test_scene_frames_grow_metadata_with_kept_lists():
    """A one-view frame and then multi-view frames whose views keep the same
    400 nodes: the kept instance records stand, but the metadata grows with
    the views (per-view records and transforms) and its device buffer is
    reallocated, so the records must be written again into the new buffer.
    A fresh context makes the first frame size the buffer.  Every view equals
    the oracle's frame."""
    c = crx.Context(0)
    try:
        scene, meshes, hosts, worlds, order = _two_mesh_scene(c, 400)
        W, H = 640, 360
        proj = wl.projection(W, H)
        view_a, _ = wl.instances_view(extent=30.0, width=W, height=H, distance=25.0)
        settings = crx.render_settings()
        v = crx.make_view(view_a, proj, W, H)
        fbw, cbw, nv = _oracle_scene_frame(meshes, hosts, worlds, order, v, view_a, proj, W, H, settings)
        assert nv > 100
        fbs = [c.alloc(8 * W * H) for _ in range(8)]
        cbs = [c.alloc(8 * W * H) for _ in range(8)]
        for count in (1, 2, 8, 1, 8):
            print(f"[case] {count} view(s) over the kept node set", flush=True)
            if count == 1:
                scene.render(v, fbs[0], cbs[0], settings)
            else:
                scene.render_views([v] * count, fbs[:count], cbs[:count], settings)
            for k in range(count):
                got_fb = fbs[k].download(np.uint64, W * H)
                got_cb = cbs[k].download(np.uint64, W * H)
                assert np.array_equal(got_fb, fbw), f"{count} views, view {k}: " + describe_mismatch(got_fb, fbw, W)
                assert np.array_equal(got_cb, cbw), f"{count} views, view {k} colours"
        scene.close()
    finally:
        c.close()


def test_scene_frames_on_two_streams(ctx):
    """Two frames on two streams before either is read back: crx_render
    returns while its resolve may still run, and the next frame's draw on the
    other stream rewrites the context's metadata (another culled set) only
    after that resolve.  Both frames equal the oracle's."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so.7")   # the runtime libcrx already loaded (same soname)
    streams = []
    for _ in range(2):
        h = C.c_void_p()
        assert hip.hipStreamCreate(C.byref(h)) == 0
        streams.append(h)
    scene, meshes, hosts, worlds, order = _two_mesh_scene(ctx)
    W, H = 640, 360
    proj = wl.projection(W, H)
    view_a, _ = wl.instances_view(extent=30.0, width=W, height=H, distance=25.0)
    view_b = crx.look_at((0.0, 0.0, 0.0), (30.0, 0.0, 5.0), (0.0, 1.0, 0.0))
    bufs = [(ctx.alloc(8 * W * H), ctx.alloc(8 * W * H)) for _ in range(2)]
    settings = crx.render_settings()
    for rnd in range(3):
        print(f"[case] two streams, round {rnd}", flush=True)
        views = [crx.make_view(view_a, proj, W, H), crx.make_view(view_b, proj, W, H)]
        for k in range(2):
            scene.render(views[k], bufs[k][0], bufs[k][1], settings, streams[k])
        for k, view in enumerate((view_a, view_b)):
            got_fb = bufs[k][0].download(np.uint64, W * H, streams[k])
            got_cb = bufs[k][1].download(np.uint64, W * H, streams[k])
            fbw, cbw, nv = _oracle_scene_frame(meshes, hosts, worlds, order, views[k], view, proj, W, H, settings)
            assert nv > 0
            assert np.array_equal(got_fb, fbw), f"stream {k}: " + describe_mismatch(got_fb, fbw, W)
            assert np.array_equal(got_cb, cbw), f"stream {k}: " + describe_mismatch(got_cb, cbw, W)
    ctx.synchronize()
    for h in streams:
        assert hip.hipStreamSynchronize(h) == 0
        assert hip.hipStreamDestroy(h) == 0
    scene.close()


@pytest.mark.parametrize("cull_order,fast_tiny,order,block,per_lane", [
    (0, 0, 0, 256, 1), (0, 1, 1, 512, 2), (1, 0, 2, 128, 2), (1, 1, 1, 512, 1), (2, 1, 0, 256, 2),
    (2, 0, 1, 512, 1), (1, 1, 1, 256, 2), (1, 1, 0, 128, 1), (1, 1, 2, 256, 4), (0, 1, 1, 512, 4),
    (2, 1, 0, 128, 4), (3, 1, 1, 256, 2), (3, 0, 2, 128, 1)])
def test_schedule_variants_bit_identical(ctx, cull_order, fast_tiny, order, block, per_lane):
    """Cull order, the filtered tiny cull, batch order, workgroup size and
    triangles per lane only change the schedule: every variant must
    reproduce the oracle exactly."""
    pos, idx, col = wl.sphere_mesh(rings=256, segments=512)
    worlds = [crx.identity()] + list(wl.instance_transforms(6, extent=3.0, smin=0.2, smax=0.6, seed=17))
    spec = SceneSpec().add(orc.HostMesh(crx, pos, idx, col), worlds)
    view, proj = wl.sphere_view(1280, 720, radius=1.3)
    t = crx.default_tuning()
    t.stage1_cull_order, t.stage1_fast_tiny, t.stage1_order, t.stage1_block = cull_order, fast_tiny, order, block
    t.stage1_tris_per_lane = per_lane
    _check(ctx, spec, 1280, 720, view, proj, tuning=t)


@pytest.mark.parametrize("scale", [1e19, 3e20])
def test_extreme_magnitudes_take_ieee_division(ctx, scale):
    """Stage 1 takes its short division sequences only for waves whose depths
    lie in [2^-60, 2^64] and whose numerators stay within 2^64.  Instances
    scaled and placed at ~1e19-1e20 units (numerators past 2^64 in some or
    all lanes), one hugging the camera plane at depth ~1e-19 (below 2^-60)
    and ordinary ones in the same draw must still match the oracle bit for
    bit."""
    pos, idx, col = wl.sphere_mesh(rings=24, segments=48, seed=81)
    view = crx.look_at((0.0, 0.0, 10.0), (0.0, 0.0, 0.0))
    vi = np.array(crx.mat4_inverse(view).m, dtype=np.float32).reshape(4, 4).T

    def at(x, y, z, s):          # an instance at view-space (x, y, z), scale s
        w = np.eye(4, dtype=np.float32) * np.float32(s)
        w[3, 3] = 1.0
        c = vi @ np.array([x, y, z, 1.0], dtype=np.float32)
        w[:3, 3] = c[:3]
        return crx.Mat4.from_numpy(w)

    worlds = [at(0.0, 0.0, -5.0, 1.0), at(2.0, 1.0, -8.0, 0.5),                 # ordinary
              at(0.2 * scale, 0.1 * scale, -scale, 0.3 * scale),                  # huge and far
              at(0.0, 0.0, -1.2 * scale, 0.05 * scale),
              at(0.0, 0.0, -2e-19, 1e-19),                                         # at the camera plane
              at(3e-19, 0.0, -5e-19, 2e-19)]
    spec = SceneSpec().add(orc.HostMesh(crx, pos, idx, col), worlds)
    _check(ctx, spec, 640, 400, view, wl.projection(640, 400))


@pytest.mark.parametrize("order,per_lane,instanced", [
    (0, 1, False), (1, 2, False), (2, 4, False), (2, 1, True), (0, 2, True), (1, 4, True)])
def test_many_items_every_batch_order(ctx, order, per_lane, instanced):
    """Thirty meshes of ragged sizes (from 7 triangles to several batches)
    in one draw: every workgroup's batch sequence crosses item boundaries,
    in the grid-stride order and in both contiguous orders, where the sweep
    starts a new pass per item."""
    rng = np.random.default_rng(71)
    spec = SceneSpec()
    for k in range(30):
        rings, segs = int(rng.integers(2, 40)), int(rng.integers(3, 60))
        pos, idx, col = wl.sphere_mesh(rings=rings, segments=segs, seed=100 + k)
        if k % 7 == 3:
            idx = idx[: 7 + k]                     # a few triangles only
        cnt = 1 + k % 3
        worlds = list(wl.instance_transforms(cnt, extent=6.0, smin=0.3, smax=0.9, seed=200 + k))
        spec.add(orc.HostMesh(crx, pos, idx, col), worlds)
    view = crx.look_at((3.0, 2.5, 12.0), (3.0, 3.0, 3.0))
    t = crx.default_tuning()
    t.stage1_order, t.stage1_tris_per_lane, t.stage1_block = order, per_lane, 128
    _check(ctx, spec, 800, 600, view, wl.projection(800, 600), instanced=instanced, tuning=t)


@pytest.mark.parametrize("per_lane", [0, 1, 2, 4])
def test_many_instances_sub_pixel(ctx, per_lane):
    """A dense mesh drawn as many small instances (the 3000-instance scene
    scaled down): most 16-triangle row groups project inside one pixel and
    are dropped from that row's instance list by the box tiny test, so the
    rows of a wave draw different instances in the same step.  Extra
    instances cross the near plane, sit behind the camera and hug the
    viewport's first and last columns and rows; every bit must still be the
    oracle's.  per_lane 0 is the per-draw choice."""
    pos, idx, col = wl.sphere_mesh(rings=128, segments=256, seed=61)
    worlds = list(wl.instance_transforms(300, extent=60.0, smin=0.2, smax=1.5, seed=62))
    view, proj = wl.instances_view(extent=60.0, width=960, height=544, distance=110.0)
    vi = crx.mat4_inverse(view)
    def at(x, y, z, s):          # an instance at view-space (x, y, z), scale s
        w = np.eye(4, dtype=np.float32) * s
        w[3, 3] = 1.0
        c = np.array(vi.m, dtype=np.float32).reshape(4, 4).T @ np.array([x, y, z, 1.0], dtype=np.float32)
        w[:3, 3] = c[:3]
        return crx.Mat4.from_numpy(w)
    f = 1.0 / math.tan(wl.FOVY / 2)
    a = 960.0 / 544.0
    for z in (-30.0, -60.0, -95.0):
        e = -z / f
        for (x, y) in ((-e * a * 0.999, 0.0), (e * a * 0.999, 0.0), (0.0, -e * 0.999), (0.0, e * 0.999),
                       (-e * a, -e), (e * a * 1.01, e * 1.01)):
            worlds.append(at(x, y, z, 0.3))
    worlds.append(at(0.0, 0.0, 0.5, 1.0))      # around the camera: crosses the near plane
    worlds.append(at(0.0, 0.0, 20.0, 2.0))     # behind the camera
    spec = SceneSpec().add(orc.HostMesh(crx, pos, idx, col), worlds)
    t = crx.default_tuning()
    t.stage1_tris_per_lane = per_lane
    _check(ctx, spec, 960, 544, view, proj, instanced=True, tuning=t)


@pytest.mark.parametrize("per_lane", [1, 2, 4])
@pytest.mark.parametrize("instanced", [True, False])
def test_non_indexed_meshes(ctx, per_lane, instanced):
    """Non-indexed meshes (triangle t = vertices 3t..3t+2) drawn in one draw
    with an indexed one, ragged triangle counts so batches end mid-wave."""
    pos, idx, col = wl.sphere_mesh(rings=61, segments=83, seed=41)
    flat_pos = pos[idx.reshape(-1)]
    flat_col = col[idx.reshape(-1)]
    pos2, idx2, col2 = wl.sphere_mesh(rings=17, segments=29, seed=42)
    worlds = list(wl.instance_transforms(5, extent=4.0, smin=0.4, smax=1.0, seed=43))
    spec = (SceneSpec()
            .add(orc.HostMesh(crx, flat_pos, None, flat_col), worlds[:3])
            .add(orc.HostMesh(crx, pos2, idx2, col2), worlds[3:])
            .add(orc.HostMesh(crx, pos2[idx2[:997].reshape(-1)], None, col2[idx2[:997].reshape(-1)]), [crx.identity()]))
    view = crx.look_at((2.0, 1.5, 9.0), (2.0, 2.0, 2.0))
    t = crx.default_tuning()
    t.stage1_tris_per_lane = per_lane
    _check(ctx, spec, 800, 600, view, wl.projection(800, 600), instanced=instanced, tuning=t)


@pytest.mark.parametrize("per_lane", [1, 2])
def test_random_index_order(ctx, per_lane):
    """Triangles in random order (no locality between neighbouring lanes)
    plus degenerate triangles that repeat an index."""
    pos, idx, col = wl.sphere_mesh(rings=96, segments=128, seed=44)
    rng = np.random.default_rng(45)
    perm = rng.permutation(len(idx))
    idx_r = idx[perm].copy()
    idx_r[::17, 1] = idx_r[::17, 0]                     # degenerate: a repeated index
    worlds = [crx.identity()]
    spec = SceneSpec().add(orc.HostMesh(crx, pos, idx_r, col), worlds)
    view, proj = wl.sphere_view(960, 540, radius=1.2)
    t = crx.default_tuning()
    t.stage1_tris_per_lane = per_lane
    _check(ctx, spec, 960, 540, view, proj, tuning=t)


@pytest.mark.parametrize("seed", [101, 202, 303, 404, 505, 606])
def test_random_scenes(ctx, seed):
    """Seeded random scenes: two to five meshes (spheres and terrain patches
    of random sizes, some non-indexed), one to forty instances each with
    random scale (some mirrored), a random camera that may sit among the
    instances, and a random schedule; every bit must be the oracle's."""
    rng = np.random.default_rng(seed)
    spec = SceneSpec()
    for m in range(int(rng.integers(2, 6))):
        if rng.random() < 0.6:
            pos, idx, col = wl.sphere_mesh(rings=int(rng.integers(4, 64)), segments=int(rng.integers(6, 96)),
                                           radius=float(rng.uniform(0.2, 2.0)), seed=seed * 10 + m)
        else:
            n = int(rng.integers(5, 80))
            pos, idx, col = wl.heightfield_mesh(n, spacing=float(rng.uniform(0.05, 0.5)),
                                                amplitude=float(rng.uniform(0.1, 2.0)), scale=16.0,
                                                seed=seed * 10 + m)
        if rng.random() < 0.25:                                   # non-indexed copy
            flat = idx.reshape(-1)
            pos, idx, col = pos[flat], None, col[flat]
        cnt = int(rng.integers(1, 41))
        worlds = list(wl.instance_transforms(cnt, extent=12.0, smin=0.2, smax=1.5, seed=seed * 100 + m))
        for k in range(0, cnt, 7):                                # mirror some
            for r in range(3):
                worlds[k].m[r] = -worlds[k].m[r]
        spec.add(orc.HostMesh(crx, pos, idx, col), worlds)
    eye = tuple(float(v) for v in rng.uniform(-4.0, 16.0, 3))
    target = tuple(float(v) for v in rng.uniform(2.0, 10.0, 3))
    W, H = int(rng.choice([320, 640, 777])), int(rng.choice([200, 400, 333]))
    view = crx.look_at(eye, target)
    t = crx.default_tuning()
    t.stage1_tris_per_lane = int(rng.choice([0, 1, 2, 4]))
    t.stage1_order = int(rng.integers(0, 3))
    t.stage1_cull_order = int(rng.integers(0, 4))
    t.stage1_fast_tiny = int(rng.integers(0, 2))
    t.stage1_block = int(rng.choice([128, 256, 512]))
    _check(ctx, spec, W, H, view, wl.projection(W, H), instanced=bool(rng.random() < 0.7), tuning=t)
