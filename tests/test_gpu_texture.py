"""TEXTURE display attribute on the GPU against the oracle, bit for bit.

The resolve samples CMesh::texture bilinearly at the mip level of the
pixel's uv footprint (resolve.cu:148-195, 783-882).  Textures here are mip
chains of seeded random texels, so any difference in the footprint, the level,
the texel addresses, the weights or their rounding changes a colour.  The
cases cover several meshes with their own textures in one draw (JPEG-coded,
missing data and missing uvs among them), an oblique terrain whose footprint
spans every level, instanced and IX_PU16 meshes, the multi-view resolve and
the scene API; the UV attribute rides along.
"""
import math

import numpy as np
import pytest

import crx
import oracle as orc
import workloads as wl
from crxtest import (CompressedMesh, GpuScene, SceneSpec, describe_mismatch, mip_chain, mip_shapes,
                     oracle_render)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = crx.Context(0)
    yield c
    c.close()


def _random_chain(w, h, seed):
    rng = np.random.default_rng(seed)
    return mip_chain([rng.integers(0, 2 ** 32, size=s, dtype=np.uint32) for s in mip_shapes(w, h)])


def _sphere_uvs(pos, repeat=(3.0, 2.0)):
    """Longitude / latitude coordinates, repeated so that uv leaves [0, 1)."""
    p = pos.astype(np.float64)
    r = np.linalg.norm(p, axis=1)
    u = (np.arctan2(p[:, 2], p[:, 0]) / (2.0 * np.pi) + 0.5) * repeat[0]
    v = np.arccos(np.clip(p[:, 1] / r, -1.0, 1.0)) / np.pi * repeat[1]
    return np.stack([u, v], axis=1).astype(np.float32)


def _terrain_uvs(pos, n, repeat=8.0):
    return np.stack([pos[:, 0] / (n - 1) * repeat, pos[:, 2] / (n - 1) * repeat - 0.25], axis=1).astype(np.float32)


def _check(ctx, spec, W, H, view, proj, attribute=crx.DISPLAY_TEXTURE, instanced=True, min_textured=1000):
    print(f"[case] W={W} H={H} instanced={instanced} attribute={attribute}", flush=True)
    want_fb, want_cb, _ = oracle_render(spec, W, H, view, proj, instanced=instanced, attribute=attribute)
    g = GpuScene(ctx, spec)
    fb, cb, _ = g.render(W, H, view, proj, instanced=instanced, attribute=attribute)
    assert np.array_equal(fb, want_fb), describe_mismatch(fb, want_fb, W)
    assert np.array_equal(cb, want_cb), describe_mismatch(cb, want_cb, W)
    # the sampled colours are not all one value: the texture was read
    cov = fb != crx.CLEAR_PIXEL
    assert int(np.count_nonzero(cov)) >= min_textured
    assert np.unique(cb[cov] & 0xFFFFFFFF).size > 50
    return fb, cb


def test_texture_meshes_with_their_own_textures(ctx):
    """Five spheres, each with its own texture and state: 256 x 192 and
    100 x 60 chains, a JPEG-coded one (magenta), uvs without texture data
    (black) and a texture without uvs (the final branch, blue)."""
    specs = [dict(tex=(256, 192), seed=1), dict(tex=(100, 60), seed=2), dict(tex=(64, 64), seed=3, jpeg=True),
             dict(tex=None, seed=4), dict(tex=(32, 48), seed=5, uvs=False)]
    spec = SceneSpec()
    worlds = []
    for k in range(len(specs)):    # a row of spheres across the view
        w = crx.identity()
        w.m[12], w.m[13] = 2.3 * (k - 2), 0.4 * (k % 2)
        worlds.append(w)
    for k, s in enumerate(specs):
        pos, idx, col = wl.sphere_mesh(rings=48, segments=96, seed=60 + k)
        chain = _random_chain(*s["tex"], seed=s["seed"]) if s["tex"] else None
        mesh = orc.HostMesh(crx, pos, idx, col, uvs=_sphere_uvs(pos) if s.get("uvs", True) else None,
                            texture=chain, texture_size=s["tex"] or (64, 64), jpeg=s.get("jpeg", False))
        spec.add(mesh, [worlds[k]])
    view, proj = crx.look_at((0.5, 1.0, 7.0), (0.0, 0.0, 0.0)), wl.projection(640, 480)
    fb, cb = _check(ctx, spec, 640, 480, view, proj)
    colors = set(int(c) for c in np.unique(cb[fb != crx.CLEAR_PIXEL] & 0xFFFFFFFF))
    assert {0xFFFF00FF, 0xFF000000, 0xFF0000FF} <= colors


def test_texture_oblique_terrain_spans_the_levels(ctx):
    """A 1025^2 terrain seen at a grazing angle at 1920 x 1080: the footprint
    grows from under a texel near the camera to many texels at the horizon,
    so every level of a 1024 x 1024 chain is sampled."""
    n = 1025
    pos, idx, col = wl.heightfield_mesh(n, amplitude=30.0, scale=256.0)
    chain = _random_chain(1024, 1024, seed=11)
    spec = SceneSpec().add(orc.HostMesh(crx, pos, idx, col, uvs=_terrain_uvs(pos, n), texture=chain,
                                        texture_size=(1024, 1024)), [crx.identity()])
    view, proj = wl.heightfield_view(n, 1920, 1080, altitude=40.0, pitch_deg=12.0, setback=10.0)
    _check(ctx, spec, 1920, 1080, view, proj, min_textured=500000)


def test_texture_instanced_and_mirrored(ctx):
    """One textured mesh drawn as 40 instances (mirrored ones among them) with
    the instanced and the per-instance kernels."""
    pos, idx, col = wl.sphere_mesh(rings=24, segments=48, seed=70)
    chain = _random_chain(128, 128, seed=12)
    worlds = list(wl.instance_transforms(40, extent=8.0, seed=71))
    for k in range(0, 40, 5):      # mirrored: the first column negated
        for e in range(3):
            worlds[k].m[e] = -worlds[k].m[e]
    spec = SceneSpec().add(orc.HostMesh(crx, pos, idx, col, uvs=_sphere_uvs(pos, (5.0, 3.0)), texture=chain,
                                        texture_size=(128, 128)), worlds)
    view, proj = wl.instances_view(extent=8.0, width=800, height=600, distance=16.0)
    for instanced in (True, False):
        _check(ctx, spec, 800, 600, view, proj, instanced=instanced)


def test_texture_compressed_mesh(ctx):
    """IX_PU16 geometry: the resolve decodes the packed indices and the
    uint16 positions before the footprint."""
    n = 257
    pos, idx, col = wl.heightfield_mesh(n, amplitude=20.0, scale=64.0)
    chain = _random_chain(300, 200, seed=13)
    spec = SceneSpec().add(CompressedMesh(pos, idx, col, uvs=_terrain_uvs(pos, n, 3.0), texture=chain,
                                          texture_size=(300, 200)), [crx.identity()])
    view, proj = wl.heightfield_view(n, 960, 540, altitude=40.0, pitch_deg=25.0, setback=20.0)
    _check(ctx, spec, 960, 540, view, proj)


def test_texture_multiview_resolve(ctx):
    """The batched multi-view resolve samples each view's footprint."""
    n = 257
    pos, idx, col = wl.heightfield_mesh(n, amplitude=20.0, scale=64.0)
    chain = _random_chain(512, 256, seed=14)
    spec = SceneSpec().add(orc.HostMesh(crx, pos, idx, col, uvs=_terrain_uvs(pos, n, 4.0), texture=chain,
                                        texture_size=(512, 256)), [crx.identity()])
    views = []
    for k, (W, H) in enumerate(((320, 200), (256, 256), (500, 280))):
        eye = (60.0 + 50.0 * k, 35.0 + 10.0 * k, -15.0)
        views.append((W, H, crx.look_at(eye, (128.0, 0.0, 128.0)), wl.projection(W, H)))
    print(f"[case] {len(views)} views TEXTURE", flush=True)
    g = GpuScene(ctx, spec)
    fbs, cbs, _ = g.render_views(views, attribute=crx.DISPLAY_TEXTURE)
    for k, (W, H, view, proj) in enumerate(views):
        want_fb, want_cb, _ = oracle_render(spec, W, H, view, proj, attribute=crx.DISPLAY_TEXTURE)
        assert np.array_equal(fbs[k], want_fb), f"view {k}\n" + describe_mismatch(fbs[k], want_fb, W)
        assert np.array_equal(cbs[k], want_cb), f"view {k} colours\n" + describe_mismatch(cbs[k], want_cb, W)


def test_uv_attribute(ctx):
    """DISPLAY_UV shades the wrapped uv of the pixel (resolve.cu:923-926)."""
    pos, idx, col = wl.sphere_mesh(rings=64, segments=128)
    worlds = list(wl.instance_transforms(5, extent=4.0, seed=3))
    spec = SceneSpec().add(orc.HostMesh(crx, pos, idx, col, uvs=_sphere_uvs(pos)), worlds)
    view, proj = wl.instances_view(extent=4.0, width=640, height=480, distance=9.0)
    _check(ctx, spec, 640, 480, view, proj, attribute=crx.DISPLAY_UV)


def test_normal_attribute_with_vertex_normals(ctx):
    """DISPLAY_NORMAL with per-vertex normals (resolve.cu:927-943), turned by
    each instance's world (mirrored and scaled ones among them), against the
    oracle; and the same through the scene API with crx_mesh_set_attributes."""
    pos, idx, col = wl.sphere_mesh(rings=48, segments=96, seed=90)
    nrm = (pos / np.linalg.norm(pos, axis=1, keepdims=True)).astype(np.float32)
    worlds = list(wl.instance_transforms(12, extent=6.0, seed=91))
    for k in range(0, 12, 4):
        for e in range(3):
            worlds[k].m[e] = -worlds[k].m[e]
    spec = SceneSpec().add(orc.HostMesh(crx, pos, idx, col, normals=nrm), worlds)
    W, H = 800, 600
    view, proj = wl.instances_view(extent=6.0, width=W, height=H, distance=12.0)
    fb, cb = _check(ctx, spec, W, H, view, proj, attribute=crx.DISPLAY_NORMAL)
    # the scene API's mesh carries the same normals
    import ctypes as C
    L = orc.lib(crx)
    mesh = crx.Mesh(ctx, pos, idx, col)
    mesh.set_attributes(normals=nrm)
    host = spec.meshes[0]
    scene = crx.Scene(ctx)
    for w in worlds:
        scene.add(mesh, w)
    v = crx.make_view(view, proj, W, H)
    settings = crx.render_settings()
    settings.attribute = crx.DISPLAY_NORMAL
    sfb, scb = ctx.alloc(8 * W * H), ctx.alloc(8 * W * H)
    print("[case] scene NORMAL", flush=True)
    scene.render(v, sfb, scb, settings)
    got_fb, got_cb = sfb.download(np.uint64, W * H), scb.download(np.uint64, W * H)
    n = len(worlds)
    node_meshes, node_world = (crx.CMesh * n)(), (crx.Mat4 * n)()
    for k in range(n):
        cm = host.cmesh(crx)
        cm.aabb = mesh.aabb()
        cm.address = 1
        node_meshes[k] = cm
        node_world[k] = worlds[k]
    uniq, nu, inst, xf, pre = (crx.CMesh * n)(), C.c_uint32(), (crx.CMesh * n)(), (crx.Mat4 * n)(), (C.c_uint64 * n)()
    nv = L.crx_oracle_prepare_frame(node_meshes, node_world, n, C.byref(v), int(settings.frustum_culling), 0,
                                    uniq, C.byref(nu), inst, xf, pre)
    assert nv > 0
    fbw, _ = orc.draw(crx, list(uniq)[:nu.value], list(inst)[:nv], list(xf)[:nv], W, H, view, proj, True)
    cbw = orc.resolve(crx, list(inst)[:nv], list(pre)[:nv], fbw, W, H, view, proj, crx.DISPLAY_NORMAL)
    assert np.array_equal(got_fb, fbw), describe_mismatch(got_fb, fbw, W)
    assert np.array_equal(got_cb, cbw), describe_mismatch(got_cb, cbw, W)
    scene.close()
    mesh.close()


def test_texture_through_the_scene_api(ctx):
    """crx_mesh_set_attributes and crx_scene_set_texture: per-node textures
    over shared meshes, against the oracle's frame of the same nodes; clearing
    a node's texture takes effect on the next frame."""
    import ctypes as C
    L = orc.lib(crx)
    pos, idx, col = wl.sphere_mesh(rings=32, segments=64, seed=21)
    uvs = _sphere_uvs(pos, (4.0, 2.0))
    mesh = crx.Mesh(ctx, pos, idx, col)
    mesh.set_attributes(uvs=uvs)
    host = orc.HostMesh(crx, pos, idx, col, uvs=uvs)
    chains = [_random_chain(96, 64, seed=30 + k) for k in range(3)]
    dev_chains = [ctx.to_device(c) for c in chains]
    scene = crx.Scene(ctx)
    worlds = list(wl.instance_transforms(12, extent=10.0, seed=31))
    for k, w in enumerate(worlds):
        scene.add(mesh, w)

    def texture(k, dev):
        t = crx.Texture()
        t.width, t.height = 96, 64
        t.data = (dev_chains[k].ptr if dev else chains[k].ctypes.data)
        return t

    W, H = 640, 400
    view, proj = wl.instances_view(extent=10.0, width=W, height=H, distance=20.0)
    v = crx.make_view(view, proj, W, H)
    settings = crx.render_settings()
    settings.attribute = crx.DISPLAY_TEXTURE
    fb = ctx.alloc(8 * W * H)
    cb = ctx.alloc(8 * W * H)
    for frame, cleared in enumerate(([], [0, 4, 8])):
        for k in range(len(worlds)):
            scene.set_texture(k, None if k in cleared else texture(k % 3, True))
        print(f"[case] scene frame {frame}", flush=True)
        scene.render(v, fb, cb, settings)
        got_fb = fb.download(np.uint64, W * H)
        got_cb = cb.download(np.uint64, W * H)
        n = len(worlds)
        node_meshes = (crx.CMesh * n)()
        node_world = (crx.Mat4 * n)()
        for k in range(n):
            cm = host.cmesh(crx)
            cm.aabb = mesh.aabb()
            cm.address = 1
            if k not in cleared:
                cm.texture = texture(k % 3, False)
            node_meshes[k] = cm
            node_world[k] = worlds[k]
        uniq = (crx.CMesh * n)()
        nu = C.c_uint32()
        inst = (crx.CMesh * n)()
        xf = (crx.Mat4 * n)()
        pre = (C.c_uint64 * n)()
        nv = L.crx_oracle_prepare_frame(node_meshes, node_world, n, C.byref(v), int(settings.frustum_culling),
                                        0, uniq, C.byref(nu), inst, xf, pre)
        assert nv > 0
        fbw, _ = orc.draw(crx, list(uniq)[:nu.value], list(inst)[:nv], list(xf)[:nv], W, H, view, proj, True)
        cbw = orc.resolve(crx, list(inst)[:nv], list(pre)[:nv], fbw, W, H, view, proj, crx.DISPLAY_TEXTURE)
        assert np.array_equal(got_fb, fbw), describe_mismatch(got_fb, fbw, W)
        assert np.array_equal(got_cb, cbw), describe_mismatch(got_cb, cbw, W)
    # the same frame recorded into a hipGraph (crx_frame_*) takes the TEXTURE resolve too
    print("[case] scene TEXTURE frame graph", flush=True)
    gfb, gcb = ctx.alloc(8 * W * H), ctx.alloc(8 * W * H)
    frame = crx.Frame(scene, [v], [gfb], [gcb], settings)
    for _ in range(3):
        frame.launch()
        frame.wait()
        assert np.array_equal(gfb.download(np.uint64, W * H), got_fb)
        assert np.array_equal(gcb.download(np.uint64, W * H), got_cb)
    frame.close()
    scene.close()
    mesh.close()


@pytest.mark.parametrize("w, h", [(1, 1), (5, 3), (100, 60), (255, 2), (1024, 3), (4096, 3), (1000, 1000),
                                  (8192, 8192)])
def test_compute_mipmap_matches_the_oracle(ctx, w, h):
    """crx_texture_compute_mipmap against the oracle's computeMipMap, bit for
    bit, and nothing written past the chain."""
    print(f"[case] mipmap {w}x{h}", flush=True)
    L = orc.lib(crx)
    rng = np.random.default_rng(w * 31 + h)
    n = crx.texture_chain_texels(w, h)
    host = np.full(n + 64, 0xDEADBEEF, dtype=np.uint32)
    host[:w * h] = rng.integers(0, 2 ** 32, size=w * h, dtype=np.uint32)
    dev = ctx.to_device(host)
    ctx.compute_mipmap(dev, w, h)
    got = dev.download(np.uint32, n + 64)
    L.crx_oracle_compute_mipmap(host.ctypes.data, w, h)
    assert np.all(got[n:] == 0xDEADBEEF)
    bad = np.nonzero(got != host)[0]
    assert bad.size == 0, f"{bad.size} texels differ, first at {bad[:4]}"


def test_texture_with_a_computed_chain(ctx):
    """A terrain textured with a chain that crx_texture_compute_mipmap built
    from a random level 0, against the oracle with the oracle's own chain."""
    n = 513
    pos, idx, col = wl.heightfield_mesh(n, amplitude=20.0, scale=128.0)
    rng = np.random.default_rng(77)
    w, h = 700, 500
    texels = crx.texture_chain_texels(w, h)
    host = np.zeros(texels, dtype=np.uint32)
    host[:w * h] = rng.integers(0, 2 ** 32, size=w * h, dtype=np.uint32)
    dev = ctx.to_device(host)
    ctx.compute_mipmap(dev, w, h)
    orc.lib(crx).crx_oracle_compute_mipmap(host.ctypes.data, w, h)
    mesh = orc.HostMesh(crx, pos, idx, col, uvs=_terrain_uvs(pos, n, 6.0), texture=host, texture_size=(w, h))
    spec = SceneSpec().add(mesh, [crx.identity()])
    W, H = 1280, 720
    view, proj = wl.heightfield_view(n, W, H, altitude=30.0, pitch_deg=15.0, setback=10.0)
    print(f"[case] computed chain {w}x{h}", flush=True)
    want_fb, want_cb, _ = oracle_render(spec, W, H, view, proj, attribute=crx.DISPLAY_TEXTURE)
    g = GpuScene(ctx, spec)
    uv_ptr = g.pointers[0][3]
    g.pointers[0] = g.pointers[0][:4] + (dev.ptr,)     # the GPU-built chain in place of the uploaded copy
    assert uv_ptr is not None
    fb, cb, _ = g.render(W, H, view, proj, attribute=crx.DISPLAY_TEXTURE)
    assert np.array_equal(fb, want_fb), describe_mismatch(fb, want_fb, W)
    assert np.array_equal(cb, want_cb), describe_mismatch(cb, want_cb, W)
    assert np.unique(cb[fb != crx.CLEAR_PIXEL] & 0xFFFFFFFF).size > 1000


def test_resolve_rejects_bad_attributes_and_textures(ctx):
    fb = ctx.alloc(8 * 16)
    cb = ctx.alloc(8 * 16)
    view, proj = wl.sphere_view(4, 4)
    target = crx.make_target(fb, cb, 4, 4, view, proj)
    with pytest.raises(crx.CrxError) as e:
        ctx.resolve_visbuffer([], [], target, 8)          # no such display attribute
    assert e.value.status == 1
    pos, idx, col = wl.sphere_mesh(rings=4, segments=8)
    host = orc.HostMesh(crx, pos, idx, col)
    g = GpuScene(ctx, SceneSpec().add(host, [crx.identity()]))
    m = crx.CMesh()
    m.numTriangles = host.num_triangles
    m.positions, m.indices = g.pointers[0][0], g.pointers[0][1]
    m.isLoaded = 1
    m.texture.data = g.pointers[0][0]
    m.texture.width, m.texture.height = 0, 8
    with pytest.raises(crx.CrxError) as e:
        ctx.resolve_visbuffer([m], [0], target, crx.DISPLAY_TEXTURE)
    assert e.value.status == 1
    scene = crx.Scene(ctx)
    mesh = crx.Mesh(ctx, pos, idx, col)
    scene.add(mesh)
    t = crx.Texture()
    t.data, t.width, t.height = g.pointers[0][0], 8, -1
    with pytest.raises(crx.CrxError) as e:
        scene.set_texture(0, t)
    assert e.value.status == 1
    with pytest.raises(ValueError):
        mesh.set_attributes(uvs=np.zeros((3, 2), np.float32))
    for bad in ((0, 4), (4, -1), (65536, 65536)):
        with pytest.raises(crx.CrxError) as e:
            ctx.compute_mipmap(g.pointers[0][0], *bad)
        assert e.value.status == 1
    scene.close()
    mesh.close()
