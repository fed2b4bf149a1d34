"""Parity at the benchmark workloads' stated sizes, one GPU frame against the
CPU oracle bit for bit (tolerance: none).

* The 400M-triangle heightfield (14142 x 14142 grid, 399,935,762 triangles)
  at 1920x1080, visibility and colour buffers.  Its index buffer is 4.8 GB,
  so indices and positions are addressed past 2^32 bytes; the far-edge view
  is resolved with TEXTURE too.
* 3000 instances of the 1M-triangle displaced sphere at 1920x1080: triangle
  ids reach 3.1e9, formed with the reference's 32-bit
  instanceIndex * numTriangles product (triangles_visbuffer.cu:252-254).
* 64 mixed-resolution views of the heightfield (256^2 to 3840x2160, 16
  close-ups that reach stages 2 and 3) in one batched pass: every view
  equals its plain single-view frame, and every 8th view the oracle's.

The generated scenes are shared by the tests of this module; every step
prints what it is doing.
"""
import os
import time

import numpy as np
import pytest

import crx
import oracle as orc
import workloads as wl
from crxtest import SceneSpec, describe_mismatch, oracle_render

pytestmark = pytest.mark.gpu

W, H = 1920, 1080
N = 14142
THREADS = int(os.environ.get("OMP_NUM_THREADS", "16"))


def say(msg):
    print(f"[fullsize] {msg}", flush=True)


@pytest.fixture(scope="module")
def ctx():
    c = crx.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def heightfield(ctx):
    t0 = time.perf_counter()
    pos, idx, col = wl.heightfield_mesh(N, threads=THREADS)
    say(f"heightfield {N}x{N}: {len(idx)} triangles, {idx.nbytes / 1e9:.2f} GB of indices, "
        f"generated in {time.perf_counter() - t0:.1f} s")
    assert len(idx) == 399_935_762 and idx.nbytes > 1 << 32
    mesh = crx.Mesh(ctx, pos, idx, col)
    scene = crx.Scene(ctx)
    scene.add(mesh, crx.identity())
    say(f"heightfield resident on the device after {time.perf_counter() - t0:.1f} s")
    yield pos, idx, col, scene
    scene.close()
    mesh.close()


def _compare(name, fb, cb, want_fb, want_cb, width):
    assert np.array_equal(fb, want_fb), f"{name}: " + describe_mismatch(fb, want_fb, width)
    assert np.array_equal(cb, want_cb), f"{name} colours: " + describe_mismatch(cb, want_cb, width)


def _far_edge_view(n):
    """The benchmark's oblique view turned around: from beyond the far edge
    (z = n - 1) looking back along -z, so the visible triangles are the last
    rows of the grid, whose indices lie past the first 2^32 bytes."""
    x = 0.5 * (n - 1)
    eye = (x, 300.0, (n - 1) + 200.0)
    pitch = np.radians(40.0)
    center = (x, 300.0 - float(np.sin(pitch)), (n - 1) + 200.0 - float(np.cos(pitch)))
    return crx.look_at(eye, center, (0.0, 1.0, 0.0)), wl.projection(W, H)


@pytest.mark.parametrize("which", ["benchmark-view", "far-edge-view"])
def test_heightfield_400m_1080p(ctx, heightfield, which):
    pos, idx, col, scene = heightfield
    view, proj = wl.heightfield_view(N, W, H) if which == "benchmark-view" else _far_edge_view(N)
    fb, cb = ctx.alloc(8 * W * H), ctx.alloc(8 * W * H)
    st = scene.render(crx.make_view(view, proj, W, H), fb, cb, crx.render_settings())
    got_fb, got_cb = fb.download(np.uint64, W * H), cb.download(np.uint64, W * H)
    say(f"{which}: GPU frame done ({st.num_nontrivial} stage-2 entries); oracle frame on {THREADS} threads ...")
    t0 = time.perf_counter()
    spec = SceneSpec().add(orc.HostMesh(crx, pos, idx, col), [crx.identity()])
    want_fb, want_cb, ost = oracle_render(spec, W, H, view, proj, attribute=crx.DISPLAY_VERTEX_COLORS,
                                          threads=THREADS)
    say(f"oracle frame in {time.perf_counter() - t0:.1f} s")
    _compare(which, got_fb, got_cb, want_fb, want_cb, W)
    assert st.num_nontrivial == ost.num_nontrivial and st.num_huge == ost.num_huge
    covered = int((got_fb != crx.CLEAR_PIXEL).sum())
    ids = got_fb[got_fb != crx.CLEAR_PIXEL] & 0xFFFFFFFFF
    say(f"bit-exact: {covered} covered pixels, triangle ids {int(ids.min())} .. {int(ids.max())}")
    assert covered > W * H // 2
    if which == "far-edge-view":
        # index bytes 12 * t lie past 2^32 for t > 357,913,941: (nearly)
        # every covered pixel's triangle
        past = int(np.count_nonzero(ids > (1 << 32) // 12))
        say(f"{past} of {ids.size} covered pixels show triangles whose indices lie past 2^32 bytes")
        assert past > 0.99 * ids.size
        _texture_resolve(ctx, scene, pos, idx, col, got_fb, view, proj)


def _texture_resolve(ctx, scene, pos, idx, col, fb_host, view, proj):
    """The TEXTURE resolve of the same visibility buffer: uvs over the whole
    grid (repeated 64 times) and a random 4096 x 4096 mip chain, against the
    oracle's resolve of that buffer."""
    from crxtest import mip_chain, mip_shapes
    say("TEXTURE resolve of the far-edge frame ...")
    uvs = np.empty((len(pos), 2), dtype=np.float32)
    np.multiply(pos[:, 0], np.float32(64.0 / (N - 1)), out=uvs[:, 0])
    np.multiply(pos[:, 2], np.float32(64.0 / (N - 1)), out=uvs[:, 1])
    rng = np.random.default_rng(5)
    chain = mip_chain([rng.integers(0, 2 ** 32, size=s, dtype=np.uint32) for s in mip_shapes(4096, 4096)])
    d_uvs, d_tex, d_fb = ctx.to_device(uvs), ctx.to_device(chain), ctx.to_device(fb_host)
    d_cb = ctx.alloc(8 * W * H)
    ctx.clear_framebuffer(d_cb, W * H)
    host = orc.HostMesh(crx, pos, idx, col, uvs=uvs, texture=chain, texture_size=(4096, 4096))
    m = host.cmesh(crx)
    m_dev = crx.CMesh.from_buffer_copy(m)
    p, i, c = scene.nodes[0][0].buffers()
    m_dev.positions, m_dev.indices, m_dev.colors = p, i, c
    m_dev.uvs, m_dev.texture.data = d_uvs.ptr, d_tex.ptr
    target = crx.make_target(d_fb, d_cb, W, H, view, proj)
    ctx.resolve_visbuffer([m_dev], [0], target, crx.DISPLAY_TEXTURE)
    got = d_cb.download(np.uint64, W * H)
    want = orc.resolve(crx, [m], [0], fb_host, W, H, view, proj, crx.DISPLAY_TEXTURE, threads=THREADS)
    assert np.array_equal(got, want), "TEXTURE: " + describe_mismatch(got, want, W)
    cov = fb_host != crx.CLEAR_PIXEL
    say(f"TEXTURE bit-exact: {np.unique(got[cov] & 0xFFFFFFFF).size} distinct colours")
    assert np.unique(got[cov] & 0xFFFFFFFF).size > 100_000
    for b in (d_uvs, d_tex, d_fb, d_cb):
        b.free()


def test_instances_3000x1m_1080p(ctx):
    pos, idx, col = wl.sphere_mesh()
    worlds = list(wl.instance_transforms(3000))
    view, proj = wl.instances_view()
    mesh = crx.Mesh(ctx, pos, idx, col)
    scene = crx.Scene(ctx)
    for w in worlds:
        scene.add(mesh, w)
    fb, cb = ctx.alloc(8 * W * H), ctx.alloc(8 * W * H)
    try:
        st = scene.render(crx.make_view(view, proj, W, H), fb, cb, crx.render_settings())
        got_fb, got_cb = fb.download(np.uint64, W * H), cb.download(np.uint64, W * H)
        say(f"GPU frame of 3000 instances done; oracle frame (3.1e9 triangle-instances) on {THREADS} threads ...")
        t0 = time.perf_counter()
        spec = SceneSpec().add(orc.HostMesh(crx, pos, idx, col), worlds)
        want_fb, want_cb, ost = oracle_render(spec, W, H, view, proj, attribute=crx.DISPLAY_VERTEX_COLORS,
                                              threads=THREADS)
        say(f"oracle frame in {time.perf_counter() - t0:.1f} s")
        _compare("instances", got_fb, got_cb, want_fb, want_cb, W)
        assert st.num_nontrivial == ost.num_nontrivial and st.num_huge == ost.num_huge
        ids = got_fb[got_fb != crx.CLEAR_PIXEL] & 0xFFFFFFFFF
        say(f"bit-exact: {ids.size} covered pixels, triangle ids up to {int(ids.max())}")
        assert ids.size > 100_000
        assert int(ids.max()) >= 1 << 31      # ids formed past the signed 32-bit range
    finally:
        scene.close()
        mesh.close()


def test_multiview_64_views_of_the_heightfield(ctx, heightfield):
    pos, idx, col, scene = heightfield
    vs = wl.multiview_views(N, pos, count=64)
    views = [crx.make_view(v, p, w, h) for (v, p, w, h, _) in vs]
    sizes = [(v.width, v.height) for v in views]
    fbs = [ctx.alloc(8 * w * h) for (w, h) in sizes]
    cbs = [ctx.alloc(8 * w * h) for (w, h) in sizes]
    settings = crx.render_settings()
    try:
        st = scene.render_views(views, fbs, cbs, settings)
        batched = [(fbs[k].download(np.uint64, w * h), cbs[k].download(np.uint64, w * h))
                   for k, (w, h) in enumerate(sizes)]
        say(f"batched frame of 64 views done ({st.num_nontrivial} stage-2, {st.num_huge} stage-3 entries)")
        assert st.num_nontrivial > 0 and st.num_huge > 0
        for k, v in enumerate(views):
            scene.render(v, fbs[k], cbs[k], settings)
            w, h = sizes[k]
            _compare(f"view {k} ({w}x{h}) batched vs plain", batched[k][0], batched[k][1],
                     fbs[k].download(np.uint64, w * h), cbs[k].download(np.uint64, w * h), w)
        say("every batched view equals its plain frame; oracle on every 8th view ...")
        spec = SceneSpec().add(orc.HostMesh(crx, pos, idx, col), [crx.identity()])
        for k in range(0, 64, 8):
            (v, p, w, h, close) = vs[k]
            t0 = time.perf_counter()
            want_fb, want_cb, _ = oracle_render(spec, w, h, v, p, attribute=crx.DISPLAY_VERTEX_COLORS,
                                                threads=THREADS)
            _compare(f"view {k} ({w}x{h}, close-up {close}) vs oracle", batched[k][0], batched[k][1],
                     want_fb, want_cb, w)
            say(f"view {k} ({w}x{h}, close-up {close}) equals the oracle ({time.perf_counter() - t0:.1f} s)")
    finally:
        for b in fbs + cbs:
            b.free()
