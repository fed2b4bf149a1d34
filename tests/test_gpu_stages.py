"""Kernel-level entry points: crx_raster_stage1/2/3 over caller-owned device
CMesh arrays, counters and queues (the reference's RasterArgs launches,
triangles_visbuffer.cu:755-835).  The visibility buffer and the queue counters
must equal crx_draw_triangles_visbuffer's and the oracle's bit for bit."""
import ctypes as C

import numpy as np
import pytest

import crx
import oracle as orc
import workloads as wl
from crxtest import CompressedMesh, GpuScene, SceneSpec, describe_mismatch, oracle_render, prepare

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = crx.Context(0)
    yield c
    c.close()


def _bytes(arr) -> np.ndarray:
    return np.frombuffer(bytes(arr), dtype=np.uint8)


class StageRun:
    """Device copies of one frame's CMesh lists and transforms, the five
    counters, the two queues and a cleared framebuffer."""

    def __init__(self, ctx, g: GpuScene, W, H, view, proj, nt_cap=1 << 20, huge_cap=1 << 18):
        fr = prepare(g.spec, view, g.pointers)
        self.ctx, self.W, self.H = ctx, W, H
        uniq = (crx.CMesh * len(fr.unique))(*fr.unique)
        inst = (crx.CMesh * len(fr.instances))(*fr.instances)
        xf = (crx.Mat4 * len(fr.transforms))(*fr.transforms)
        self.d_uniq, self.d_inst, self.d_xf = ctx.to_device(_bytes(uniq)), ctx.to_device(_bytes(inst)), ctx.to_device(_bytes(xf))
        self.counters = ctx.to_device(np.full(8, 0xDEADBEEF, dtype=np.uint32))
        self.nt = ctx.alloc(8 * nt_cap)
        self.huge = ctx.alloc(16 * huge_cap)
        self.fb = ctx.alloc(8 * W * H)
        ctx.clear_framebuffer(self.fb, W * H)
        a = crx.RasterArgs()
        a.meshes, a.numMeshes = self.d_uniq.ptr, len(fr.unique)
        a.instances, a.numInstances = self.d_inst.ptr, len(fr.instances)
        a.transforms = self.d_xf.ptr
        c = self.counters.ptr
        a.numProcessedBatches, a.numProcessedBatches_nontrivial = c, c + 4
        a.hugeTrianglesCounter, a.numProcessedHugeTriangles = c + 8, c + 12
        a.nontrivialTrianglesCounter = c + 16
        a.hugeTriangles, a.nontrivialTrianglesList = self.huge.ptr, self.nt.ptr
        a.target = crx.make_target(self.fb, None, W, H, view, proj)
        a.nontrivialCapacity, a.hugeCapacity = nt_cap, huge_cap
        self.args = a

    def run(self, flags):
        self.ctx.raster_stage1(self.args, flags)
        self.ctx.raster_stage2(self.args)
        self.ctx.raster_stage3(self.args)
        fb = self.fb.download(np.uint64, self.W * self.H)
        c = self.counters.download(np.uint32, 8)
        # (numProcessedBatches, cursor2, huge counter, cursor3, nontrivial counter); word 5.. untouched
        return fb, c


def _compare(ctx, spec, W, H, view, proj, instanced, compressed=False):
    print(f"[case] stages W={W} H={H} instanced={instanced} compressed={compressed}", flush=True)
    g = GpuScene(ctx, spec)
    want_fb, _, st = g.render(W, H, view, proj, instanced=instanced)
    run = StageRun(ctx, g, W, H, view, proj)
    flags = (crx.RASTER_INSTANCED if instanced else 0) | (crx.RASTER_COMPRESSED if compressed else 0)
    fb, c = run.run(flags)
    assert np.array_equal(fb, want_fb), describe_mismatch(fb, want_fb, W)
    assert (int(c[4]), int(c[2])) == (st.num_nontrivial, st.num_huge)
    assert c[5] == 0xDEADBEEF                # nothing outside the five counters is written
    ora_fb, _, ost = oracle_render(spec, W, H, view, proj, instanced=instanced)
    assert np.array_equal(fb, ora_fb), describe_mismatch(fb, ora_fb, W)
    return fb, st


def test_stages_sphere_and_instances(ctx):
    pos, idx, col = wl.sphere_mesh(rings=96, segments=192, seed=51)
    worlds = [crx.identity()] + list(wl.instance_transforms(7, extent=4.0, smin=0.3, smax=0.8, seed=52))
    for k in (2, 5):
        for r in range(3):
            worlds[k].m[r] = -worlds[k].m[r]                    # mirrored instances
    spec = SceneSpec().add(orc.HostMesh(crx, pos, idx, col), worlds)
    view, proj = wl.sphere_view(1280, 720, radius=1.6)
    for instanced in (True, False):
        _compare(ctx, spec, 1280, 720, view, proj, instanced)


def test_stages_closeup_all_three_stages(ctx):
    hp, hi, hc = wl.heightfield_mesh(129, amplitude=6.0, scale=32.0)
    pos, idx, col = wl.sphere_mesh(rings=24, segments=48, seed=53)
    spec = (SceneSpec().add(orc.HostMesh(crx, hp, hi, hc), [crx.identity()])
            .add(orc.HostMesh(crx, pos, idx, col), list(wl.instance_transforms(5, extent=30.0, seed=54))))
    W, H = 960, 544
    view = crx.look_at((64.0, 4.5, 20.0), (64.0, 2.0, 40.0))
    _, st = _compare(ctx, spec, W, H, view, wl.projection(W, H), True)
    assert st.num_nontrivial > 0 and st.num_huge > 0


def test_stages_compressed_and_non_indexed(ctx):
    pos, idx, col = wl.sphere_mesh(rings=64, segments=96, seed=55)
    spec = SceneSpec().add(CompressedMesh(pos, idx, col), list(wl.instance_transforms(4, extent=3.0, seed=56)))
    view, proj = wl.sphere_view(800, 600, radius=2.0)
    _compare(ctx, spec, 800, 600, view, proj, True, compressed=True)
    flat = SceneSpec().add(orc.HostMesh(crx, pos[idx.reshape(-1)], None, col[idx.reshape(-1)]),
                           list(wl.instance_transforms(3, extent=3.0, seed=57)))
    _compare(ctx, flat, 800, 600, view, proj, False)


def test_stages_order_and_format_rules(ctx):
    pos, idx, col = wl.sphere_mesh(rings=32, segments=64, seed=58)
    spec = SceneSpec().add(orc.HostMesh(crx, pos, idx, col), [crx.identity()])
    g = GpuScene(ctx, spec)
    view, proj = wl.sphere_view(320, 240)
    run = StageRun(ctx, g, 320, 240, view, proj)
    # a format that differs from the flags draws nothing
    fb, c = run.run(crx.RASTER_INSTANCED | crx.RASTER_COMPRESSED)
    assert (fb == crx.CLEAR_PIXEL).all() and c[4] == 0 and c[2] == 0
    # an instance range outside the instance list draws nothing for that mesh
    bad = crx.CMesh.from_buffer_copy(bytes(run.d_uniq.download(np.uint8, C.sizeof(crx.CMesh))))
    bad.instances.count = 5
    run.d_uniq.upload(_bytes(bad))
    ctx.clear_framebuffer(run.fb, 320 * 240)
    fb, c = run.run(crx.RASTER_INSTANCED)
    assert (fb == crx.CLEAR_PIXEL).all()
    # the same args with flags that match draw the frame
    bad.instances.count = 1
    run.d_uniq.upload(_bytes(bad))
    fb, c = run.run(crx.RASTER_INSTANCED)
    want, _, _ = g.render(320, 240, view, proj)
    assert np.array_equal(fb, want)


def _interleave_scenes(ctx):
    """Two draws with different args (scenes, views, sizes, counters and
    queues), both reaching stages 2 and 3, and their single-draw buffers."""
    hp, hi, hc = wl.heightfield_mesh(65, amplitude=4.0, scale=16.0)
    pos, idx, col = wl.sphere_mesh(rings=24, segments=48, seed=61)
    spec_a = SceneSpec().add(orc.HostMesh(crx, hp, hi, hc), [crx.identity()])
    spec_b = (SceneSpec().add(orc.HostMesh(crx, pos, idx, col), list(wl.instance_transforms(6, extent=6.0, seed=62)))
              .add(orc.HostMesh(crx, hp, hi, hc), [crx.identity()]))
    ga, gb = GpuScene(ctx, spec_a), GpuScene(ctx, spec_b)
    va = crx.look_at((32.0, 3.5, 8.0), (32.0, 1.0, 30.0))
    vb = crx.look_at((20.0, 2.5, 10.0), (40.0, 0.5, 30.0))
    (Wa, Ha), (Wb, Hb) = (640, 360), (512, 384)
    pa, pb = wl.projection(Wa, Ha), wl.projection(Wb, Hb)
    want_a, _, sa = ga.render(Wa, Ha, va, pa)
    want_b, _, sb = gb.render(Wb, Hb, vb, pb)
    assert sa.num_huge > 0 and sb.num_huge > 0
    ra, rb = StageRun(ctx, ga, Wa, Ha, va, pa), StageRun(ctx, gb, Wb, Hb, vb, pb)
    return (ra, want_a, sa), (rb, want_b, sb)


def _check_interleaved(a, b):
    for run, want, st in (a, b):
        fb = run.fb.download(np.uint64, run.W * run.H)
        c = run.counters.download(np.uint32, 8)
        assert np.array_equal(fb, want), describe_mismatch(fb, want, run.W)
        assert (int(c[4]), int(c[2])) == (st.num_nontrivial, st.num_huge)


def test_stages_depend_only_on_their_args(ctx):
    """Like the reference's stage kernels, each call depends only on its
    args: two draws interleaved stage by stage on one context each equal
    their own single draw."""
    a, b = _interleave_scenes(ctx)
    ra, rb = a[0], b[0]
    ctx.raster_stage1(ra.args, crx.RASTER_INSTANCED)
    ctx.raster_stage1(rb.args, crx.RASTER_INSTANCED)
    ctx.raster_stage2(ra.args)
    ctx.raster_stage2(rb.args)
    ctx.raster_stage3(rb.args)
    ctx.raster_stage3(ra.args)
    _check_interleaved(a, b)


def test_stages_interleaved_on_two_streams(ctx):
    """The same interleaving with each draw on its own stream: a call on the
    other stream waits for the previous call's launches before its records
    replace theirs."""
    hip = C.CDLL("libamdhip64.so.7")   # the runtime libcrx already loaded (same soname)
    streams = []
    for _ in range(2):
        h = C.c_void_p()
        assert hip.hipStreamCreate(C.byref(h)) == 0
        streams.append(h)
    a, b = _interleave_scenes(ctx)
    ra, rb = a[0], b[0]
    for rnd in range(3):
        print(f"[case] interleaved stages on two streams, round {rnd}", flush=True)
        ctx.clear_framebuffer(ra.fb, ra.W * ra.H, streams[0])
        ctx.clear_framebuffer(rb.fb, rb.W * rb.H, streams[1])
        ctx.raster_stage1(ra.args, crx.RASTER_INSTANCED, streams[0])
        ctx.raster_stage1(rb.args, crx.RASTER_INSTANCED, streams[1])
        ctx.raster_stage2(ra.args, streams[0])
        ctx.raster_stage2(rb.args, streams[1])
        ctx.raster_stage3(rb.args, streams[1])
        ctx.raster_stage3(ra.args, streams[0])
        for h in streams:
            ctx.synchronize(h)
        _check_interleaved(a, b)
    for h in streams:
        assert hip.hipStreamDestroy(h) == 0


def test_stages_queue_capacity_reports_overflow(ctx):
    hp, hi, hc = wl.heightfield_mesh(65, amplitude=4.0, scale=16.0)
    spec = SceneSpec().add(orc.HostMesh(crx, hp, hi, hc), [crx.identity()])
    W, H = 640, 360
    view = crx.look_at((32.0, 3.5, 8.0), (32.0, 1.0, 30.0))
    g = GpuScene(ctx, spec)
    run = StageRun(ctx, g, W, H, view, wl.projection(W, H), nt_cap=4, huge_cap=2)
    fb, c = run.run(crx.RASTER_INSTANCED)
    _, _, st = g.render(W, H, view, wl.projection(W, H))
    assert c[4] == st.num_nontrivial and st.num_nontrivial > 4   # counts every entry, stores the first 4


def test_atomic_min_probe(ctx):
    """The atomicMin probe writes only inside its 2^logEntries entries and
    leaves each entry at the minimum of the values aimed at it."""
    n = 1 << 12
    buf = ctx.alloc(8 * (n + 64))
    ctx.clear_framebuffer(buf, n + 64)
    for scattered in (True, False):
        issued = ctx.probe_atomic_min(buf, 12, 4, scattered)
        assert issued == ctx.num_cus * 8 * 256 * 4
    out = buf.download(np.uint64, n + 64)
    assert (out[n:] == crx.CLEAR_PIXEL).all()                 # nothing past the probed entries
    assert (out[:n] != crx.CLEAR_PIXEL).sum() > n // 2
    with pytest.raises(crx.CrxError):
        ctx.probe_atomic_min(buf, 40, 1, True)


def test_read_probe(ctx):
    """The streaming-read probe reads ragged sizes (vector loop and tail),
    leaves its source alone, and rejects unaligned sizes."""
    n = 4099 * 16
    src = ctx.to_device(np.arange(n // 4, dtype=np.uint32))
    sink = ctx.alloc(16)
    sink.upload(np.full(4, 7, dtype=np.uint32))
    for nbytes in (16, 4096, n):
        ctx.probe_read(src, nbytes, sink)
    ctx.synchronize()
    assert np.array_equal(src.download(np.uint32, n // 4), np.arange(n // 4, dtype=np.uint32))
    with pytest.raises(crx.CrxError):
        ctx.probe_read(src, 20, sink)


def test_division_sequences_are_ieee(ctx):
    """Stage 1 computes reciprocals as v_rcp_f32 plus one FMA Newton step and
    quotients with Markstein's correction; both must equal the IEEE division
    bit for bit: every float in the reciprocal's range (4.2e9 values), 2^30
    hashed quotient operand pairs, and 2^30 pairs with a numerator below
    2^-60 through the screen term q * 0.5 + 0.5."""
    rcp_bad, div_bad, small_bad = ctx.check_division(1 << 30)
    assert rcp_bad == 0 and div_bad == 0 and small_bad == 0, (rcp_bad, div_bad, small_bad)
