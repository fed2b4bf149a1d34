"""crx_raster_visbuffer_noalloc: the three stages of one RasterArgs with the
draw records in the library's static device scratch.  Its visibility buffer
and queue counters must equal those of crx_raster_stage1/2/3 over the same
args bit for bit, on two calls in a row (the second with a smaller input, so
the scratch holds the first call's records when it starts), and args whose
records exceed the scratch are rejected."""
import ctypes as C

import numpy as np
import pytest

import crx
import oracle as orc
import workloads as wl
from crxtest import GpuScene, SceneSpec, describe_mismatch, prepare

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = crx.Context(0)
    yield c
    c.close()


def _bytes(arr) -> np.ndarray:
    return np.frombuffer(bytes(arr), dtype=np.uint8)


class Args:
    """Device CMesh lists, transforms, counters, queues and a framebuffer of one frame."""

    def __init__(self, ctx, g, W, H, view, proj):
        fr = prepare(g.spec, view, g.pointers)
        self.ctx, self.W, self.H = ctx, W, H
        uniq = (crx.CMesh * len(fr.unique))(*fr.unique)
        inst = (crx.CMesh * len(fr.instances))(*fr.instances)
        xf = (crx.Mat4 * len(fr.transforms))(*fr.transforms)
        self.keep = [ctx.to_device(_bytes(uniq)), ctx.to_device(_bytes(inst)), ctx.to_device(_bytes(xf))]
        self.counters = ctx.to_device(np.full(8, 0xDEADBEEF, dtype=np.uint32))
        self.nt = ctx.alloc(8 * (1 << 20))
        self.huge = ctx.alloc(16 * (1 << 18))
        self.fb = ctx.alloc(8 * W * H)
        a = crx.RasterArgs()
        a.meshes, a.numMeshes = self.keep[0].ptr, len(fr.unique)
        a.instances, a.numInstances = self.keep[1].ptr, len(fr.instances)
        a.transforms = self.keep[2].ptr
        c = self.counters.ptr
        a.numProcessedBatches, a.numProcessedBatches_nontrivial = c, c + 4
        a.hugeTrianglesCounter, a.numProcessedHugeTriangles = c + 8, c + 12
        a.nontrivialTrianglesCounter = c + 16
        a.hugeTriangles, a.nontrivialTrianglesList = self.huge.ptr, self.nt.ptr
        a.target = crx.make_target(self.fb, None, W, H, view, proj)
        a.nontrivialCapacity, a.hugeCapacity = 1 << 20, 1 << 18
        self.args = a

    def result(self):
        return self.fb.download(np.uint64, self.W * self.H), self.counters.download(np.uint32, 8)


def _scene(rings, instances, seed):
    hp, hi, hc = wl.heightfield_mesh(65, amplitude=4.0, scale=16.0)
    pos, idx, col = wl.sphere_mesh(rings=rings, segments=2 * rings, seed=seed)
    worlds = list(wl.instance_transforms(instances, extent=8.0, seed=seed + 1))
    for k in range(0, instances, 3):
        for r in range(3):
            worlds[k].m[r] = -worlds[k].m[r]                # mirrored instances
    return (SceneSpec().add(orc.HostMesh(crx, hp, hi, hc), [crx.identity()])
            .add(orc.HostMesh(crx, pos, idx, col), worlds))


def _run_both(ctx, spec, W, H, view, flags):
    g = GpuScene(ctx, spec)
    proj = wl.projection(W, H)
    ref = Args(ctx, g, W, H, view, proj)
    ctx.clear_framebuffer(ref.fb, W * H)
    ctx.raster_stage1(ref.args, flags)
    ctx.raster_stage2(ref.args)
    ctx.raster_stage3(ref.args)
    want_fb, want_c = ref.result()
    got = Args(ctx, g, W, H, view, proj)
    ctx.clear_framebuffer(got.fb, W * H)
    ctx.raster_visbuffer_noalloc(got.args, flags)
    fb, c = got.result()
    return fb, c, want_fb, want_c


@pytest.mark.parametrize("flags", [crx.RASTER_INSTANCED, 0])
def test_noalloc_equals_the_stage_calls(ctx, flags):
    view = crx.look_at((32.0, 3.5, 8.0), (32.0, 1.0, 30.0))
    # two calls in a row, the second with fewer instances, a smaller mesh and a smaller target
    for (rings, count, W, H, seed) in ((48, 40, 960, 544, 71), (16, 5, 320, 240, 73)):
        print(f"[case] noalloc flags={flags} rings={rings} instances={count} {W}x{H}", flush=True)
        spec = _scene(rings, count, seed)
        fb, c, want_fb, want_c = _run_both(ctx, spec, W, H, view, flags)
        assert np.array_equal(fb, want_fb), describe_mismatch(fb, want_fb, W)
        assert np.array_equal(c[:5], want_c[:5]) and c[5] == 0xDEADBEEF
        assert want_c[4] > 0                                       # stage 2 ran
        if W == 960:
            assert want_c[2] > 0                                   # and stage 3
        assert int((fb != crx.CLEAR_PIXEL).sum()) > W * H // 10


def test_noalloc_rejects_records_beyond_the_scratch(ctx):
    """20,000 meshes drawn over one instance each take 6.9 MB of records."""
    pos, idx, col = wl.sphere_mesh(rings=4, segments=8, seed=75)
    spec = SceneSpec().add(orc.HostMesh(crx, pos, idx, col), list(wl.instance_transforms(20000, seed=76)))
    g = GpuScene(ctx, spec)
    W, H = 64, 64
    a = Args(ctx, g, W, H, crx.look_at((0.0, 0.0, 50.0), (0.0, 0.0, 0.0)), wl.projection(W, H))
    with pytest.raises(crx.CrxError) as e:
        ctx.raster_visbuffer_noalloc(a.args, 0)
    assert "static scratch" in str(e.value)
    # the stage calls draw the same args (they allocate their records)
    ctx.raster_stage1(a.args, 0)
    ctx.raster_stage2(a.args)
    ctx.raster_stage3(a.args)
    ctx.synchronize()
