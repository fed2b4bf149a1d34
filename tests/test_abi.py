"""The C-ABI boundary without a GPU: the libraries load, export every symbol
the headers declare, lay out the reference's structs byte for byte, run their
host logic (matrices, tuning, error paths) and never link the oracle."""
import ctypes as C
import re
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

import crx
import oracle as orc

ROOT = Path(__file__).resolve().parent.parent
INCLUDE = ROOT / "include"


def declared_functions(header: Path):
    text = header.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(crx_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("header,lib", [("crx.h", "libcrx.so"), ("crx_gen.h", "libcrx_gen.so")])
def test_library_exports_every_declared_symbol(header, lib):
    path = ROOT / "curast-ooc_amd" / "lib" / lib
    handle = C.CDLL(str(path))
    names = declared_functions(INCLUDE / header)
    assert len(names) > 5
    missing = [n for n in names if not hasattr(handle, n)]
    assert not missing, missing


def test_product_does_not_link_the_oracle():
    out = subprocess.run(["readelf", "-d", str(ROOT / "curast-ooc_amd" / "lib" / "libcrx.so")],
                         stdout=subprocess.PIPE, text=True, check=True).stdout
    assert "oracle" not in out
    syms = subprocess.run(["nm", "-D", str(ROOT / "curast-ooc_amd" / "lib" / "libcrx.so")],
                          stdout=subprocess.PIPE, text=True, check=True).stdout
    assert "crx_oracle" not in syms


# Offsets of the reference's CMesh / RenderTarget (src/kernels/HostDeviceInterface.h:131-143,
# :196-234) with GLM's packed vec3/mat4 (4-byte alignment).
CMESH_OFFSETS = {"numTriangles": 0, "indices": 8, "positions": 16, "cummulativeTriangleCount": 24, "aabb": 32,
                 "compressionFactor": 56, "index_min": 68, "bitsPerIndex": 72, "instances": 76, "world": 84,
                 "id": 148, "uvs": 152, "normals": 160, "colors": 168, "firstTriangle": 176, "numVertices": 180,
                 "index_max": 184, "address": 192, "texture": 200, "isLoaded": 248, "flipTriangles": 249,
                 "compressed": 250, "impostor": 256}
TARGET_OFFSETS = {"framebuffer": 0, "colorbuffer": 8, "width": 16, "height": 20, "view": 24, "viewI": 88,
                  "proj": 152, "cameraPos": 216, "f": 228, "aspect": 232, "debug": 236}
# RasterArgs (HostDeviceInterface.h:309-323, 336 B) followed by the two queue capacities.
ARGS_OFFSETS = {"meshes": 0, "numMeshes": 8, "instances": 16, "numInstances": 24, "transforms": 32,
                "numProcessedBatches": 40, "numProcessedBatches_nontrivial": 48, "hugeTriangles": 56,
                "hugeTrianglesCounter": 64, "numProcessedHugeTriangles": 72, "nontrivialTrianglesCounter": 80,
                "nontrivialTrianglesList": 88, "target": 96, "nontrivialCapacity": 336, "hugeCapacity": 340}


def test_struct_layout_matches_reference(tmp_path):
    fields = ([("crx_cmesh", k) for k in CMESH_OFFSETS] + [("crx_render_target", k) for k in TARGET_OFFSETS] +
              [("crx_raster_args", k) for k in ARGS_OFFSETS])
    src = ['#include <stdio.h>', '#include <stddef.h>', '#include "crx.h"', "int main(void){"]
    src.append('printf("%zu %zu %zu\\n", sizeof(crx_cmesh), sizeof(crx_render_target), sizeof(crx_raster_args));')
    for s, f in fields:
        src.append(f'printf("%zu\\n", offsetof({s}, {f}));')
    src.append("return 0;}")
    (tmp_path / "probe.c").write_text("\n".join(src))
    subprocess.run(["gcc", "-std=c99", f"-I{INCLUDE}", str(tmp_path / "probe.c"), "-o", str(tmp_path / "probe")],
                   check=True)
    out = subprocess.run([str(tmp_path / "probe")], stdout=subprocess.PIPE, text=True, check=True).stdout.split()
    assert out[0] == "280" and out[1] == "240" and out[2] == "344"
    got = [int(x) for x in out[3:]]
    want = list(CMESH_OFFSETS.values()) + list(TARGET_OFFSETS.values()) + list(ARGS_OFFSETS.values())
    assert got == want
    # the Python binding mirrors the same layout
    for name, off in CMESH_OFFSETS.items():
        assert getattr(crx.CMesh, name).offset == off
    for name, off in TARGET_OFFSETS.items():
        assert getattr(crx.RenderTarget, name).offset == off
    for name, off in ARGS_OFFSETS.items():
        assert getattr(crx.RasterArgs, name).offset == off


def test_tuning_defaults_are_the_reference_constants():
    t = crx.default_tuning()
    # TILE_SIZE, TRIANGLES_PER_SWEEP, THRESHOLD_SMALL, MAX_* (HostDeviceInterface.h:261-266)
    assert (t.small_threshold, t.tile_size, t.max_nontrivial, t.max_huge) == (128, 64, 5_000_000, 5_000_000)


def test_no_device_error_path():
    """Without a GPU the context constructor reports CRX_NO_DEVICE with a
    message instead of exiting (the reference exits, CURuntime.h:38-59)."""
    try:
        import torch  # noqa: F401
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a GPU is present")
    with pytest.raises(crx.CrxError) as e:
        crx.Context(0)
    assert e.value.status == 5
    assert "device" in str(e.value)
    assert crx.lib().crx_status_string(3) == b"CRX_QUEUE_OVERFLOW"
    assert crx.lib().crx_abi_version() == 1


def test_host_matrix_math_matches_oracle_bit_for_bit():
    """The product's host GLM-order math and the oracle's independent
    restatement agree bit for bit (view*world, inverse(view))."""
    L = orc.lib(crx)
    rng = np.random.default_rng(7)
    for _ in range(300):
        a = crx.Mat4.from_numpy(rng.normal(size=(4, 4)).astype(np.float32))
        b = crx.Mat4.from_numpy(rng.normal(size=(4, 4)).astype(np.float32))
        p1, p2 = crx.Mat4(), crx.Mat4()
        crx.lib().crx_mat4_mul(C.byref(a), C.byref(b), C.byref(p1))
        L.crx_oracle_mat4_mul(C.byref(a), C.byref(b), C.byref(p2))
        assert bytes(p1) == bytes(p2)
        i1, i2 = crx.Mat4(), crx.Mat4()
        crx.lib().crx_mat4_inverse(C.byref(a), C.byref(i1))
        L.crx_oracle_mat4_inverse(C.byref(a), C.byref(i2))
        assert bytes(i1) == bytes(i2)


def test_projection_matrix_formula():
    """Camera::createProjectionMatrix: [f/a,0,0,0; 0,f,0,0; 0,0,0,-1; 0,0,near,0]
    column-major with f = 1/tan(fovy/2) in float."""
    fovy, aspect, near = np.float32(np.radians(60)), np.float32(16 / 9), np.float32(0.01)
    m = np.frombuffer(bytes(crx.create_projection_matrix(near, fovy, aspect)), dtype=np.float32)
    f = np.float32(1.0) / np.float32(np.tan(np.float32(fovy / np.float32(2))))
    want = np.zeros(16, dtype=np.float32)
    want[0], want[5], want[11], want[14] = f / aspect, f, -1.0, near
    assert abs(float(m[5]) - float(f)) <= float(np.spacing(f))
    np.testing.assert_allclose(m, want, rtol=1e-6)


def test_look_at_is_a_rigid_view_matrix():
    v = crx.look_at((3.0, 4.0, 5.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0)).to_numpy().astype(np.float64)
    R = v[:3, :3]
    np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-6)
    eye = v @ np.array([3, 4, 5, 1.0])
    np.testing.assert_allclose(eye[:3], 0, atol=1e-5)
    target = v @ np.array([0, 0, 0, 1.0])
    assert target[2] < 0 and abs(target[0]) < 1e-5 and abs(target[1]) < 1e-5


def test_cpu_suite_has_no_gpu_side_effects():
    """The oracle and the generators run without the HIP runtime touching a device."""
    import workloads as wl
    pos, idx, col = wl.sphere_mesh(rings=8, segments=16)
    assert pos.shape == (9 * 16, 3) and idx.shape == (2 * 8 * 16, 3)
    assert idx.max() < len(pos)


def test_texture_chain_texels_host_logic():
    """crx_texture_chain_texels (host only): the texels of levels 0 ..
    max(floor(log2(max(w, h))) - 1, 7), each ((w + 1) / 2) x ((h + 1) / 2) of
    the level before; non-positive sizes and 2^31 texels or more are refused."""
    def want(w, h):
        L = max(w, h).bit_length() - 1
        total = 0
        for _ in range(max(L, 8)):
            total += w * h
            w, h = (w + 1) // 2, (h + 1) // 2
        return total
    for w, h in [(1, 1), (1, 7), (5, 3), (64, 64), (255, 2), (256, 9), (1024, 3), (4096, 3), (16384, 16384),
                 (46340, 46340), (1, 1 << 30)]:
        assert crx.texture_chain_texels(w, h) == want(w, h), (w, h)
    assert crx.texture_chain_texels(256, 256) == sum((256 >> k) ** 2 for k in range(8))
    for bad in ((0, 4), (4, 0), (-3, 4), (65536, 65536), (2, 1 << 30)):
        with pytest.raises(crx.CrxError) as e:
            crx.texture_chain_texels(*bad)
        assert e.value.status == 1


def _views_for_cull():
    import math
    import workloads as wl
    out = []
    rng = np.random.default_rng(7)
    for k in range(24):
        eye = tuple(float(x) for x in rng.uniform(-300.0, 300.0, 3))
        tgt = tuple(float(x) for x in rng.uniform(-100.0, 100.0, 3))
        W, H = [(1920, 1080), (256, 256), (640, 200), (3840, 2160)][k % 4]
        fovy = math.radians([60.0, 20.0, 90.0, 5.0][k % 4])
        out.append(crx.make_view(crx.look_at(eye, tgt), wl.projection(W, H, fovy=fovy), W, H))
    return out


def _cull_boxes(seed):
    """A grid of tiles (neighbouring ids are neighbours in space), scattered
    boxes, flat, inverted and huge ones."""
    rng = np.random.default_rng(seed)
    boxes = []
    for r in range(30):
        for c in range(30):
            lo = np.array([c * 20.0 - 300.0, rng.uniform(-20, 0), r * 20.0 - 300.0], np.float32)
            boxes.append((lo, lo + np.array([20.0, rng.uniform(0, 40), 20.0], np.float32)))
    for _ in range(400):
        lo = rng.uniform(-400, 400, 3).astype(np.float32)
        boxes.append((lo, lo + rng.uniform(0, 60, 3).astype(np.float32)))
    for _ in range(40):
        lo = rng.uniform(-200, 200, 3).astype(np.float32)
        boxes.append((lo, lo))                                    # a point
        boxes.append((lo + 5.0, lo))                              # inverted
        boxes.append((lo - 5000.0, lo + 5000.0))                  # huge
    out = []
    for lo, hi in boxes:
        b = crx.Box3()
        b.min.x, b.min.y, b.min.z = (float(v) for v in lo)
        b.max.x, b.max.y, b.max.z = (float(v) for v in hi)
        out.append(b)
    return out


def test_frustum_cull_matches_the_oracle_node_cull():
    """crx_frustum_cull (the range hierarchy the server culls nodes with)
    keeps exactly the boxes the oracle's SNTriangles::update test keeps (the
    oracle rebuilds each box from its transformed corners, so the boxes here
    are proper ones: min <= max)."""
    L = orc.lib(crx)
    boxes = [b for b in _cull_boxes(1) if b.min.x <= b.max.x and b.min.y <= b.max.y and b.min.z <= b.max.z]
    n = len(boxes)
    node_meshes = (crx.CMesh * n)()
    node_world = (crx.Mat4 * n)()
    for i, b in enumerate(boxes):
        node_meshes[i].numTriangles = 1
        node_meshes[i].aabb = b
        node_meshes[i].isLoaded = 1
        node_meshes[i].address = i + 1
        node_world[i] = crx.identity()
    for v in _views_for_cull():
        got = crx.frustum_cull(v, boxes)
        uniq, nu, inst = (crx.CMesh * n)(), C.c_uint32(), (crx.CMesh * n)()
        xf, pre = (crx.Mat4 * n)(), (C.c_uint64 * n)()
        nv = L.crx_oracle_prepare_frame(node_meshes, node_world, n, C.byref(v), 1, 1, uniq, C.byref(nu), inst, xf, pre)
        want = sorted(int(inst[k].address) - 1 for k in range(nv))
        assert got == want
        assert got == sorted(got)


def test_frustum_cull_hierarchy_equals_each_box_test():
    """The hierarchy's shortcuts (dropping or keeping a whole range) give every
    box the answer of its own test, also for non-finite and inverted boxes:
    the cull of the set equals the boxes kept when each is culled alone."""
    boxes = _cull_boxes(2)
    nan, inf = float("nan"), float("inf")
    for i, (lo, hi) in enumerate([((nan, 0, 0), (1, 1, 1)), ((-inf, -1, -1), (1, 1, 1)), ((0, 0, 0), (inf, inf, inf)),
                                  ((-inf, -inf, -inf), (inf, inf, inf))]):
        b = crx.Box3()
        b.min.x, b.min.y, b.min.z = lo
        b.max.x, b.max.y, b.max.z = hi
        boxes.insert(100 + 37 * i, b)
    for v in _views_for_cull():
        alone = [i for i, b in enumerate(boxes) if crx.frustum_cull(v, [b]) == [0]]
        assert crx.frustum_cull(v, boxes) == alone


def test_frustum_cull_non_finite_view_equals_each_box_test():
    """A view matrix holding an infinity or a NaN gives non-finite planes;
    the hierarchy then tests every box on its own, so the cull of the set
    still equals the boxes kept when each is culled alone."""
    boxes = _cull_boxes(3)
    for k, bad in enumerate([float("inf"), float("-inf"), float("nan")]):
        v = _views_for_cull()[k]
        v.view.m[4 * k + 1] = bad
        alone = [i for i, b in enumerate(boxes) if crx.frustum_cull(v, [b]) == [0]]
        assert crx.frustum_cull(v, boxes) == alone
