"""IX_PU16 compression (CPU): the compressor against hand-computed bit
layouts and the loader's formulas, and the oracle's compressed path against
its uncompressed path on the decoded positions (the decode is the only
difference, so the buffers must match bit for bit)."""
import ctypes as C
import math

import numpy as np

import crx
import oracle as orc
import workloads as wl
from crxtest import CompressedMesh, SceneSpec, oracle_render

L = orc.lib(crx)


def _read(words, off, bits):
    return L.crx_oracle_bitedit_read(words.ctypes.data_as(C.POINTER(C.c_uint32)), off, bits)


def test_bit_layout_known_answer():
    # indices 5..12: index_min 5, range 7 -> 3 bits; values 0,1,2,7,3,4 packed LSB first
    pos = np.arange(13 * 3, dtype=np.float32).reshape(13, 3)
    idx = np.array([[5, 6, 7], [12, 8, 9]], dtype=np.uint32)
    q, packed, meta = crx.compress_mesh(pos, idx)
    assert (meta.index_min, meta.index_max, meta.bitsPerIndex, meta.compressed) == (5, 12, 3, 1)
    want = 0
    for k, v in enumerate([0, 1, 2, 7, 3, 4]):
        want |= v << (3 * k)
    assert int(packed[0]) == want and len(packed) == 1      # 18 bits round up to one word


def test_straddling_indices_and_widths():
    rng = np.random.default_rng(3)
    for span in (1, 2, 3, 1000, (1 << 17) + 5, (1 << 20) - 1):
        base = 77
        nv = base + span
        idx = rng.integers(base, nv, size=(333, 3)).astype(np.uint32)
        idx[0, 0], idx[0, 1] = base, nv - 1
        pos = rng.normal(size=(nv, 3)).astype(np.float32)
        q, packed, meta = crx.compress_mesh(pos, idx)
        bits = meta.bitsPerIndex
        assert bits == math.ceil(math.log2(span))            # range + 1 = span values
        flat = idx.reshape(-1)
        got = np.array([_read(packed, bits * i, bits) + meta.index_min for i in range(len(flat))])
        assert np.array_equal(got, flat)
        assert len(packed) == max(1, ((len(flat) * bits + 7) // 8 + 3) // 4)


def test_quantisation_formula():
    rng = np.random.default_rng(4)
    pos = (rng.normal(size=(500, 3)) * [10, 0.5, 300]).astype(np.float32)
    idx = rng.integers(0, 500, size=(100, 3)).astype(np.uint32)
    q, _, meta = crx.compress_mesh(pos, idx)
    lo, hi = pos.min(axis=0), pos.max(axis=0)
    size = (hi - lo).astype(np.float32)
    want = np.clip(65536.0 * (pos - lo).astype(np.float32).astype(np.float64) / size.astype(np.float64),
                   0.0, 65535.0).astype(np.uint16)
    assert np.array_equal(q, want)
    assert (meta.aabb.min.x, meta.aabb.max.z) == (float(lo[0]), float(hi[2]))
    assert meta.compressionFactor.y == float(np.float32(size[1]) / np.float32(65536.0))
    # decoded positions stay within one quantisation step of the originals
    # (plus the float32 rounding of aabb.max - aabb.min and of p - aabb.min)
    dec = q.astype(np.float64) * (size / 65536.0) + lo
    slack = 4 * np.spacing(np.abs(pos).max(axis=0))
    assert np.all(np.abs(dec - pos) <= size / 65536.0 + slack)


def test_flat_axis_quantises_to_zero():
    pos = np.array([[0, 1, 0], [1, 1, 0], [0, 1, 1]], dtype=np.float32)
    q, _, meta = crx.compress_mesh(pos, np.array([[0, 1, 2]], dtype=np.uint32))
    assert np.all(q[:, 1] == 0) and meta.compressionFactor.y == 0.0


def test_oracle_compressed_equals_decoded_uncompressed():
    """Stages 1-3 and resolve: the compressed mesh renders exactly like an
    uncompressed mesh holding the decoded float positions."""
    pos, idx, col = wl.heightfield_mesh(97, amplitude=6.0, scale=24.0)
    cm = CompressedMesh(pos, idx, col)
    plain = orc.HostMesh(crx, cm.decoded_positions(), idx, col)
    for view, W, H in (
        (crx.look_at((48.0, 30.0, -20.0), (48.0, 0.0, 48.0)), 160, 120),     # small triangles (stage 1)
        (crx.look_at((48.0, 3.0, 30.0), (48.0, 1.0, 50.0)), 200, 150),       # close-up: stages 2 and 3
    ):
        proj = wl.projection(W, H)
        a, ca, sa = oracle_render(SceneSpec().add(cm, [crx.identity()]), W, H, view, proj,
                                  attribute=crx.DISPLAY_VERTEX_COLORS)
        b, cb, sb = oracle_render(SceneSpec().add(plain, [crx.identity()]), W, H, view, proj,
                                  attribute=crx.DISPLAY_VERTEX_COLORS)
        assert (sa.num_nontrivial, sa.num_huge) == (sb.num_nontrivial, sb.num_huge)
        assert np.array_equal(a, b) and np.array_equal(ca, cb)
        assert (a != crx.CLEAR_PIXEL).sum() > 1000
