"""ctypes binding of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this module, and only as the checker / the timed CPU baseline.  Structures are
the ones of crx.py (include/crx.h); pointer members must be HOST pointers.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent


class OracleStats(C.Structure):
    _fields_ = [("num_triangles", C.c_uint64), ("num_nontrivial", C.c_uint64), ("num_huge", C.c_uint64),
                ("num_stage1_px", C.c_uint64), ("num_stage2_px", C.c_uint64), ("num_stage3_px", C.c_uint64)]


_lib = None


def lib(crx):
    """Load liboracle.so; `crx` is the crx module (for the shared structs)."""
    global _lib
    if _lib is not None:
        return _lib
    path = HERE / "liboracle.so"
    if not path.exists():
        raise ImportError(f"{path} is missing: build it with `python oracle/build.py`")
    L = C.CDLL(str(path))
    VP = C.c_void_p
    P = C.POINTER
    sigs = {
        "crx_oracle_clear": (None, [VP, C.c_uint64]),
        "crx_oracle_compute_mipmap": (C.c_int, [VP, C.c_int32, C.c_int32]),
        "crx_oracle_draw": (C.c_int, [P(crx.CMesh), C.c_uint32, P(crx.CMesh), C.c_uint32, P(crx.Mat4),
                                      P(crx.RenderTarget), C.c_int, P(crx.Tuning), C.c_int, P(OracleStats)]),
        "crx_oracle_stage1": (C.c_uint64, [P(crx.CMesh), C.c_uint32, P(crx.CMesh), C.c_uint32, P(crx.Mat4),
                                           P(crx.RenderTarget), C.c_int, P(crx.Tuning), C.c_int, VP, C.c_uint64]),
        "crx_oracle_stage2": (C.c_uint64, [P(crx.CMesh), P(crx.Mat4), P(crx.RenderTarget), P(crx.Tuning),
                                           VP, C.c_uint64, C.c_int, VP, C.c_uint64]),
        "crx_oracle_stage3": (None, [P(crx.CMesh), P(crx.Mat4), P(crx.RenderTarget), P(crx.Tuning),
                                     VP, C.c_uint64, C.c_int]),
        "crx_oracle_resolve": (C.c_int, [P(crx.CMesh), C.c_uint32, P(C.c_uint64), P(crx.RenderTarget),
                                         C.c_int, C.c_int]),
        "crx_oracle_prepare_frame": (C.c_uint32, [P(crx.CMesh), P(crx.Mat4), C.c_uint32, P(crx.View), C.c_int,
                                                  C.c_int, P(crx.CMesh), P(C.c_uint32), P(crx.CMesh),
                                                  P(crx.Mat4), P(C.c_uint64)]),
        "crx_oracle_pack_pixel": (C.c_uint64, [C.c_float, C.c_uint64]),
        "crx_oracle_unpack_pixel": (None, [C.c_uint64, P(C.c_float), P(C.c_uint64)]),
        "crx_oracle_transform": (None, [P(crx.Mat4), P(C.c_float), C.c_float, C.c_float, P(C.c_float),
                                        P(C.c_float)]),
        "crx_oracle_mat4_mul": (None, [P(crx.Mat4), P(crx.Mat4), P(crx.Mat4)]),
        "crx_oracle_mat4_inverse": (None, [P(crx.Mat4), P(crx.Mat4)]),
        "crx_oracle_mat4_vec4": (None, [P(crx.Mat4), P(C.c_float), P(C.c_float)]),
        "crx_oracle_stage2_bbox": (None, [P(C.c_float), P(C.c_float), P(C.c_float), C.c_float, C.c_float,
                                          C.c_float, C.c_float, P(C.c_float)]),
        "crx_oracle_intersect_mt": (C.c_float, [P(C.c_float), P(C.c_float), P(C.c_float), P(C.c_float)]),
        "crx_oracle_bitedit_read": (C.c_uint32, [VP, C.c_uint64, C.c_uint64]),
        "crx_oracle_bitedit_write": (None, [VP, C.c_uint64, C.c_uint64, C.c_uint32]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


class HostMesh:
    """Host-side arrays of one mesh plus a CMesh pointing at them."""

    def __init__(self, crx, positions, indices, colors=None, uvs=None, texture=None, texture_size=None,
                 jpeg=False, normals=None):
        """texture: the RGBA8 mip chain as one flat uint32 array (level 0 of
        texture_size = (width, height) first); jpeg marks it JPEG-coded."""
        self.positions = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
        self.indices = None if indices is None else np.ascontiguousarray(indices, dtype=np.uint32).reshape(-1, 3)
        self.colors = None if colors is None else np.ascontiguousarray(colors, dtype=np.uint32).reshape(-1)
        self.uvs = None if uvs is None else np.ascontiguousarray(uvs, dtype=np.float32).reshape(-1, 2)
        self.normals = None if normals is None else np.ascontiguousarray(normals, dtype=np.float32).reshape(-1, 3)
        self.texture = None if texture is None else np.ascontiguousarray(texture, dtype=np.uint32).reshape(-1)
        self.texture_size = (0, 0) if texture_size is None else (int(texture_size[0]), int(texture_size[1]))
        self.jpeg = bool(jpeg)
        # stands in for CMesh::texture.huffmanTables, which is only tested for NULL
        self._huffman = np.zeros(1, dtype=np.uint32)
        self.num_triangles = len(self.indices) if self.indices is not None else len(self.positions) // 3

    def set_texture(self, m, uvs_ptr, texture_ptr):
        """Fill CMesh m's uvs and texture from device or host addresses."""
        m.uvs = uvs_ptr
        m.texture.width, m.texture.height = self.texture_size
        m.texture.data = texture_ptr
        m.texture.huffmanTables = self._huffman.ctypes.data if self.jpeg else None

    def cmesh(self, crx, cumm=0, offset=0, count=1, world=None, mesh_id=0):
        m = crx.CMesh()
        m.numTriangles = self.num_triangles
        m.indices = self.indices.ctypes.data if self.indices is not None else None
        m.positions = self.positions.ctypes.data
        m.colors = self.colors.ctypes.data if self.colors is not None else None
        m.cummulativeTriangleCount = cumm
        m.instances.offset = offset
        m.instances.count = count
        m.world = world if world is not None else crx.identity()
        m.id = mesh_id
        m.numVertices = len(self.positions)
        m.isLoaded = 1
        self.set_texture(m, self.uvs.ctypes.data if self.uvs is not None else None,
                         self.texture.ctypes.data if self.texture is not None else None)
        m.normals = self.normals.ctypes.data if self.normals is not None else None
        return m


def draw(crx, meshes, instances, transforms, width, height, view, proj, instanced, tuning=None,
         threads=0, fb=None):
    """Run the three stages on the CPU; returns (framebuffer u64 array, stats)."""
    L = lib(crx)
    if fb is None:
        fb = np.empty(width * height, dtype=np.uint64)
        L.crx_oracle_clear(fb.ctypes.data, fb.size)
    t = crx.RenderTarget()
    t.framebuffer = fb.ctypes.data
    t.width, t.height = width, height
    t.view, t.proj = view, proj
    tu = tuning if tuning is not None else crx.default_tuning()
    m_arr = crx._as_array(crx.CMesh, meshes)
    i_arr = crx._as_array(crx.CMesh, instances)
    x_arr = crx._as_array(crx.Mat4, transforms)
    st = OracleStats()
    rc = L.crx_oracle_draw(m_arr, len(m_arr) if m_arr is not None else 0, i_arr, len(i_arr), x_arr,
                           C.byref(t), int(bool(instanced)), C.byref(tu), threads, C.byref(st))
    if rc != 0:
        raise RuntimeError(f"oracle draw failed ({rc})")
    return fb, st


def resolve(crx, instances, prefix, fb, width, height, view, proj, attribute, threads=0, cb=None):
    L = lib(crx)
    if cb is None:
        cb = np.empty(width * height, dtype=np.uint64)
        L.crx_oracle_clear(cb.ctypes.data, cb.size)
    t = crx.RenderTarget()
    t.framebuffer = fb.ctypes.data
    t.colorbuffer = cb.ctypes.data
    t.width, t.height = width, height
    t.view, t.proj = view, proj
    i_arr = crx._as_array(crx.CMesh, instances)
    pre = np.ascontiguousarray(np.asarray(prefix, dtype=np.uint64))
    rc = L.crx_oracle_resolve(i_arr, len(i_arr) if i_arr is not None else 0,
                              pre.ctypes.data_as(C.POINTER(C.c_uint64)), C.byref(t), int(attribute), threads)
    if rc != 0:
        raise RuntimeError(f"oracle resolve failed ({rc})")
    return cb
