"""Python binding of the gfx950 visibility-buffer rasterizer (ctypes over include/crx.h).

Mirrors the host interface of CuRast's triangle path:

* ``CMesh`` / ``RenderTarget`` / ``HugeTriangle`` are the reference's POD
  structs (src/kernels/HostDeviceInterface.h:131-143, :196-234, :254-259),
  byte-compatible with the C-ABI.
* ``Context.clear_framebuffer``, ``Context.draw_triangles_visbuffer`` and
  ``Context.resolve_visbuffer`` replace the launches of kernel_clearFramebuffer,
  drawTrianglesVisbuffer and kernel_resolve_visbuffer_to_colorbuffer2D
  (src/CuRast_render.h:211-264, :491-512, :580-592).
* ``Scene`` / ``Mesh`` / ``create_projection_matrix`` mirror Scene, SNTriangles,
  Mesh and Camera::createProjectionMatrix (src/scene/*.h, src/VKRenderer.h:141).

The native library must be built (``python curast-ooc_amd/build.py``); there is
no fallback path: importing this module without it raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

_PKG = Path(__file__).resolve().parent
_LIBDIR = _PKG / "lib"

# ----------------------------------------------------------------------------
# POD structs (must match include/crx.h byte for byte)
# ----------------------------------------------------------------------------


class Vec2(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float)]


class Vec3(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]


class Mat4(C.Structure):
    """Column-major float 4x4 (glm::mat4): element (col c, row r) is m[4*c + r]."""
    _fields_ = [("m", C.c_float * 16)]

    @classmethod
    def from_numpy(cls, a) -> "Mat4":
        a = np.asarray(a, dtype=np.float32)
        if a.shape == (4, 4):          # row-major math matrix -> column-major storage
            a = a.T.reshape(16)
        out = cls()
        C.memmove(out.m, np.ascontiguousarray(a.reshape(16), dtype=np.float32).ctypes.data, 64)
        return out

    def to_numpy(self) -> np.ndarray:
        """Row-major 4x4 (M[r, c])."""
        return np.frombuffer(bytes(self.m), dtype=np.float32).reshape(4, 4).T.copy()


class Box3(C.Structure):
    _fields_ = [("min", Vec3), ("max", Vec3)]


class Texture(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("data", C.c_void_p),
                ("huffmanTables", C.c_void_p), ("quanttables", C.c_void_p),
                ("mcuPositions", C.c_void_p), ("handle", C.c_uint32)]


class _Instances(C.Structure):
    _fields_ = [("offset", C.c_int32), ("count", C.c_int32)]


class _Impostor(C.Structure):
    _fields_ = [("positions", C.c_void_p), ("colors", C.c_void_p), ("numPoints", C.c_uint32)]


class CMesh(C.Structure):
    """CMesh (HostDeviceInterface.h:196-234), 280 bytes."""
    _fields_ = [
        ("numTriangles", C.c_uint32),
        ("indices", C.c_void_p),
        ("positions", C.c_void_p),
        ("cummulativeTriangleCount", C.c_uint64),
        ("aabb", Box3),
        ("compressionFactor", Vec3),
        ("index_min", C.c_uint32),
        ("bitsPerIndex", C.c_uint32),
        ("instances", _Instances),
        ("world", Mat4),
        ("id", C.c_int32),
        ("uvs", C.c_void_p),
        ("normals", C.c_void_p),
        ("colors", C.c_void_p),
        ("firstTriangle", C.c_uint32),
        ("numVertices", C.c_uint32),
        ("index_max", C.c_uint32),
        ("address", C.c_uint64),
        ("texture", Texture),
        ("isLoaded", C.c_uint8),
        ("flipTriangles", C.c_uint8),
        ("compressed", C.c_uint8),
        ("impostor", _Impostor),
    ]


class RenderTarget(C.Structure):
    """RenderTarget (HostDeviceInterface.h:131-143), 240 bytes."""
    _fields_ = [
        ("framebuffer", C.c_void_p),
        ("colorbuffer", C.c_void_p),
        ("width", C.c_int32),
        ("height", C.c_int32),
        ("view", Mat4),
        ("viewI", Mat4),
        ("proj", Mat4),
        ("cameraPos", Vec3),
        ("f", C.c_float),
        ("aspect", C.c_float),
        ("debug", C.c_uint8),
    ]


class HugeTriangle(C.Structure):
    _fields_ = [("meshIndex", C.c_int32), ("triangleIndex", C.c_int32),
                ("tile_x", C.c_int32), ("tile_y", C.c_int32)]


class RasterArgs(C.Structure):
    """RasterArgs (HostDeviceInterface.h:309-323) plus the two queue
    capacities; every pointer is a device address."""
    _fields_ = [
        ("meshes", C.c_void_p),
        ("numMeshes", C.c_uint32),
        ("instances", C.c_void_p),
        ("numInstances", C.c_uint32),
        ("transforms", C.c_void_p),
        ("numProcessedBatches", C.c_void_p),
        ("numProcessedBatches_nontrivial", C.c_void_p),
        ("hugeTriangles", C.c_void_p),
        ("hugeTrianglesCounter", C.c_void_p),
        ("numProcessedHugeTriangles", C.c_void_p),
        ("nontrivialTrianglesCounter", C.c_void_p),
        ("nontrivialTrianglesList", C.c_void_p),
        ("target", RenderTarget),
        ("nontrivialCapacity", C.c_uint32),
        ("hugeCapacity", C.c_uint32),
    ]


RASTER_INSTANCED = 1
RASTER_COMPRESSED = 2


class Tuning(C.Structure):
    _fields_ = [("small_threshold", C.c_uint32), ("tile_size", C.c_uint32),
                ("max_nontrivial", C.c_uint32), ("max_huge", C.c_uint32),
                ("stage1_block", C.c_uint32), ("stage1_waves_per_cu", C.c_uint32),
                ("stage2_waves_per_cu", C.c_uint32), ("stage3_waves_per_cu", C.c_uint32),
                ("stage1_cull_order", C.c_uint32), ("stage1_order", C.c_uint32),
                ("stage1_fast_tiny", C.c_uint32), ("stage1_tris_per_lane", C.c_uint32),
                ("stage1_layout", C.c_uint32), ("stage1_strip_rows", C.c_uint32),
                ("stage1_strip_batches", C.c_uint32)]


class DrawStats(C.Structure):
    _fields_ = [("num_triangles", C.c_uint64), ("num_nontrivial", C.c_uint32),
                ("num_huge", C.c_uint32), ("redraws", C.c_uint32), ("stage1_kernel", C.c_uint32)]


class AutotuneResult(C.Structure):
    _fields_ = [("tuning", Tuning), ("start_ms", C.c_float), ("best_ms", C.c_float),
                ("trials", C.c_uint32), ("reserved", C.c_uint32)]


TUNE_SCHEDULE, TUNE_ROUTING, TUNE_QUEUES = 1, 2, 4


class ServerConfig(C.Structure):
    _fields_ = [("node_cache_bytes", C.c_uint64), ("frame_bytes", C.c_uint64), ("max_batch_views", C.c_uint32),
                ("lanes", C.c_uint32), ("resolve", C.c_int32), ("prefetch", C.c_uint32),
                ("shared_stream", C.c_uint32), ("lane_backlog", C.c_uint32), ("merge_backlog", C.c_uint32),
                ("reserved", C.c_uint32)]


class RequestResult(C.Structure):
    _fields_ = [("framebuffer", C.c_void_p), ("colorbuffer", C.c_void_p), ("submit_time", C.c_double),
                ("start_time", C.c_double), ("done_time", C.c_double), ("batch_views", C.c_uint32),
                ("uploads", C.c_uint32), ("status", C.c_int), ("reserved", C.c_uint32)]


class ServerStats(C.Structure):
    _fields_ = [("requests", C.c_uint64), ("batches", C.c_uint64), ("uploads", C.c_uint64),
                ("upload_bytes", C.c_uint64), ("evictions", C.c_uint64), ("resident_nodes", C.c_uint32),
                ("cache_slots", C.c_uint32), ("max_batch", C.c_uint32), ("prefetches", C.c_uint32)]


class Timings(C.Structure):
    _fields_ = [("clear_ms", C.c_float), ("stage1_ms", C.c_float), ("stage2_ms", C.c_float),
                ("stage3_ms", C.c_float), ("resolve_ms", C.c_float), ("pipeline_ms", C.c_float)]


class View(C.Structure):
    _fields_ = [("view", Mat4), ("proj", Mat4), ("width", C.c_int32), ("height", C.c_int32)]


class RenderSettings(C.Structure):
    _fields_ = [("instanced", C.c_int32), ("frustum_culling", C.c_int32),
                ("resolve", C.c_int32), ("attribute", C.c_int32)]


# DisplayAttribute (HostDeviceInterface.h:277-286)
DISPLAY_NONE, DISPLAY_TEXTURE, DISPLAY_UV, DISPLAY_NORMAL = 0, 1, 2, 3
DISPLAY_VERTEX_COLORS, DISPLAY_TRIANGLE_ID, DISPLAY_MESH_ID, DISPLAY_STAGE = 4, 5, 6, 7

CLEAR_PIXEL = 0xFFFFFFF000000000

STATUS = {0: "CRX_OK", 1: "CRX_INVALID_ARG", 2: "CRX_OUT_OF_MEMORY", 3: "CRX_QUEUE_OVERFLOW",
          4: "CRX_HIP_ERROR", 5: "CRX_NO_DEVICE", 6: "CRX_UNSUPPORTED", 7: "CRX_REJECTED"}

assert C.sizeof(CMesh) == 280, C.sizeof(CMesh)
assert C.sizeof(RenderTarget) == 240, C.sizeof(RenderTarget)


class CrxError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS.get(status, status)}: {message}")
        self.status = status


# ----------------------------------------------------------------------------
# Library loading (fails loudly: there is no CPU fallback for the product path)
# ----------------------------------------------------------------------------

_lib = None
_gen = None

_VP = C.c_void_p
_SIGS = {
    "crx_abi_version": (C.c_int, []),
    "crx_status_string": (C.c_char_p, [C.c_int]),
    "crx_last_error": (C.c_char_p, []),
    "crx_tuning_default": (None, [C.POINTER(Tuning)]),
    "crx_context_create": (C.c_int, [C.c_int, C.POINTER(Tuning), C.POINTER(_VP)]),
    "crx_context_destroy": (C.c_int, [_VP]),
    "crx_context_set_tuning": (C.c_int, [_VP, C.POINTER(Tuning)]),
    "crx_context_get_tuning": (C.c_int, [_VP, C.POINTER(Tuning)]),
    "crx_context_num_cus": (C.c_int, [_VP, C.POINTER(C.c_int)]),
    "crx_context_enable_timing": (C.c_int, [_VP, C.c_int]),
    "crx_context_timings": (C.c_int, [_VP, C.POINTER(Timings)]),
    "crx_device_alloc": (C.c_int, [_VP, C.c_size_t, C.POINTER(_VP)]),
    "crx_device_free": (C.c_int, [_VP, _VP]),
    "crx_memcpy_htod": (C.c_int, [_VP, _VP, _VP, C.c_size_t, _VP]),
    "crx_memcpy_dtoh": (C.c_int, [_VP, _VP, _VP, C.c_size_t, _VP]),
    "crx_stream_synchronize": (C.c_int, [_VP, _VP]),
    "crx_copy_buffer": (C.c_int, [_VP, _VP, _VP, C.c_size_t, _VP]),
    "crx_probe_atomic_min": (C.c_int, [_VP, _VP, C.c_uint32, C.c_uint32, C.c_int, C.POINTER(C.c_uint64), _VP]),
    "crx_probe_read": (C.c_int, [_VP, _VP, C.c_size_t, _VP, _VP]),
    "crx_probe_valu": (C.c_int, [_VP, C.c_uint32, C.c_int, _VP, C.POINTER(C.c_uint64), _VP]),
    "crx_texture_chain_texels": (C.c_int, [C.c_int32, C.c_int32, C.POINTER(C.c_uint64)]),
    "crx_texture_compute_mipmap": (C.c_int, [_VP, _VP, C.c_int32, C.c_int32, _VP]),
    "crx_check_division": (C.c_int, [_VP, C.c_uint64, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                     C.POINTER(C.c_uint64)]),
    "crx_clear_framebuffer": (C.c_int, [_VP, C.c_uint32, _VP]),
    "crx_raster_stage1": (C.c_int, [_VP, C.POINTER(RasterArgs), C.c_uint32, _VP]),
    "crx_raster_stage2": (C.c_int, [_VP, C.POINTER(RasterArgs), _VP]),
    "crx_raster_stage3": (C.c_int, [_VP, C.POINTER(RasterArgs), _VP]),
    "crx_raster_visbuffer_noalloc": (C.c_int, [_VP, C.POINTER(RasterArgs), C.c_uint32, _VP]),
    "crx_draw_triangles_visbuffer": (C.c_int, [_VP, C.POINTER(CMesh), C.c_uint32, C.POINTER(CMesh),
                                               C.c_uint32, C.POINTER(Mat4), C.POINTER(RenderTarget),
                                               C.c_int, _VP, C.POINTER(DrawStats)]),
    "crx_resolve_visbuffer": (C.c_int, [_VP, C.POINTER(CMesh), C.c_uint32, C.POINTER(C.c_uint64),
                                        C.POINTER(RenderTarget), C.c_int, _VP]),
    "crx_draw_triangles_visbuffer_views": (C.c_int, [_VP, C.POINTER(CMesh), C.c_uint32, C.POINTER(CMesh),
                                                     C.c_uint32, C.POINTER(Mat4), C.POINTER(RenderTarget),
                                                     C.c_uint32, C.c_int, _VP, C.POINTER(DrawStats)]),
    "crx_resolve_visbuffer_views": (C.c_int, [_VP, C.POINTER(CMesh), C.c_uint32, C.POINTER(C.c_uint64),
                                              C.POINTER(RenderTarget), C.c_uint32, C.c_int, _VP]),
    "crx_clear_views": (C.c_int, [_VP, C.POINTER(RenderTarget), C.c_uint32, C.c_int, _VP]),
    "crx_render_views": (C.c_int, [_VP, _VP, C.POINTER(View), C.c_uint32, C.POINTER(RenderSettings),
                                   C.POINTER(_VP), C.POINTER(_VP), _VP, C.POINTER(DrawStats)]),
    "crx_frame_create": (C.c_int, [_VP, _VP, C.POINTER(View), C.c_uint32, C.POINTER(RenderSettings),
                                   C.POINTER(_VP), C.POINTER(_VP), C.POINTER(_VP)]),
    "crx_frame_set_views": (C.c_int, [_VP, C.POINTER(View)]),
    "crx_frame_launch": (C.c_int, [_VP, _VP]),
    "crx_frame_wait": (C.c_int, [_VP, C.POINTER(DrawStats)]),
    "crx_frame_destroy": (C.c_int, [_VP]),
    "crx_server_config_default": (None, [C.POINTER(ServerConfig)]),
    "crx_host_alloc": (C.c_int, [_VP, C.c_size_t, C.POINTER(_VP)]),
    "crx_host_free": (C.c_int, [_VP, _VP]),
    "crx_server_create": (C.c_int, [_VP, C.POINTER(ServerConfig), C.POINTER(_VP)]),
    "crx_server_destroy": (C.c_int, [_VP]),
    "crx_server_add_node": (C.c_int, [_VP, _VP, C.c_uint32, _VP, C.c_uint32, _VP, C.POINTER(Box3),
                                      C.POINTER(Mat4), C.c_int, C.POINTER(C.c_uint32)]),
    "crx_server_start": (C.c_int, [_VP]),
    "crx_server_submit": (C.c_int, [_VP, C.POINTER(View), C.POINTER(C.c_uint64)]),
    "crx_server_wait": (C.c_int, [_VP, C.c_uint64, C.POINTER(RequestResult)]),
    "crx_server_release": (C.c_int, [_VP, C.c_uint64]),
    "crx_server_clock": (C.c_double, [_VP]),
    "crx_server_get_stats": (C.c_int, [_VP, C.POINTER(ServerStats)]),
    "crx_server_view_nodes": (C.c_int, [_VP, C.POINTER(View), C.POINTER(C.c_uint32), C.c_uint32,
                                        C.POINTER(C.c_uint32)]),
    "crx_autotune": (C.c_int, [_VP, _VP, C.POINTER(View), C.c_uint32, C.POINTER(_VP), C.c_uint32, C.c_uint32,
                               C.POINTER(AutotuneResult)]),
    "crx_camera_projection": (None, [C.c_float, C.c_float, C.c_float, C.POINTER(Mat4)]),
    "crx_camera_look_at": (None, [C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float),
                                  C.POINTER(Mat4)]),
    "crx_mat4_inverse": (None, [C.POINTER(Mat4), C.POINTER(Mat4)]),
    "crx_mat4_mul": (None, [C.POINTER(Mat4), C.POINTER(Mat4), C.POINTER(Mat4)]),
    "crx_mesh_create": (C.c_int, [_VP, _VP, C.c_uint32, _VP, C.c_uint32, _VP, C.POINTER(_VP)]),
    "crx_compress_mesh": (C.c_int, [_VP, C.c_uint32, _VP, C.c_uint32, _VP, _VP, C.POINTER(C.c_uint64),
                                    C.POINTER(CMesh)]),
    "crx_mesh_create_compressed": (C.c_int, [_VP, _VP, C.c_uint32, _VP, C.c_uint32, _VP, C.POINTER(_VP)]),
    "crx_mesh_create_device": (C.c_int, [_VP, _VP, C.c_uint32, _VP, C.c_uint32, _VP, C.POINTER(Box3),
                                         C.POINTER(_VP)]),
    "crx_mesh_destroy": (C.c_int, [_VP]),
    "crx_mesh_info": (C.c_int, [_VP, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(Box3)]),
    "crx_mesh_row_triangles": (C.c_int, [_VP, C.POINTER(C.c_uint32)]),
    "crx_mesh_buffers": (C.c_int, [_VP, C.POINTER(_VP), C.POINTER(_VP), C.POINTER(_VP)]),
    "crx_scene_create": (C.c_int, [_VP, C.POINTER(_VP)]),
    "crx_scene_destroy": (C.c_int, [_VP]),
    "crx_scene_add_node": (C.c_int, [_VP, _VP, C.POINTER(Mat4), C.POINTER(C.c_uint32)]),
    "crx_scene_set_transform": (C.c_int, [_VP, C.c_uint32, C.POINTER(Mat4)]),
    "crx_scene_set_texture": (C.c_int, [_VP, C.c_uint32, C.POINTER(Texture)]),
    "crx_frustum_cull": (C.c_int, [C.POINTER(View), C.POINTER(Box3), C.c_uint32, C.POINTER(C.c_uint32), C.c_uint32,
                                   C.POINTER(C.c_uint32)]),
    "crx_mesh_set_attributes": (C.c_int, [_VP, _VP, _VP]),
    "crx_scene_num_nodes": (C.c_int, [_VP, C.POINTER(C.c_uint32)]),
    "crx_render_settings_default": (None, [C.POINTER(RenderSettings)]),
    "crx_render": (C.c_int, [_VP, _VP, C.POINTER(View), C.POINTER(RenderSettings), _VP, _VP, _VP,
                             C.POINTER(DrawStats)]),
}

_GEN_SIGS = {
    "crx_gen_splitmix64": (C.c_uint64, [C.c_uint64, C.c_uint64]),
    "crx_gen_hash64": (C.c_uint64, [_VP, C.c_size_t]),
    "crx_gen_sphere": (None, [C.c_uint32, C.c_uint32, C.c_float, C.c_float, C.c_uint64, _VP, _VP, _VP]),
    "crx_gen_heightfield": (None, [C.c_uint32, C.c_float, C.c_float, C.c_float, C.c_uint32, C.c_uint64,
                                   _VP, _VP, _VP, C.c_int]),
    "crx_gen_heightfield_tile": (None, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_float,
                                        C.c_float, C.c_float, C.c_uint32, C.c_uint64, _VP, _VP, _VP, C.c_int]),
    "crx_gen_heightfield_indices": (C.c_uint64, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                                 C.c_uint32, _VP]),
    "crx_gen_instances": (None, [C.c_uint32, C.c_float, C.c_float, C.c_float, C.c_uint64, C.POINTER(Mat4)]),
}


def library_path() -> Path:
    return Path(os.environ.get("CRX_LIBRARY", str(_LIBDIR / "libcrx.so")))


def lib():
    """The loaded libcrx.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        path = library_path()
        if not path.exists():
            raise ImportError(f"{path} is missing: build it with `python curast-ooc_amd/build.py`")
        handle = C.CDLL(str(path))
        for name, (res, args) in _SIGS.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def gen():
    """The loaded libcrx_gen.so (synthetic workload generators)."""
    global _gen
    if _gen is None:
        path = _LIBDIR / "libcrx_gen.so"
        if not path.exists():
            raise ImportError(f"{path} is missing: build it with `python curast-ooc_amd/build.py`")
        handle = C.CDLL(str(path))
        for name, (res, args) in _GEN_SIGS.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _gen = handle
    return _gen


def _check(status: int) -> None:
    if status != 0:
        raise CrxError(status, lib().crx_last_error().decode())


def _ptr(a: np.ndarray | None):
    return None if a is None else C.c_void_p(a.ctypes.data)


# ----------------------------------------------------------------------------
# Context, device memory
# ----------------------------------------------------------------------------


def default_tuning() -> Tuning:
    t = Tuning()
    lib().crx_tuning_default(C.byref(t))
    return t


class DeviceBuffer:
    """Device allocation owned by a context (freed on close / garbage collection)."""

    def __init__(self, ctx: "Context", nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        _check(lib().crx_device_alloc(ctx.handle, self.nbytes, C.byref(p)))
        self.ptr = p.value

    def upload(self, array: np.ndarray, stream=None) -> "DeviceBuffer":
        a = np.ascontiguousarray(array)
        if a.nbytes > self.nbytes:
            raise ValueError("upload larger than the buffer")
        _check(lib().crx_memcpy_htod(self.ctx.handle, C.c_void_p(self.ptr), _ptr(a), a.nbytes, stream))
        return self

    def download(self, dtype, count: int | None = None, stream=None) -> np.ndarray:
        dtype = np.dtype(dtype)
        n = self.nbytes // dtype.itemsize if count is None else int(count)
        out = np.empty(n, dtype=dtype)
        _check(lib().crx_memcpy_dtoh(self.ctx.handle, _ptr(out), C.c_void_p(self.ptr), out.nbytes, stream))
        return out

    def free(self) -> None:
        if self.ptr and self.ctx.handle:
            lib().crx_device_free(self.ctx.handle, C.c_void_p(self.ptr))
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    """One rasterizer context on one device (queues, counters, staging)."""

    def __init__(self, device: int = 0, tuning: Tuning | None = None):
        h = C.c_void_p()
        _check(lib().crx_context_create(device, C.byref(tuning) if tuning is not None else None, C.byref(h)))
        self.handle = h.value
        self._buffers: list[DeviceBuffer] = []

    def close(self) -> None:
        if self.handle:
            for b in self._buffers:
                b.free()
            self._buffers.clear()
            lib().crx_context_destroy(C.c_void_p(self.handle))
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def tuning(self) -> Tuning:
        t = Tuning()
        _check(lib().crx_context_get_tuning(self.handle, C.byref(t)))
        return t

    @tuning.setter
    def tuning(self, t: Tuning) -> None:
        _check(lib().crx_context_set_tuning(self.handle, C.byref(t)))

    @property
    def num_cus(self) -> int:
        n = C.c_int()
        _check(lib().crx_context_num_cus(self.handle, C.byref(n)))
        return n.value

    def enable_timing(self, on: bool = True) -> None:
        _check(lib().crx_context_enable_timing(self.handle, int(on)))

    def timings(self) -> Timings:
        t = Timings()
        _check(lib().crx_context_timings(self.handle, C.byref(t)))
        return t

    def alloc(self, nbytes: int) -> DeviceBuffer:
        b = DeviceBuffer(self, nbytes)
        self._buffers.append(b)
        return b

    def to_device(self, array: np.ndarray) -> DeviceBuffer:
        a = np.ascontiguousarray(array)
        return self.alloc(max(a.nbytes, 1)).upload(a)

    def synchronize(self, stream=None) -> None:
        _check(lib().crx_stream_synchronize(self.handle, stream))

    def copy(self, dst: DeviceBuffer, src: DeviceBuffer, nbytes: int, stream=None) -> None:
        _check(lib().crx_copy_buffer(self.handle, C.c_void_p(dst.ptr), C.c_void_p(src.ptr), nbytes, stream))

    def check_division(self, pairs: int) -> tuple[int, int, int]:
        """Mismatches of stage 1's division sequences against the IEEE
        division: (reciprocal over every float in range, quotient over
        `pairs` operand pairs, screen term over `pairs` small numerators)."""
        r, d, t = C.c_uint64(), C.c_uint64(), C.c_uint64()
        _check(lib().crx_check_division(self.handle, pairs, C.byref(r), C.byref(d), C.byref(t)))
        return r.value, d.value, t.value

    def probe_read(self, src: DeviceBuffer, nbytes: int, sink: DeviceBuffer, stream=None) -> None:
        """Enqueue the streaming-read bandwidth probe over nbytes of src."""
        _check(lib().crx_probe_read(self.handle, C.c_void_p(src.ptr), nbytes, C.c_void_p(sink.ptr), stream))

    def probe_atomic_min(self, buf: DeviceBuffer, log_entries: int, per_lane: int, scattered: bool,
                         stream=None) -> int:
        """Enqueue the 64-bit atomicMin probe; returns the number of atomics."""
        n = C.c_uint64()
        _check(lib().crx_probe_atomic_min(self.handle, C.c_void_p(buf.ptr), log_entries, per_lane,
                                          int(bool(scattered)), C.byref(n), stream))
        return n.value

    def probe_valu(self, iters: int, mode: int, sink: DeviceBuffer, stream=None) -> int:
        """Enqueue the VALU issue-rate probe (mode 0 fp32 fma, 1 int32
        add/xor, 2 fma on 32 lanes, 3 fma on 1 lane, 4 fma and int
        interleaved, 5 fma and int waves alternating); returns the number of
        waves launched."""
        n = C.c_uint64()
        _check(lib().crx_probe_valu(self.handle, iters, mode, C.c_void_p(sink.ptr), C.byref(n), stream))
        return n.value

    def compute_mipmap(self, data: DeviceBuffer | int, width: int, height: int, stream=None) -> None:
        """Levels 1 .. of the RGBA8 mip chain in `data` from its level 0
        (computeMipMap, textureTools.cu:138-178); data holds
        texture_chain_texels(width, height) texels."""
        ptr = data.ptr if isinstance(data, DeviceBuffer) else data
        _check(lib().crx_texture_compute_mipmap(self.handle, C.c_void_p(ptr), width, height, stream))

    # --- kernel-level entry points ------------------------------------------------

    def clear_framebuffer(self, fb: DeviceBuffer | int, num_pixels: int, stream=None) -> None:
        ptr = fb.ptr if isinstance(fb, DeviceBuffer) else fb
        _check(lib().crx_clear_framebuffer(C.c_void_p(ptr), num_pixels, stream))

    def raster_stage1(self, args: RasterArgs, flags: int, stream=None) -> None:
        """kernel_stage1_drawSmallTriangles_* over device CMesh arrays (crx_raster_stage1)."""
        _check(lib().crx_raster_stage1(self.handle, C.byref(args), int(flags), stream))

    def raster_stage2(self, args: RasterArgs, stream=None) -> None:
        _check(lib().crx_raster_stage2(self.handle, C.byref(args), stream))

    def raster_stage3(self, args: RasterArgs, stream=None) -> None:
        _check(lib().crx_raster_stage3(self.handle, C.byref(args), stream))

    def raster_visbuffer_noalloc(self, args: RasterArgs, flags: int, stream=None) -> None:
        """The three stages of `args` with the draw records in the library's
        static device scratch (crx_raster_visbuffer_noalloc): no allocation."""
        _check(lib().crx_raster_visbuffer_noalloc(self.handle, C.byref(args), int(flags), stream))

    def draw_triangles_visbuffer(self, meshes, instances, transforms, target: RenderTarget,
                                 instanced: bool, stream=None) -> DrawStats:
        """drawTrianglesVisbuffer (CuRast_render.h:211-264).

        meshes / instances: ctypes arrays (or lists) of CMesh holding device
        pointers; transforms: one world-to-view Mat4 per instance."""
        m_arr = _as_array(CMesh, meshes)
        i_arr = _as_array(CMesh, instances)
        t_arr = _as_array(Mat4, transforms)
        st = DrawStats()
        _check(lib().crx_draw_triangles_visbuffer(
            self.handle, m_arr, len(m_arr) if m_arr is not None else 0, i_arr,
            len(i_arr) if i_arr is not None else 0, t_arr, C.byref(target), int(bool(instanced)),
            stream, C.byref(st)))
        return st

    def resolve_visbuffer(self, instances, prefix, target: RenderTarget,
                          attribute: int = DISPLAY_VERTEX_COLORS, stream=None) -> None:
        i_arr = _as_array(CMesh, instances)
        pre = np.ascontiguousarray(np.asarray(prefix, dtype=np.uint64))
        n = len(i_arr) if i_arr is not None else 0
        _check(lib().crx_resolve_visbuffer(self.handle, i_arr, n,
                                           pre.ctypes.data_as(C.POINTER(C.c_uint64)) if n else None,
                                           C.byref(target), int(attribute), stream))

    # --- batched multi-view ---------------------------------------------------------

    def draw_triangles_visbuffer_views(self, meshes, instances, transforms, targets,
                                       instanced: bool, stream=None) -> DrawStats:
        """One pass over the geometry for len(targets) views.

        transforms: len(targets) * len(instances) world-to-view matrices,
        view-major (transforms[v * n + i] = view_v * world_i)."""
        m_arr = _as_array(CMesh, meshes)
        i_arr = _as_array(CMesh, instances)
        t_arr = _as_array(Mat4, transforms)
        g_arr = _as_array(RenderTarget, targets)
        st = DrawStats()
        _check(lib().crx_draw_triangles_visbuffer_views(
            self.handle, m_arr, len(m_arr) if m_arr is not None else 0, i_arr,
            len(i_arr) if i_arr is not None else 0, t_arr, g_arr, len(g_arr) if g_arr is not None else 0,
            int(bool(instanced)), stream, C.byref(st)))
        return st

    def resolve_visbuffer_views(self, instances, prefix, targets, attribute: int = DISPLAY_VERTEX_COLORS,
                                stream=None) -> None:
        i_arr = _as_array(CMesh, instances)
        g_arr = _as_array(RenderTarget, targets)
        pre = np.ascontiguousarray(np.asarray(prefix, dtype=np.uint64))
        n = len(i_arr) if i_arr is not None else 0
        _check(lib().crx_resolve_visbuffer_views(self.handle, i_arr, n,
                                                 pre.ctypes.data_as(C.POINTER(C.c_uint64)) if n else None,
                                                 g_arr, len(g_arr) if g_arr is not None else 0,
                                                 int(attribute), stream))

    def autotune(self, scene: "Scene", views, fbs, flags: int = TUNE_SCHEDULE, reps: int = 5) -> AutotuneResult:
        """Coordinate-descent search of the tuning on this scene and these
        views (crx_autotune); the context keeps the best tuning."""
        v_arr = _as_array(View, views)
        n = len(v_arr)
        fb_arr = (C.c_void_p * n)(*[b.ptr for b in fbs])
        res = AutotuneResult()
        _check(lib().crx_autotune(self.handle, scene.handle, v_arr, n, fb_arr, flags, reps, C.byref(res)))
        return res

    def clear_views(self, targets, colorbuffers: bool = False, stream=None) -> None:
        g_arr = _as_array(RenderTarget, targets)
        _check(lib().crx_clear_views(self.handle, g_arr, len(g_arr) if g_arr is not None else 0,
                                     int(bool(colorbuffers)), stream))


def _as_array(ctype, items):
    if items is None:
        return None
    if isinstance(items, C.Array):
        return items
    items = list(items)
    if not items:
        return None
    arr = (ctype * len(items))()
    for i, it in enumerate(items):
        arr[i] = it
    return arr


# ----------------------------------------------------------------------------
# Camera / matrices (float, GLM order)
# ----------------------------------------------------------------------------


def create_projection_matrix(near: float, fovy_rad: float, aspect: float) -> Mat4:
    """Camera::createProjectionMatrix (VKRenderer.h:141-148)."""
    m = Mat4()
    lib().crx_camera_projection(near, fovy_rad, aspect, C.byref(m))
    return m


def look_at(eye, center, up=(0.0, 1.0, 0.0)) -> Mat4:
    e = (C.c_float * 3)(*eye)
    c = (C.c_float * 3)(*center)
    u = (C.c_float * 3)(*up)
    m = Mat4()
    lib().crx_camera_look_at(e, c, u, C.byref(m))
    return m


def mat4_inverse(m: Mat4) -> Mat4:
    out = Mat4()
    lib().crx_mat4_inverse(C.byref(m), C.byref(out))
    return out


def mat4_mul(a: Mat4, b: Mat4) -> Mat4:
    out = Mat4()
    lib().crx_mat4_mul(C.byref(a), C.byref(b), C.byref(out))
    return out


def identity() -> Mat4:
    return Mat4.from_numpy(np.eye(4, dtype=np.float32))


def make_target(fb, cb, width: int, height: int, view: Mat4, proj: Mat4) -> RenderTarget:
    t = RenderTarget()
    t.framebuffer = fb.ptr if isinstance(fb, DeviceBuffer) else fb
    t.colorbuffer = (cb.ptr if isinstance(cb, DeviceBuffer) else cb) if cb is not None else None
    t.width = width
    t.height = height
    t.view = view
    t.viewI = mat4_inverse(view)
    t.proj = proj
    t.f = proj.m[5]
    t.aspect = float(np.float32(width) / np.float32(height))
    return t


# ----------------------------------------------------------------------------
# Scene API (Scene / SNTriangles / Mesh)
# ----------------------------------------------------------------------------


class Mesh:
    """Mesh (Mesh.h:4-25): positions (V,3) float32, indices (T,3) uint32 or None,
    colors (V,) uint32 RGBA8 or None; copied to device memory owned by the mesh."""

    def __init__(self, ctx: Context, positions: np.ndarray, indices: np.ndarray | None,
                 colors: np.ndarray | None = None, num_triangles: int | None = None, compressed: bool = False):
        pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
        idx = None if indices is None else np.ascontiguousarray(indices, dtype=np.uint32).reshape(-1)
        col = None if colors is None else np.ascontiguousarray(colors, dtype=np.uint32).reshape(-1)
        nt = (len(idx) // 3 if idx is not None else len(pos) // 3) if num_triangles is None else num_triangles
        h = C.c_void_p()
        create = lib().crx_mesh_create_compressed if compressed else lib().crx_mesh_create
        _check(create(ctx.handle, _ptr(pos), len(pos), _ptr(idx), nt, _ptr(col), C.byref(h)))
        self.ctx = ctx
        self.handle = h.value
        self.num_triangles = nt
        self.num_vertices = len(pos)

    def buffers(self):
        p, i, c = C.c_void_p(), C.c_void_p(), C.c_void_p()
        _check(lib().crx_mesh_buffers(self.handle, C.byref(p), C.byref(i), C.byref(c)))
        return p.value, i.value, c.value

    def set_attributes(self, uvs: np.ndarray | None = None, normals: np.ndarray | None = None) -> None:
        """Per-vertex uvs (V,2) and normals (V,3) float32 for the resolve
        (Mesh.h:17-19); None leaves an attribute as it is."""
        uv = None if uvs is None else np.ascontiguousarray(uvs, dtype=np.float32).reshape(-1, 2)
        nr = None if normals is None else np.ascontiguousarray(normals, dtype=np.float32).reshape(-1, 3)
        for a in (uv, nr):
            if a is not None and len(a) != self.num_vertices:
                raise ValueError("one attribute per vertex")
        _check(lib().crx_mesh_set_attributes(C.c_void_p(self.handle), _ptr(uv), _ptr(nr)))

    def aabb(self) -> Box3:
        b = Box3()
        _check(lib().crx_mesh_info(self.handle, None, None, C.byref(b)))
        return b

    def row_triangles(self) -> int:
        """Triangles per row found in the index buffer (0 = none)."""
        n = C.c_uint32()
        _check(lib().crx_mesh_row_triangles(self.handle, C.byref(n)))
        return n.value

    def close(self) -> None:
        if self.handle:
            lib().crx_mesh_destroy(C.c_void_p(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Scene:
    """Scene graph of SNTriangles nodes (flat: each node carries its global transform)."""

    def __init__(self, ctx: Context):
        h = C.c_void_p()
        _check(lib().crx_scene_create(ctx.handle, C.byref(h)))
        self.ctx = ctx
        self.handle = h.value
        self.nodes: list[tuple[Mesh, Mat4]] = []

    def add(self, mesh: Mesh, world: Mat4 | None = None) -> int:
        w = identity() if world is None else world
        nid = C.c_uint32()
        _check(lib().crx_scene_add_node(self.handle, mesh.handle, C.byref(w), C.byref(nid)))
        self.nodes.append((mesh, w))
        return nid.value

    def set_transform(self, node_id: int, world: Mat4) -> None:
        _check(lib().crx_scene_set_transform(self.handle, node_id, C.byref(world)))
        self.nodes[node_id] = (self.nodes[node_id][0], world)

    def set_texture(self, node_id: int, texture: Texture | None) -> None:
        """SNTriangles::texture of a node (a device mip chain); None clears it."""
        _check(lib().crx_scene_set_texture(self.handle, node_id, C.byref(texture) if texture is not None else None))

    def render(self, view: View, fb: DeviceBuffer, cb: DeviceBuffer | None,
               settings: RenderSettings | None = None, stream=None) -> DrawStats:
        """CuRast::draw for one view (CuRast_render.h:266-592)."""
        st = DrawStats()
        s = settings if settings is not None else render_settings()
        _check(lib().crx_render(self.ctx.handle, self.handle, C.byref(view), C.byref(s),
                                C.c_void_p(fb.ptr), C.c_void_p(cb.ptr) if cb is not None else None,
                                stream, C.byref(st)))
        return st

    def render_views(self, views, fbs, cbs=None, settings: RenderSettings | None = None,
                     stream=None) -> DrawStats:
        """All views in one batched pass; each view's buffers equal a
        single-view draw of the shared (union-culled) instance list."""
        st = DrawStats()
        s = settings if settings is not None else render_settings()
        v_arr = _as_array(View, views)
        n = len(v_arr)
        fb_arr = (C.c_void_p * n)(*[b.ptr for b in fbs])
        cb_arr = (C.c_void_p * n)(*[b.ptr for b in cbs]) if cbs is not None else None
        _check(lib().crx_render_views(self.ctx.handle, self.handle, v_arr, n, C.byref(s), fb_arr, cb_arr,
                                      stream, C.byref(st)))
        return st

    def close(self) -> None:
        if self.handle:
            lib().crx_scene_destroy(C.c_void_p(self.handle))
            self.handle = None


def frustum_cull(view: View, boxes) -> list[int]:
    """Indices of the world boxes the view's node cull keeps, ascending
    (SNTriangles::update's test through the range hierarchy)."""
    arr = _as_array(Box3, boxes)
    n = len(arr) if arr is not None else 0
    ids = (C.c_uint32 * max(n, 1))()
    cnt = C.c_uint32()
    _check(lib().crx_frustum_cull(C.byref(view), arr, n, ids, n, C.byref(cnt)))
    return list(ids[:cnt.value])


def texture_chain_texels(width: int, height: int) -> int:
    """Texels of a width x height texture's mip chain (levels 0 .. max(L-1, 7))."""
    n = C.c_uint64()
    _check(lib().crx_texture_chain_texels(width, height, C.byref(n)))
    return n.value


def compress_mesh(positions: np.ndarray, indices: np.ndarray):
    """IX_PU16 compression (crx_compress_mesh): returns (qpositions uint16 [V,3],
    packed index words uint32, CMesh with the compression fields set)."""
    pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
    idx = np.ascontiguousarray(indices, dtype=np.uint32).reshape(-1, 3)
    q = np.empty((len(pos), 3), dtype=np.uint16)
    packed = np.zeros(3 * len(idx) + 1, dtype=np.uint32)
    words = C.c_uint64()
    meta = CMesh()
    _check(lib().crx_compress_mesh(_ptr(pos), len(pos), _ptr(idx), len(idx), _ptr(q), _ptr(packed),
                                   C.byref(words), C.byref(meta)))
    return q, packed[:max(1, words.value)].copy(), meta


class PinnedBuffer:
    """Page-locked host memory (crx_host_alloc) the server streams node geometry from."""

    def __init__(self, ctx: "Context", nbytes: int):
        p = C.c_void_p()
        _check(lib().crx_host_alloc(ctx.handle, nbytes, C.byref(p)))
        self.ctx = ctx
        self.ptr = p.value
        self.nbytes = nbytes

    def array(self, dtype, count: int, offset: int = 0) -> np.ndarray:
        """numpy view of `count` elements at byte `offset` (no copy)."""
        dt = np.dtype(dtype)
        assert offset + count * dt.itemsize <= self.nbytes
        buf = (C.c_char * (count * dt.itemsize)).from_address(self.ptr + offset)
        return np.frombuffer(buf, dtype=dt, count=count)

    def free(self) -> None:
        if self.ptr:
            lib().crx_host_free(self.ctx.handle, C.c_void_p(self.ptr))
            self.ptr = None


def server_config(**kw) -> ServerConfig:
    c = ServerConfig()
    lib().crx_server_config_default(C.byref(c))
    for k, v in kw.items():
        setattr(c, k, v)
    return c


class Server:
    """Continuous-batching render server over an out-of-core node store (crx_server_*)."""

    def __init__(self, ctx: "Context", config: ServerConfig | None = None):
        h = C.c_void_p()
        cfg = config if config is not None else server_config()
        _check(lib().crx_server_create(ctx.handle, C.byref(cfg), C.byref(h)))
        self.ctx = ctx
        self.handle = h.value
        self.config = cfg

    def add_node(self, positions_ptr: int, num_vertices: int, indices_ptr, num_triangles: int, colors_ptr,
                 aabb: Box3, world: Mat4 | None = None, resident: bool = False) -> int:
        nid = C.c_uint32()
        w = identity() if world is None else world
        _check(lib().crx_server_add_node(self.handle, C.c_void_p(positions_ptr), num_vertices,
                                         C.c_void_p(indices_ptr) if indices_ptr else None, num_triangles,
                                         C.c_void_p(colors_ptr) if colors_ptr else None, C.byref(aabb),
                                         C.byref(w), int(bool(resident)), C.byref(nid)))
        return nid.value

    def start(self) -> None:
        _check(lib().crx_server_start(self.handle))

    def submit(self, view: View) -> int:
        t = C.c_uint64()
        _check(lib().crx_server_submit(self.handle, C.byref(view), C.byref(t)))
        return t.value

    def wait(self, ticket: int) -> RequestResult:
        r = RequestResult()
        _check(lib().crx_server_wait(self.handle, ticket, C.byref(r)))
        return r

    def release(self, ticket: int) -> None:
        _check(lib().crx_server_release(self.handle, ticket))

    def clock(self) -> float:
        return float(lib().crx_server_clock(self.handle))

    def stats(self) -> ServerStats:
        st = ServerStats()
        _check(lib().crx_server_get_stats(self.handle, C.byref(st)))
        return st

    def view_nodes(self, view: View) -> list[int]:
        n = C.c_uint32()
        _check(lib().crx_server_view_nodes(self.handle, C.byref(view), None, 0, C.byref(n)))
        ids = (C.c_uint32 * max(1, n.value))()
        _check(lib().crx_server_view_nodes(self.handle, C.byref(view), ids, n.value, C.byref(n)))
        return list(ids[:n.value])

    def download(self, ptr: int, count: int) -> np.ndarray:
        out = np.empty(count, dtype=np.uint64)
        _check(lib().crx_memcpy_dtoh(self.ctx.handle, out.ctypes.data_as(C.c_void_p), C.c_void_p(ptr),
                                     8 * count, None))
        return out

    def close(self) -> None:
        if self.handle:
            lib().crx_server_destroy(C.c_void_p(self.handle))
            self.handle = None


class Frame:
    """A scene frame recorded once into a hipGraph (crx_frame_*): clear,
    stages 1-3 and resolve for fixed view sizes and buffers."""

    def __init__(self, scene: "Scene", views, fbs, cbs=None, settings: RenderSettings | None = None):
        v_arr = _as_array(View, views)
        n = len(v_arr)
        self.n = n
        fb_arr = (C.c_void_p * n)(*[b.ptr for b in fbs])
        cb_arr = (C.c_void_p * n)(*[b.ptr for b in cbs]) if cbs is not None else None
        s = settings if settings is not None else render_settings()
        h = C.c_void_p()
        _check(lib().crx_frame_create(scene.ctx.handle, scene.handle, v_arr, n, C.byref(s), fb_arr, cb_arr,
                                      C.byref(h)))
        self.handle = h.value

    def set_views(self, views) -> None:
        v_arr = _as_array(View, views)
        assert len(v_arr) == self.n
        _check(lib().crx_frame_set_views(self.handle, v_arr))

    def launch(self, stream=None) -> None:
        _check(lib().crx_frame_launch(self.handle, stream))

    def wait(self) -> DrawStats:
        st = DrawStats()
        _check(lib().crx_frame_wait(self.handle, C.byref(st)))
        return st

    def close(self) -> None:
        if self.handle:
            lib().crx_frame_destroy(C.c_void_p(self.handle))
            self.handle = None


def render_settings(**kw) -> RenderSettings:
    s = RenderSettings()
    lib().crx_render_settings_default(C.byref(s))
    for k, v in kw.items():
        setattr(s, k, v)
    return s


def make_view(view: Mat4, proj: Mat4, width: int, height: int) -> View:
    v = View()
    v.view = view
    v.proj = proj
    v.width = width
    v.height = height
    return v
