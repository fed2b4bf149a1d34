"""Shared helpers of the test-suite: build GPU and oracle copies of one scene
from the same host arrays and compare their buffers."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

import crx
import oracle as orc


class CompressedMesh:
    """IX_PU16 copy of a mesh: 3 x uint16 positions, bit-packed indices and
    the CMesh compression fields (crx_compress_mesh)."""

    def __init__(self, positions, indices, colors=None, uvs=None, texture=None, texture_size=None, jpeg=False):
        self.positions, self.indices, self.meta = crx.compress_mesh(positions, indices)
        self.colors = None if colors is None else np.ascontiguousarray(colors, dtype=np.uint32).reshape(-1)
        self.uvs = None if uvs is None else np.ascontiguousarray(uvs, dtype=np.float32).reshape(-1, 2)
        self.texture = None if texture is None else np.ascontiguousarray(texture, dtype=np.uint32).reshape(-1)
        self.texture_size = (0, 0) if texture_size is None else (int(texture_size[0]), int(texture_size[1]))
        self.jpeg = bool(jpeg)
        self._huffman = np.zeros(1, dtype=np.uint32)
        self.num_triangles = len(np.asarray(indices).reshape(-1, 3))

    set_texture = orc.HostMesh.set_texture

    def decoded_positions(self):
        """Float positions as the raster stages decode them: X * factor + min
        per axis in float32, a product then a sum."""
        cf = np.array([self.meta.compressionFactor.x, self.meta.compressionFactor.y,
                       self.meta.compressionFactor.z], dtype=np.float32)
        lo = np.array([self.meta.aabb.min.x, self.meta.aabb.min.y, self.meta.aabb.min.z], dtype=np.float32)
        return (self.positions.astype(np.float32) * cf) + lo


def _apply_compression(m, mesh):
    meta = getattr(mesh, "meta", None)
    if meta is None:
        return
    m.compressed = 1
    m.bitsPerIndex = meta.bitsPerIndex
    m.index_min = meta.index_min
    m.index_max = meta.index_max
    m.compressionFactor = meta.compressionFactor
    m.aabb = meta.aabb


@dataclass
class SceneSpec:
    """Meshes and, per mesh, the world matrices of its instances."""
    meshes: list = field(default_factory=list)        # list[orc.HostMesh]
    worlds: list = field(default_factory=list)        # list[list[crx.Mat4]]

    def add(self, mesh: "orc.HostMesh", worlds):
        self.meshes.append(mesh)
        self.worlds.append(list(worlds))
        return self


@dataclass
class Frame:
    """CMesh lists as CuRast::draw prepares them (CuRast_render.h:356-435)."""
    unique: list
    instances: list
    transforms: list
    prefix: np.ndarray


def mip_chain(levels):
    """One flat uint32 array of RGBA8 mip levels, level 0 first; each level k
    is an (h, w) array of ((w_(k-1) + 1) // 2, (h_(k-1) + 1) // 2) texels."""
    for a, b in zip(levels, levels[1:]):
        assert b.shape == ((a.shape[0] + 1) // 2, (a.shape[1] + 1) // 2), (a.shape, b.shape)
    return np.concatenate([np.ascontiguousarray(l, dtype=np.uint32).reshape(-1) for l in levels])


def mip_shapes(w, h, n=8):
    """(h, w) of the first n levels of a w x h texture."""
    out = []
    for _ in range(n):
        out.append((h, w))
        w, h = (w + 1) // 2, (h + 1) // 2
    return out


def prepare(spec: SceneSpec, view: crx.Mat4, pointers, instanced_ids=True) -> Frame:
    """pointers[k] = (positions, indices, colors[, uvs, texture[, normals]]) addresses of mesh k."""
    unique, instances, transforms, prefix = [], [], [], []
    total = 0
    for k, (mesh, worlds) in enumerate(zip(spec.meshes, spec.worlds)):
        first = len(instances)
        for w in worlds:
            m = crx.CMesh()
            p, i, c = pointers[k][:3]
            uv, tex = pointers[k][3:5] if len(pointers[k]) > 3 else (None, None)
            mesh.set_texture(m, uv, tex)
            m.normals = pointers[k][5] if len(pointers[k]) > 5 else None
            m.numTriangles = mesh.num_triangles
            m.positions = p
            m.indices = i
            m.colors = c
            m.cummulativeTriangleCount = total
            m.instances.offset = len(instances)
            m.instances.count = 1
            m.world = w
            m.id = len(instances)
            m.numVertices = len(mesh.positions)
            m.isLoaded = 1
            _apply_compression(m, mesh)
            prefix.append(total)
            total += mesh.num_triangles
            instances.append(m)
            transforms.append(crx.mat4_mul(view, w))
        if worlds:
            u = crx.CMesh.from_buffer_copy(instances[first])
            u.instances.count = len(worlds)
            unique.append(u)
    return Frame(unique, instances, transforms, np.asarray(prefix, dtype=np.uint64))


def _uvs(mesh):
    return getattr(mesh, "uvs", None)


def _texture(mesh):
    return getattr(mesh, "texture", None)


def _normals(mesh):
    return getattr(mesh, "normals", None)


def host_pointers(spec: SceneSpec):
    out = []
    for m in spec.meshes:
        out.append((m.positions.ctypes.data,
                    m.indices.ctypes.data if m.indices is not None else None,
                    m.colors.ctypes.data if m.colors is not None else None,
                    _uvs(m).ctypes.data if _uvs(m) is not None else None,
                    _texture(m).ctypes.data if _texture(m) is not None else None,
                    _normals(m).ctypes.data if _normals(m) is not None else None))
    return out


def oracle_render(spec: SceneSpec, W, H, view, proj, instanced=True, tuning=None, attribute=None,
                  threads=0):
    fr = prepare(spec, view, host_pointers(spec))
    fb, st = orc.draw(crx, fr.unique, fr.instances, fr.transforms, W, H, view, proj, instanced,
                      tuning=tuning, threads=threads)
    cb = None
    if attribute is not None:
        cb = orc.resolve(crx, fr.instances, fr.prefix, fb, W, H, view, proj, attribute, threads=threads)
    return fb, cb, st


class GpuScene:
    """Device copies of a SceneSpec's meshes on one context."""

    def __init__(self, ctx: crx.Context, spec: SceneSpec):
        self.ctx = ctx
        self.spec = spec
        self.pointers = []
        for m in spec.meshes:
            p = ctx.to_device(m.positions).ptr
            i = ctx.to_device(m.indices).ptr if m.indices is not None else None
            c = ctx.to_device(m.colors).ptr if m.colors is not None else None
            uv = ctx.to_device(_uvs(m)).ptr if _uvs(m) is not None else None
            tex = ctx.to_device(_texture(m)).ptr if _texture(m) is not None else None
            nr = ctx.to_device(_normals(m)).ptr if _normals(m) is not None else None
            self.pointers.append((p, i, c, uv, tex, nr))

    def render(self, W, H, view, proj, instanced=True, attribute=None, fb=None, cb=None, stream=None):
        ctx = self.ctx
        fr = prepare(self.spec, view, self.pointers)
        fb = fb if fb is not None else ctx.alloc(8 * W * H)
        cbuf = cb if cb is not None else (ctx.alloc(8 * W * H) if attribute is not None else None)
        ctx.clear_framebuffer(fb, W * H, stream)
        if cbuf is not None:
            ctx.clear_framebuffer(cbuf, W * H, stream)
        target = crx.make_target(fb, cbuf, W, H, view, proj)
        st = ctx.draw_triangles_visbuffer(fr.unique, fr.instances, fr.transforms, target, instanced, stream)
        out_cb = None
        if attribute is not None:
            ctx.resolve_visbuffer(fr.instances, fr.prefix, target, attribute, stream)
            out_cb = cbuf.download(np.uint64, W * H, stream)
        out_fb = fb.download(np.uint64, W * H, stream)
        return out_fb, out_cb, st

    def render_views(self, views, instanced=True, attribute=None, stream=None):
        """One batched draw (and resolve) of all `views` = [(W, H, view, proj)].
        Returns per-view framebuffers, colour buffers (or None) and the stats."""
        ctx = self.ctx
        frames = [prepare(self.spec, v, self.pointers) for (_, _, v, _) in views]
        fr0 = frames[0]
        transforms = [t for fr in frames for t in fr.transforms]
        fbs = [ctx.alloc(8 * W * H) for (W, H, _, _) in views]
        cbs = [ctx.alloc(8 * W * H) for (W, H, _, _) in views] if attribute is not None else None
        targets = [crx.make_target(fbs[k], cbs[k] if cbs else None, W, H, v, p)
                   for k, (W, H, v, p) in enumerate(views)]
        ctx.clear_views(targets, colorbuffers=cbs is not None, stream=stream)
        st = ctx.draw_triangles_visbuffer_views(fr0.unique, fr0.instances, transforms, targets, instanced, stream)
        out_cb = None
        if attribute is not None:
            ctx.resolve_visbuffer_views(fr0.instances, fr0.prefix, targets, attribute, stream)
            out_cb = [cbs[k].download(np.uint64, W * H, stream) for k, (W, H, _, _) in enumerate(views)]
        out_fb = [fbs[k].download(np.uint64, W * H, stream) for k, (W, H, _, _) in enumerate(views)]
        return out_fb, out_cb, st


def describe_mismatch(a: np.ndarray, b: np.ndarray, W: int) -> str:
    bad = np.nonzero(a != b)[0]
    if bad.size == 0:
        return "identical"
    lines = [f"{bad.size} of {a.size} pixels differ"]
    for i in bad[:8]:
        lines.append(f"  px ({i % W},{i // W}): got {int(a[i]):#018x} want {int(b[i]):#018x}")
    return "\n".join(lines)


def rgba(cb: np.ndarray, W: int, H: int) -> np.ndarray:
    return (cb & 0xFFFFFFFF).astype(np.uint32).view(np.uint8).reshape(H, W, 4)
