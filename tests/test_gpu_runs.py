"""Stage 1 over pairs of consecutive triangles (k_stage1_pairs).

A lane draws two consecutive triangles; a wave whose pairs all share two
vertices in one of the two quad patterns gathers each shared vertex once and
computes its depth and view y row once; a mesh stored in rows is walked in
units of consecutive rows, each drawn by one wave.  These change which
lane draws which triangle and in what order, never a result: every layout,
run length and strip height must give the oracle's buffers bit for bit, on
grids whose rows are long enough for strips, on meshes with no vertex
sharing between neighbours (shuffled triangle order), on non-indexed meshes
and on rows shorter than four batches, for float meshes and for IX_PU16
meshes (bit-packed indices cut out through LDS, 3 x uint16 positions).
"""
import numpy as np
import pytest

import crx
import oracle as orc
import workloads as wl
from crxtest import CompressedMesh, GpuScene, SceneSpec, describe_mismatch, oracle_render

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = crx.Context(0)
    yield c
    c.close()


def _render(ctx, meshes_worlds, W, H, view, proj, tuning):
    ctx.tuning = tuning
    try:
        scene = crx.Scene(ctx)
        for mesh, worlds in meshes_worlds:
            for w in worlds:
                scene.add(mesh, w)
        fb, cb = ctx.alloc(8 * W * H), ctx.alloc(8 * W * H)
        st = scene.render(crx.make_view(view, proj, W, H), fb, cb, crx.render_settings())
        out = fb.download(np.uint64, W * H), cb.download(np.uint64, W * H), st
        fb.free()
        cb.free()
        scene.close()
        return out
    finally:
        ctx.tuning = crx.default_tuning()


def test_row_length_detected(ctx):
    hp, hi, hc = wl.heightfield_mesh(1200, amplitude=20.0, scale=200.0)
    sp, si, sc = wl.sphere_mesh()
    rng = np.random.default_rng(3)
    shuffled = si[rng.permutation(len(si))]
    meshes = [crx.Mesh(ctx, hp, hi, hc), crx.Mesh(ctx, sp, si, sc), crx.Mesh(ctx, sp, shuffled, sc),
              crx.Mesh(ctx, sp[si.reshape(-1)], None, sc[si.reshape(-1)])]
    got = [m.row_triangles() for m in meshes]
    for m in meshes:
        m.close()
    # a vertex's first use in the next row comes one row and one triangle
    # later; the row length is that distance rounded down to even
    assert got[0] == 2 * 1199 and got[1] == 2 * 1024
    assert got[2] == 0 and got[3] == 0
    # a compressed mesh keeps the row length of its index order
    packed = crx.Mesh(ctx, hp, hi, hc, compressed=True)
    try:
        assert packed.row_triangles() == 2 * 1199
    finally:
        packed.close()


def _terrain_views():
    W, H = 1280, 720
    oblique = crx.look_at((600.0, 120.0, -80.0), (600.0, 0.0, 500.0))
    close = crx.look_at((600.0, 6.0, 590.0), (600.0, 0.0, 640.0))     # stages 2 and 3
    return W, H, [oblique, close]


def test_layouts_bit_identical_on_a_grid(ctx):
    """A 1200 x 1200 grid (2398-triangle rows as detected: 19 strips of 128,
    the last one partly filled) seen obliquely and from close up, with units
    of 1 to 32 rows."""
    hp, hi, hc = wl.heightfield_mesh(1200, amplitude=20.0, scale=200.0)
    mesh = crx.Mesh(ctx, hp, hi, hc)
    spec = SceneSpec().add(orc.HostMesh(crx, hp, hi, hc), [crx.identity()])
    W, H, views = _terrain_views()
    proj = wl.projection(W, H)
    try:
        for vk, view in enumerate(views):
            want_fb, want_cb, ost = oracle_render(spec, W, H, view, proj, attribute=crx.DISPLAY_VERTEX_COLORS)
            if vk == 1:
                assert ost.num_nontrivial > 0 and ost.num_huge > 0
            for layout in (0, 1, 2):
                for per_lane in (2, 4):
                    for rows in ((1, 3, 8, 32) if per_lane == 2 and layout else (8,)):
                        # strips of 1 or 3 batches (2398 = 6 x 384 + 94: the last strip's
                        # second and third columns lie past the row's end)
                        for cols in ((1, 3) if per_lane == 2 and layout == 2 else (1,)):
                            print(f"[case] view {vk} layout {layout} tris/lane {per_lane} strip rows {rows} "
                                  f"strip batches {cols}", flush=True)
                            t = crx.default_tuning()
                            t.stage1_layout, t.stage1_tris_per_lane, t.stage1_strip_rows = layout, per_lane, rows
                            t.stage1_strip_batches = cols
                            # a grid of 4 waves per CU: 2.9M triangles are enough batches per wave
                            t.stage1_waves_per_cu = 4
                            fb, cb, st = _render(ctx, [(mesh, [crx.identity()])], W, H, view, proj, t)
                            assert st.stage1_kernel == (2 if layout and per_lane == 2 else 0)
                            assert np.array_equal(fb, want_fb), describe_mismatch(fb, want_fb, W)
                            assert np.array_equal(cb, want_cb), describe_mismatch(cb, want_cb, W)
                            assert (st.num_nontrivial, st.num_huge) == (ost.num_nontrivial, ost.num_huge)
    finally:
        mesh.close()


def test_runs_without_shared_vertices_and_non_indexed(ctx):
    """Triangles in random order (no neighbour shares a vertex), a
    non-indexed copy, and the indexed sphere (no row hint: a contiguous
    walk), drawn together over a few instances through
    crx_draw_triangles_visbuffer."""
    sp, si, sc = wl.sphere_mesh(rings=256, segments=512, seed=91)
    rng = np.random.default_rng(92)
    shuffled = si[rng.permutation(len(si))]
    flat_p, flat_c = sp[si.reshape(-1)], sc[si.reshape(-1)]
    worlds = list(wl.instance_transforms(3, extent=3.0, smin=0.4, smax=0.9, seed=93))
    spec = (SceneSpec().add(orc.HostMesh(crx, sp, shuffled, sc), worlds[:1])
            .add(orc.HostMesh(crx, flat_p, None, flat_c), worlds[1:2])
            .add(orc.HostMesh(crx, sp, si, sc), worlds[2:]))
    W, H = 960, 544
    proj = wl.projection(W, H)
    view = crx.look_at((1.5, 1.5, 6.0), (1.5, 1.5, 0.0))
    want_fb, want_cb, _ = oracle_render(spec, W, H, view, proj, attribute=crx.DISPLAY_VERTEX_COLORS)
    assert int((want_fb != crx.CLEAR_PIXEL).sum()) > 10_000
    g = GpuScene(ctx, spec)
    for instanced in (True, False):
        for layout in (1, 2):
            for per_lane in (2, 4):
                print(f"[case] mixed meshes instanced {instanced} layout {layout} tris/lane {per_lane}", flush=True)
                t = crx.default_tuning()
                t.stage1_layout, t.stage1_tris_per_lane = layout, per_lane
                t.stage1_waves_per_cu, t.stage1_strip_rows = 1, 2    # 786K triangles: a small grid
                ctx.tuning = t
                try:
                    fb, cb, st = g.render(W, H, view, proj, instanced=instanced, attribute=crx.DISPLAY_VERTEX_COLORS)
                finally:
                    ctx.tuning = crx.default_tuning()
                assert st.stage1_kernel == (2 if per_lane == 2 else 0)
                assert np.array_equal(fb, want_fb), describe_mismatch(fb, want_fb, W)
                assert np.array_equal(cb, want_cb)


def test_compressed_layouts_bit_identical_on_a_grid(ctx):
    """The IX_PU16 copy of the 1200 x 1200 grid (21-bit indices: a strip's
    6 x 128 indices start at any bit of a word), both views, units of 1 and
    8 rows, against the oracle's compressed path."""
    hp, hi, hc = wl.heightfield_mesh(1200, amplitude=20.0, scale=200.0)
    mesh = crx.Mesh(ctx, hp, hi, hc, compressed=True)
    spec = SceneSpec().add(CompressedMesh(hp, hi, hc), [crx.identity()])
    assert spec.meshes[0].meta.bitsPerIndex == 21
    W, H, views = _terrain_views()
    proj = wl.projection(W, H)
    try:
        for vk, view in enumerate(views):
            want_fb, want_cb, ost = oracle_render(spec, W, H, view, proj, attribute=crx.DISPLAY_VERTEX_COLORS)
            for layout in (0, 1, 2):
                for rows in ((1, 8) if layout else (8,)):
                    print(f"[case] compressed view {vk} layout {layout} strip rows {rows}", flush=True)
                    t = crx.default_tuning()
                    t.stage1_layout, t.stage1_tris_per_lane, t.stage1_strip_rows = layout, 2, rows
                    t.stage1_strip_batches = 2 if rows == 8 else 1
                    t.stage1_waves_per_cu = 4
                    fb, cb, st = _render(ctx, [(mesh, [crx.identity()])], W, H, view, proj, t)
                    assert st.stage1_kernel == (2 if layout else 0)
                    assert np.array_equal(fb, want_fb), describe_mismatch(fb, want_fb, W)
                    assert np.array_equal(cb, want_cb), describe_mismatch(cb, want_cb, W)
                    assert (st.num_nontrivial, st.num_huge) == (ost.num_nontrivial, ost.num_huge)
    finally:
        mesh.close()


def test_compressed_runs_without_shared_vertices(ctx):
    """IX_PU16 copies of the shuffled sphere (no pattern) and the sphere in
    order (the fan pattern, a contiguous walk: no row hint through raw
    pointers), drawn in one draw over a few instances."""
    sp, si, sc = wl.sphere_mesh(rings=256, segments=512, seed=94)
    rng = np.random.default_rng(95)
    shuffled = si[rng.permutation(len(si))]
    worlds = list(wl.instance_transforms(3, extent=3.0, smin=0.4, smax=0.9, seed=96))
    spec = (SceneSpec().add(CompressedMesh(sp, shuffled, sc), worlds[:1])
            .add(CompressedMesh(sp, si, sc), worlds[1:]))
    W, H = 960, 544
    proj = wl.projection(W, H)
    view = crx.look_at((1.5, 1.5, 6.0), (1.5, 1.5, 0.0))
    want_fb, want_cb, _ = oracle_render(spec, W, H, view, proj, attribute=crx.DISPLAY_VERTEX_COLORS)
    assert int((want_fb != crx.CLEAR_PIXEL).sum()) > 10_000
    g = GpuScene(ctx, spec)
    for instanced in (True, False):
        for layout in (0, 1, 2):
            print(f"[case] compressed mixed meshes instanced {instanced} layout {layout}", flush=True)
            t = crx.default_tuning()
            t.stage1_layout, t.stage1_tris_per_lane = layout, 2
            t.stage1_waves_per_cu, t.stage1_strip_rows = 1, 2
            ctx.tuning = t
            try:
                fb, cb, st = g.render(W, H, view, proj, instanced=instanced, attribute=crx.DISPLAY_VERTEX_COLORS)
            finally:
                ctx.tuning = crx.default_tuning()
            assert st.stage1_kernel == (2 if layout else 0)
            assert np.array_equal(fb, want_fb), describe_mismatch(fb, want_fb, W)
            assert np.array_equal(cb, want_cb)
