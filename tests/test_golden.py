"""Golden fixtures (tests/golden/, written by tests/golden/make_golden.py).

They pin the seeded generators by array hashes and the CPU oracle by full
visibility / colour buffers of small scenes that exercise all three stages,
mirrored instances and a non-default routing, plus the 1080p sphere (config
1) by whole-buffer hashes and sampled pixels.  On the GPU the same scenes are
rendered through the C-ABI and compared with the fixtures directly.
"""
import json
from pathlib import Path

import numpy as np
import pytest

import crx
import workloads as wl
from crxtest import GpuScene, oracle_render

HERE = Path(__file__).resolve().parent / "golden"
META = json.loads((HERE / "golden.json").read_text())


def _scenes():
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_golden", HERE / "make_golden.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_generator_hashes():
    mod = _scenes()
    assert mod.generator_hashes() == {k: (v if isinstance(v, int) else list(v)) for k, v in META["generators"].items()}


@pytest.mark.parametrize("name", ["sphere_64", "terrain_closeup_64x48", "instanced_mirrored_64"])
def test_oracle_matches_small_fixtures(name):
    mod = _scenes()
    spec, W, H, view, proj, inst, tuning = mod.scenes()[name]
    fb, cb, st = oracle_render(spec, W, H, view, proj, instanced=inst, tuning=tuning,
                               attribute=crx.DISPLAY_VERTEX_COLORS)
    assert np.array_equal(fb.reshape(H, W), np.load(HERE / f"{name}_fb.npy"))
    assert np.array_equal(cb.reshape(H, W), np.load(HERE / f"{name}_cb.npy"))
    assert (st.num_nontrivial, st.num_huge) == (META[name]["nontrivial"], META[name]["huge"])


def test_fixture_scenes_cover_all_stages():
    assert META["terrain_closeup_64x48"]["nontrivial"] > 0 and META["terrain_closeup_64x48"]["huge"] > 0
    assert (np.load(HERE / "sphere_64_fb.npy") != crx.CLEAR_PIXEL).sum() > 300


def test_oracle_matches_sphere_1080p_fixture():
    mod = _scenes()
    fb, cb, st = mod.c1_frame()
    g = META["sphere_1m_1080p"]
    for i, vfb, vcb in g["samples"]:
        assert int(fb[i]) == vfb and int(cb[i]) == vcb
    assert wl.hash64(fb) == g["fb_hash"] and wl.hash64(cb) == g["cb_hash"]
    crop = (cb.reshape(1080, 1920)[524:556, 944:976] & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    assert np.array_equal(crop, np.load(HERE / "sphere_1m_1080p_rgba_crop.npy"))


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["sphere_64", "terrain_closeup_64x48", "instanced_mirrored_64"])
def test_gpu_matches_small_fixtures(name):
    mod = _scenes()
    spec, W, H, view, proj, inst, tuning = mod.scenes()[name]
    ctx = crx.Context(0)
    try:
        if tuning is not None:
            ctx.tuning = tuning
        fb, cb, st = GpuScene(ctx, spec).render(W, H, view, proj, instanced=inst, attribute=crx.DISPLAY_VERTEX_COLORS)
    finally:
        ctx.close()
    assert np.array_equal(fb.reshape(H, W), np.load(HERE / f"{name}_fb.npy"))
    assert np.array_equal(cb.reshape(H, W), np.load(HERE / f"{name}_cb.npy"))


@pytest.mark.gpu
def test_gpu_matches_sphere_1080p_fixture():
    import oracle as orc
    from crxtest import SceneSpec
    pos, idx, col = wl.sphere_mesh()
    spec = SceneSpec().add(orc.HostMesh(crx, pos, idx, col), [crx.identity()])
    view, proj = wl.sphere_view(1920, 1080)
    ctx = crx.Context(0)
    try:
        fb, cb, st = GpuScene(ctx, spec).render(1920, 1080, view, proj, attribute=crx.DISPLAY_VERTEX_COLORS)
    finally:
        ctx.close()
    g = META["sphere_1m_1080p"]
    assert wl.hash64(fb) == g["fb_hash"] and wl.hash64(cb) == g["cb_hash"]
