"""Regenerate the golden fixtures of tests/test_golden.py.

    python tests/golden/make_golden.py

The fixtures pin the seeded workload generators (array hashes) and the CPU
oracle's output on small scenes (full visibility buffers <= 64x64 px, the
1080p sphere as a whole-buffer hash plus sampled pixels, a resolve crop).
They are data only: inputs are the seeded generators, outputs the oracle's.
"""
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
for p in (ROOT / "curast-ooc_amd", ROOT / "oracle", ROOT / "tests"):
    sys.path.insert(0, str(p))

import crx  # noqa: E402
import oracle as orc  # noqa: E402
import workloads as wl  # noqa: E402
from crxtest import SceneSpec, oracle_render  # noqa: E402


def scenes():
    """name -> (spec, W, H, view, proj, instanced, tuning)"""
    out = {}
    pos, idx, col = wl.sphere_mesh(rings=32, segments=64)
    view, proj = wl.sphere_view(64, 64)
    out["sphere_64"] = (SceneSpec().add(orc.HostMesh(crx, pos, idx, col), [crx.identity()]), 64, 64, view, proj, True,
                        None)
    hp, hi, hc = wl.heightfield_mesh(33, amplitude=3.0, scale=8.0)
    view = crx.look_at((16.0, 2.0, 10.0), (16.0, -1.0, 14.0), (0.0, 1.0, 0.0))
    t = crx.default_tuning()
    t.small_threshold, t.tile_size = 32, 16          # routes more triangles to stages 2 and 3
    out["terrain_closeup_64x48"] = (SceneSpec().add(orc.HostMesh(crx, hp, hi, hc), [crx.identity()]), 64, 48, view,
                                    wl.projection(64, 48), True, t)
    worlds = list(wl.instance_transforms(12, extent=4.0, seed=31))
    for k in (0, 5, 9):
        for r in range(3):
            worlds[k].m[r] = -worlds[k].m[r]
    spos, sidx, scol = wl.sphere_mesh(rings=12, segments=24, seed=32)
    view, proj = wl.instances_view(extent=4.0, width=64, height=64, distance=9.0)
    out["instanced_mirrored_64"] = (SceneSpec().add(orc.HostMesh(crx, spos, sidx, scol), worlds), 64, 64, view, proj,
                                    True, None)
    return out


def generator_hashes():
    pos, idx, col = wl.sphere_mesh()
    hp, hi, hc = wl.heightfield_mesh(257)
    inst = wl.instance_transforms(8)
    return {
        "sphere_512x1024": [wl.hash64(pos), wl.hash64(idx), wl.hash64(col)],
        "heightfield_257": [wl.hash64(hp), wl.hash64(hi), wl.hash64(hc)],
        "instances_8": wl.hash64(np.frombuffer(bytes(inst), dtype=np.float32)),
    }


def c1_frame():
    pos, idx, col = wl.sphere_mesh()
    spec = SceneSpec().add(orc.HostMesh(crx, pos, idx, col), [crx.identity()])
    view, proj = wl.sphere_view(1920, 1080)
    fb, cb, st = oracle_render(spec, 1920, 1080, view, proj, attribute=crx.DISPLAY_VERTEX_COLORS)
    return fb, cb, st


def main():
    meta = {"generators": {k: v for k, v in generator_hashes().items()}}
    for name, (spec, W, H, view, proj, inst, tuning) in scenes().items():
        fb, cb, st = oracle_render(spec, W, H, view, proj, instanced=inst, tuning=tuning,
                                   attribute=crx.DISPLAY_VERTEX_COLORS)
        np.save(HERE / f"{name}_fb.npy", fb.reshape(H, W))
        np.save(HERE / f"{name}_cb.npy", cb.reshape(H, W))
        meta[name] = {"nontrivial": int(st.num_nontrivial), "huge": int(st.num_huge)}
    fb, cb, st = c1_frame()
    rng = np.random.default_rng(0)
    covered = np.nonzero(fb != crx.CLEAR_PIXEL)[0]
    sample = np.sort(rng.choice(covered, 64, replace=False))
    meta["sphere_1m_1080p"] = {
        "fb_hash": wl.hash64(fb), "cb_hash": wl.hash64(cb), "covered": int(covered.size),
        "nontrivial": int(st.num_nontrivial),
        "samples": [[int(i), int(fb[i]), int(cb[i])] for i in sample],
    }
    crop = (cb.reshape(1080, 1920)[524:556, 944:976] & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    np.save(HERE / "sphere_1m_1080p_rgba_crop.npy", crop)
    (HERE / "golden.json").write_text(json.dumps(meta, indent=1))
    print("wrote", HERE)


if __name__ == "__main__":
    main()
