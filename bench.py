#!/usr/bin/env python3
"""Benchmark of the gfx950 visibility-buffer rasterizer on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload heightfield|sphere|instances]

A step is one frame at 1920x1080: clear, the three raster stages and the
resolve to RGBA8 vertex colours, over a scene already resident in HBM.  The
default workload is the 400M-triangle heightfield (14142 x 14142 grid, seeded
6-octave value noise, one oblique view).  Prints one JSON line:

* value       triangles per second over the whole timed frames (wall clock)
* roofline    stage 1 (the dominant kernel) against the 8 TB/s HBM peak, with
              its algorithmic bytes 12*T (indices) + 12*V (positions) per
              launch over the stage-1 duration measured with HIP events on
              the stream the kernel runs on
* cpu_baseline  the CPU oracle timed on this host (OpenMP) rendering the same
              frame, repeated until at least 10 s of wall time
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "curast-ooc_amd"))

HBM_PEAK_GBS = 8000.0
WIDTH, HEIGHT = 1920, 1080


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def build_workload(name: str):
    import crx
    import workloads as wl
    if name == "heightfield":
        n = int(os.environ.get("CRX_BENCH_N", "14142"))
        pos, idx, col = wl.heightfield_mesh(n, threads=int(os.environ.get("OMP_NUM_THREADS", "16")))
        view, proj = wl.heightfield_view(n, WIDTH, HEIGHT)
        worlds = [crx.identity()]
        desc = {"workload": f"heightfield-{n}x{n}-1080p", "grid": n}
    elif name == "sphere":
        pos, idx, col = wl.sphere_mesh()
        view, proj = wl.sphere_view(WIDTH, HEIGHT)
        worlds = [crx.identity()]
        desc = {"workload": "displaced-sphere-1M-1080p"}
    elif name == "instances":
        pos, idx, col = wl.sphere_mesh()
        count = int(os.environ.get("CRX_BENCH_INSTANCES", "3000"))
        worlds = list(wl.instance_transforms(count))
        view, proj = wl.instances_view()
        desc = {"workload": f"sphere-1M-x{count}-instances-1080p", "instances": count}
    else:
        raise SystemExit(f"unknown workload {name}")
    return pos, idx, col, worlds, view, proj, desc


def stage1_bytes(T: int, V: int) -> int:
    """Algorithmic HBM bytes of one stage-1 launch: every index once, every vertex once."""
    return 12 * T + 12 * V


def load_traffic(workload: str):
    """HBM bytes per stage-1 launch measured with rocprofv3 --pmc (profiles/)."""
    path = ROOT / "profiles" / f"stage1_{workload}_traffic.json"
    if not path.exists():
        return None
    try:
        return json.loads(path.read_text()).get("traffic_bytes")
    except (OSError, ValueError):
        return None


def cpu_baseline(pos, idx, col, worlds, view, proj, threads: int):
    """The CPU oracle rendering the same frame (stages + resolve), repeated for >= 10 s."""
    sys.path.insert(0, str(ROOT / "oracle"))
    sys.path.insert(0, str(ROOT / "tests"))
    import crx
    import oracle as orc
    from crxtest import SceneSpec, oracle_render
    spec = SceneSpec().add(orc.HostMesh(crx, pos, idx, col), worlds)
    T = len(idx) * len(worlds)
    reps, t0 = 0, time.perf_counter()
    while True:
        oracle_render(spec, WIDTH, HEIGHT, view, proj, attribute=crx.DISPLAY_VERTEX_COLORS, threads=threads)
        reps += 1
        el = time.perf_counter() - t0
        if el >= 10.0 or reps >= 50:
            break
    return {"value": T * reps / el, "unit": "triangles/s", "cores": threads, "kind": "port",
            "sample": f"{reps} full frame(s) of the same workload (stages + resolve) on the CPU oracle, "
                      f"{threads} OpenMP threads, {el:.1f} s"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="heightfield", choices=["heightfield", "sphere", "instances"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--probe-copy", action="store_true",
                    help="also run the 1 GiB copy kernel that calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE")
    ap.add_argument("--tuning", default="",
                    help="comma-separated crx_tuning overrides, e.g. stage1_order=0,stage1_block=256")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: --gpus {args.gpus} is not supported "
              "(the visibility-buffer path here is single-GPU; see DESIGN.md for how it would shard)",
              file=sys.stderr)
        return 2

    import crx
    t0 = time.perf_counter()
    pos, idx, col, worlds, view, proj, desc = build_workload(args.workload)
    log(f"[bench] generated {desc['workload']} in {time.perf_counter() - t0:.1f} s")
    ctx = crx.Context(0)
    if args.tuning:
        tu = ctx.tuning
        for kv in args.tuning.split(","):
            k, v = kv.split("=")
            setattr(tu, k.strip(), int(v))
        ctx.tuning = tu
        log(f"[bench] tuning overrides: {args.tuning}")
    mesh = crx.Mesh(ctx, pos, idx, col)
    scene = crx.Scene(ctx)
    for w in worlds:
        scene.add(mesh, w)
    v = crx.make_view(view, proj, WIDTH, HEIGHT)
    fb = ctx.alloc(8 * WIDTH * HEIGHT)
    cb = ctx.alloc(8 * WIDTH * HEIGHT)
    settings = crx.render_settings(instanced=1, frustum_culling=1, resolve=1,
                                   attribute=crx.DISPLAY_VERTEX_COLORS)
    T = len(idx) * len(worlds)
    V = len(pos)
    log(f"[bench] scene resident: {T} triangle-instances, {V} vertices; warmup {args.warmup}")

    for _ in range(args.warmup):
        scene.render(v, fb, cb, settings)
    ctx.synchronize()

    if args.probe_copy:
        nbytes = 1 << 30
        src, dst = ctx.alloc(nbytes), ctx.alloc(nbytes)
        for _ in range(3):
            ctx.copy(dst, src, nbytes)
        ctx.synchronize()
        src.free()
        dst.free()
        log(f"[bench] copy probe: 3 x k_copy of {nbytes} bytes")

    # per-stage device times (HIP events on the draw stream), separate from the wall-clock run
    ctx.enable_timing(True)
    stage = {"clear": [], "stage1": [], "stage2": [], "stage3": [], "resolve": [], "pipeline": []}
    for _ in range(min(args.steps, 60)):
        scene.render(v, fb, cb, settings)
        t = ctx.timings()
        stage["clear"].append(t.clear_ms)
        stage["stage1"].append(t.stage1_ms)
        stage["stage2"].append(t.stage2_ms)
        stage["stage3"].append(t.stage3_ms)
        stage["resolve"].append(t.resolve_ms)
        stage["pipeline"].append(t.pipeline_ms)
    ctx.enable_timing(False)
    med = {k: float(np.median(x)) for k, x in stage.items()}
    mean_s1 = float(np.mean(stage["stage1"]))

    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        scene.render(v, fb, cb, settings)
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps
    value = T * args.steps / elapsed
    log(f"[bench] {args.steps} frames in {elapsed:.3f} s: {ms_per_step:.3f} ms/frame; stages {med}")

    algo = stage1_bytes(len(idx), V) * (len(worlds) if args.workload != "instances" else 1)
    if args.workload == "instances":
        algo = stage1_bytes(len(idx), V) + 64 * len(worlds)
    achieved = algo / (mean_s1 * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": load_traffic(args.workload),
                "kernel": "k_stage1", "algorithmic_bytes_per_launch": algo,
                "stage1_ms_mean": round(mean_s1, 4)}

    cpu = None
    if not args.no_cpu_baseline:
        threads = int(os.environ.get("OMP_NUM_THREADS", "16"))
        log(f"[bench] CPU oracle baseline on {threads} threads ...")
        cpu = cpu_baseline(pos, idx, col, worlds, view, proj, threads)
        log(f"[bench] CPU baseline {cpu['value']:.3e} triangles/s ({cpu['sample']})")

    out = {
        "metric": "triangles/sec & ms/frame @1080p",
        "value": value,
        "unit": "triangles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": dict(desc, width=WIDTH, height=HEIGHT, triangles=T, vertices=V,
                       frame="clear + stage1 + stage2 + stage3 + resolve(VERTEX_COLORS)"),
        "stage_ms_median": {k: round(x, 4) for k, x in med.items()},
        "pipeline_triangles_per_s": T / (med["pipeline"] * 1e-3),
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
