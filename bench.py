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
* parity      the benchmarked frame's visibility and colour buffers against
              the CPU oracle's frame of the same workload, bit for bit
* cpu_baseline  the CPU oracle timed on this host rendering the same frame:
              on every CPU this process may use (capped by OMP_NUM_THREADS,
              the host's CPU share per GPU), and on one thread
* batched_multiview  64 mixed-resolution views of the heightfield: the
              batched frame (eager and as a captured hipGraph) against the
              plain single-view frame looped over the views
* instances   3000 instances of the 1M-triangle sphere: frame time and stage
              1 against the VALU issue rate
* serving     the 2G-triangle tiled heightfield served out of core at 200
              Poisson requests/s: p50 / p99 latency of batched and plain
              serving, and saturated throughput
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


def kernel_source_sha16() -> str:
    """Hash of the kernel sources and of the host code that plans their
    launches: a counter summary taken on other sources is stale."""
    import hashlib
    h = hashlib.sha256()
    for f in ("crx_kernels.hip", "crx_device.h", "crx_api.cpp"):
        h.update((ROOT / "curast-ooc_amd" / "csrc" / f).read_bytes())
    return h.hexdigest()[:16]


def load_profile(name: str):
    """A figure measured with rocprofv3 counters (profiles/<name>.json) if it
    was taken on the current kernel sources, else None."""
    path = ROOT / "profiles" / f"{name}.json"
    if not path.exists():
        return None
    try:
        d = json.loads(path.read_text())
    except (OSError, ValueError):
        return None
    return d if d.get("kernel_source_sha16") == kernel_source_sha16() else None


def cpu_info():
    """CPUs this process may use on this host and the CPU model."""
    model = ""
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    affinity = len(os.sched_getaffinity(0))
    quota = None
    try:
        q, period = Path("/sys/fs/cgroup/cpu.max").read_text().split()
        if q != "max":
            quota = round(int(q) / int(period), 2)
    except (OSError, ValueError):
        pass
    share = int(os.environ["OMP_NUM_THREADS"]) if os.environ.get("OMP_NUM_THREADS", "").isdigit() else affinity
    return {"model": model, "nproc": os.cpu_count(), "affinity": affinity, "cgroup_cpus": quota,
            "threads": max(1, min(affinity, share))}


def bench_multiview(ctx, scene, pos, n: int, steps: int, warmup: int, count: int = 64, autotune: bool = True):
    """Batched multi-view workload: `count` seeded views of the heightfield
    (mixed resolutions, a quarter close-ups), each a full frame (clear, three
    stages, resolve).  With `autotune` the context's tuning (stage-1 schedule,
    routing thresholds, tile size, queue depths) is first searched by
    crx_autotune on the batched frame; both paths then use it.  Batched = one
    crx_render_views call per step; plain = crx_render once per view.
    Returns views/s of both."""
    import crx
    import workloads as wl
    vs = wl.multiview_views(n, pos, count=count)
    views = [crx.make_view(v, p, W, H) for (v, p, W, H, _) in vs]
    fbs = [ctx.alloc(8 * v.width * v.height) for v in views]
    cbs = [ctx.alloc(8 * v.width * v.height) for v in views]
    settings = crx.render_settings(instanced=1, frustum_culling=1, resolve=1, attribute=crx.DISPLAY_VERTEX_COLORS)
    pixels = sum(v.width * v.height for v in views)
    tuned = None
    if autotune:
        t0 = time.perf_counter()
        res = ctx.autotune(scene, views, fbs, flags=crx.TUNE_SCHEDULE | crx.TUNE_ROUTING | crx.TUNE_QUEUES, reps=3)
        t = res.tuning
        tuned = {"start_ms": round(res.start_ms, 3), "best_ms": round(res.best_ms, 3), "trials": res.trials,
                 "search_s": round(time.perf_counter() - t0, 2),
                 "tuning": {f: getattr(t, f) for f, _ in t._fields_ if f != "reserved"}}

    def batched():
        return scene.render_views(views, fbs, cbs, settings)

    def plain():
        for k, v in enumerate(views):
            scene.render(v, fbs[k], cbs[k], settings)

    def timed(fn, reps):
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        ctx.synchronize()
        return time.perf_counter() - t0

    for _ in range(warmup):
        batched()
    # the batched frame must equal the plain frames bit for bit
    batched()
    ctx.synchronize()
    got = [b.download(np.uint64, v.width * v.height) for b, v in zip(fbs, views)]
    plain()
    ctx.synchronize()
    same = all(np.array_equal(g, b.download(np.uint64, v.width * v.height)) for g, b, v in zip(got, fbs, views))
    ctx.enable_timing(True)
    st = batched()
    t = ctx.timings()
    ctx.enable_timing(False)
    stage = {"clear": t.clear_ms, "stage1": t.stage1_ms, "stage2": t.stage2_ms, "stage3": t.stage3_ms,
             "resolve": t.resolve_ms}
    reps_b = max(3, steps // 10)
    tb = timed(batched, reps_b)
    reps_p = max(2, steps // 40)
    tp = timed(plain, reps_p)
    # the same batched frame recorded once into a hipGraph (crx_frame_*):
    # one graph launch per frame, no per-frame host work
    graph_fbs = [ctx.alloc(8 * v.width * v.height) for v in views]
    graph_cbs = [ctx.alloc(8 * v.width * v.height) for v in views]
    frame = crx.Frame(scene, views, graph_fbs, graph_cbs, settings)
    frame.launch()
    frame.wait()
    graph_same = all(np.array_equal(g, b.download(np.uint64, v.width * v.height))
                     for g, b, v in zip(got, graph_fbs, views))

    def graph():
        frame.launch()
        frame.wait()

    tg = timed(graph, reps_b)
    frame.close()
    for b in graph_fbs + graph_cbs:
        b.free()
    out = {
        "views": count, "pixels": pixels, "closeups": int(sum(c for *_, c in vs)),
        "batched_views_per_s": count * reps_b / tb, "batched_ms_per_frame": 1e3 * tb / reps_b,
        "plain_views_per_s": count * reps_p / tp, "plain_ms_per_frame": 1e3 * tp / reps_p,
        "speedup": (count * reps_b / tb) / (count * reps_p / tp),
        "bit_identical_to_plain": bool(same),
        "graph_views_per_s": count * reps_b / tg, "graph_ms_per_frame": 1e3 * tg / reps_b,
        "graph_speedup": (count * reps_b / tg) / (count * reps_p / tp),
        "graph_bit_identical": bool(graph_same),
        "batched_stage_ms": {k: round(x, 4) for k, x in stage.items()},
        "batched_nontrivial": st.num_nontrivial, "batched_huge": st.num_huge,
        "autotune": tuned,
    }
    for b in fbs + cbs:
        b.free()
    ctx.tuning = crx.default_tuning()
    return out


def bench_atomics(ctx, log_entries: int = 21, per_lane: int = 64, reps: int = 5):
    """Measured peak of 64-bit atomicMin on a 16 MiB (1080p-sized) buffer,
    scattered and coalesced (crx_probe_atomic_min), in atomics/s."""
    buf = ctx.alloc(8 << log_entries)
    ctx.clear_framebuffer(buf, 1 << log_entries)
    out = {}
    for scattered in (True, False):
        ctx.probe_atomic_min(buf, log_entries, per_lane, scattered)
        ctx.synchronize()
        t0 = time.perf_counter()
        n = 0
        for _ in range(reps):
            n += ctx.probe_atomic_min(buf, log_entries, per_lane, scattered)
        ctx.synchronize()
        dt = time.perf_counter() - t0
        out["scattered" if scattered else "coalesced"] = {"atomics": n, "seconds": round(dt, 5),
                                                          "atomics_per_s": n / dt}
    buf.free()
    return out


def bench_hbm(ctx, nbytes: int = 1 << 30, reps: int = 20):
    """Measured HBM bandwidth of plain streaming kernels over nbytes of
    device memory: a read (crx_probe_read) and a copy (crx_copy_buffer, read
    plus write), in GB/s of bytes moved, from wall-clock time over `reps`
    launches between stream synchronisations."""
    src, dst, sink = ctx.alloc(nbytes), ctx.alloc(nbytes), ctx.alloc(16)
    out = {"bytes": nbytes}
    for name, fn, moved in (("read", lambda: ctx.probe_read(src, nbytes, sink), nbytes),
                            ("copy", lambda: ctx.copy(dst, src, nbytes), 2 * nbytes)):
        fn()
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        ctx.synchronize()
        dt = (time.perf_counter() - t0) / reps
        out[f"{name}_ms"] = round(dt * 1e3, 4)
        out[f"{name}_GBps"] = round(moved / dt / 1e9, 1)
    for b in (src, dst, sink):
        b.free()
    return out


VALU_PROBE_ITERS = 16384


def bench_valu(ctx, iters: int = VALU_PROBE_ITERS, reps: int = 10):
    """Measured VALU issue rate (crx_probe_valu): v_fma_f32 chains and
    32-bit add/xor chains on every lane, and the fma chains on 32 lanes and
    on 1 lane of each wave, as wall-clock launch times; the
    instructions per launch come from the SQ_INSTS_VALU counter of the same
    launch, recorded with the instanced stage-1 counters."""
    sink = ctx.alloc(16)
    out = {"iters": iters}
    # VALU instructions per wave and round as compiled: 8 v_fma_f32, or 8 v_add_u32 + 8 v_xor_b32;
    # the fma chains also with 32 lanes and with 1 lane of each wave active
    for mode, name, per_round in ((0, "fma_f32", 8), (1, "add_xor_u32", 16), (2, "fma_f32_32_lanes", 8),
                                  (3, "fma_f32_1_lane", 8), (4, "fma_and_int_interleaved", 12),
                                  (5, "fma_and_int_alternate_waves", 12)):
        ctx.probe_valu(iters, mode, sink)
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            waves = ctx.probe_valu(iters, mode, sink)
        ctx.synchronize()
        dt = (time.perf_counter() - t0) / reps
        out[name] = {"waves": waves, "ms": round(dt * 1e3, 4),
                     "nominal_G_insts_per_s": round(waves * per_round * iters / dt / 1e9, 1)}
    sink.free()
    return out


SERVING_LEGS = (
    # name, views per batch, lanes, paced (else every request queued at once), lanes share one stream
    ("batched", 64, 2, True, 0),
    ("batched_saturated", 64, 2, False, 0),
    ("plain", 1, 1, True, 0),
    ("plain_saturated", 1, 1, False, 0),
    ("batched_again", 64, 2, True, 0),
    ("plain_again", 1, 1, True, 0),
)


def bench_serving(ctx, n: int, tile: int, duration: float, rate: float, legs=SERVING_LEGS):
    """Out-of-core serving workload: an n x n heightfield cut into tile x tile
    nodes in pinned host memory, half of them device-resident at start, the
    device node cache holding half of the scene; a seeded Poisson request
    stream (`rate` per second for `duration` seconds, 8 moving cameras, mixed
    resolutions) is served by crx_server with continuous batching, then by the
    same server limited to one view per batch and one lane (the plain path).
    Reports throughput and p50 / p99 latency (submit to completion) of both."""
    import threading
    import crx
    import workloads as wl
    t0 = time.perf_counter()
    nodes, bufs = wl.tiled_heightfield(ctx, n, tile, threads=int(os.environ.get("OMP_NUM_THREADS", "16")))
    node_bytes = [(nd["nv"] * 12 + 255) // 256 * 256 + (nd["nt"] * 12 + 255) // 256 * 256 +
                  (nd["nv"] * 4 + 255) // 256 * 256 for nd in nodes]
    total_tris = sum(nd["nt"] for nd in nodes)
    log(f"[serve] {len(nodes)} nodes, {total_tris} triangles, {sum(node_bytes) / 1e9:.1f} GB pinned "
        f"in {time.perf_counter() - t0:.1f} s")
    ground = wl.ground_lookup(nodes, tile)
    reqs = wl.serving_requests(n, ground, duration=duration, rate=rate)
    views = [crx.make_view(v, p, W, H) for (_, v, p, W, H, _) in reqs]
    arrivals = [t for (t, *_) in reqs]

    def run(max_batch, nlanes, paced=True, shared_stream=1, upto=None):
        cfg = crx.server_config(node_cache_bytes=max(node_bytes) * (len(nodes) // 2), frame_bytes=8 << 30,
                                max_batch_views=max_batch, lanes=nlanes, resolve=1, shared_stream=shared_stream)
        srv = crx.Server(ctx, cfg)
        for k, nd in enumerate(nodes):
            srv.add_node(nd["pos"], nd["nv"], nd["idx"], nd["nt"], nd["col"], nd["aabb"], resident=(k % 2 == 0))
        t1 = time.perf_counter()
        srv.start()
        log(f"[serve] server started (batch {max_batch}, lanes {nlanes}) in {time.perf_counter() - t1:.1f} s")
        tickets = []
        lat, batches, ups, results = [], [], [], {}
        done = threading.Event()

        def reaper(nreq):
            k = 0
            while k < nreq:
                while k >= len(tickets):
                    time.sleep(0.0005)
                r = srv.wait(tickets[k])
                results[k] = r.status
                lat.append(r.done_time - r.submit_time)
                batches.append(r.batch_views)
                ups.append(r.uploads)
                srv.release(tickets[k])
                k += 1
            done.set()

        nreq = len(views) if upto is None else sum(1 for t in arrivals if t < upto)
        th = threading.Thread(target=reaper, args=(nreq,))
        th.start()
        base = time.perf_counter()
        for k, v in enumerate(views[:nreq]):
            delay = arrivals[k] - (time.perf_counter() - base) if paced else 0.0
            if delay > 0:
                time.sleep(delay)
            tickets.append(srv.submit(v))
        th.join()
        wall = time.perf_counter() - base
        st = srv.stats()
        srv.close()
        lat_ms = np.array(lat) * 1e3
        ok = sum(1 for s in results.values() if s == 0)
        pct = lambda a, q: round(float(np.percentile(a, q)), 3) if len(a) else None
        # requests whose batch streamed no node in, and requests after the
        # first second (the cache warmed up along the cameras' paths)
        no_up = lat_ms[np.array(ups) == 0]
        warm = lat_ms[np.array(arrivals[:len(lat_ms)]) >= 1.0] if paced else lat_ms
        return {"requests": nreq, "ok": ok, "wall_s": round(wall, 3), "views_per_s": nreq / wall,
                "p50_ms": pct(lat_ms, 50), "p90_ms": pct(lat_ms, 90), "p99_ms": pct(lat_ms, 99),
                "p999_ms": pct(lat_ms, 99.9), "max_ms": round(float(lat_ms.max()), 3),
                "requests_with_uploads": int(np.count_nonzero(np.array(ups))),
                "p99_without_uploads_ms": pct(no_up, 99), "p99_after_first_second_ms": pct(warm, 99),
                "mean_ms": round(float(lat_ms.mean()), 3), "mean_batch": round(float(np.mean(batches)), 2),
                "max_batch": st.max_batch, "uploads": st.uploads, "prefetches": st.prefetches,
                "upload_gb": round(st.upload_bytes / 1e9, 3),
                "evictions": st.evictions, "cache_slots": st.cache_slots}

    out = {"workload": f"tiled-heightfield-{n}x{n}-{len(nodes)}nodes", "triangles": total_tris,
           "pinned_gb": round(sum(node_bytes) / 1e9, 2), "rate_per_s": rate, "duration_s": duration}
    # the first 2 s of the stream once on the plain path, unreported: the
    # first server after the scene is pinned pays one-time host costs
    run(1, 1, paced=True, shared_stream=0, upto=2.0)
    for name, max_batch, nlanes, paced, shared in legs:
        out[name] = run(max_batch, nlanes, paced=paced, shared_stream=shared)
        log(f"[serve] {name} {out[name]}")
    if "batched_saturated" in out and "plain_saturated" in out:
        out["capacity_ratio"] = out["batched_saturated"]["views_per_s"] / out["plain_saturated"]["views_per_s"]
    for b in bufs:
        b.free()
    return out


def cpu_baseline(pos, idx, col, worlds, view, proj, compressed: bool = False, seconds: float = 8.0):
    """The CPU oracle rendering the same frame (stages + resolve): repeated
    for >= `seconds` on every CPU this process may use (up to the host's CPU
    share per GPU, OMP_NUM_THREADS), then once on one thread.  Also returns
    the last frame's buffers (the parity check of the GPU frame)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    sys.path.insert(0, str(ROOT / "tests"))
    import crx
    import oracle as orc
    from crxtest import CompressedMesh, SceneSpec, oracle_render
    info = cpu_info()
    threads = info["threads"]
    mesh = CompressedMesh(pos, idx, col) if compressed else orc.HostMesh(crx, pos, idx, col)
    spec = SceneSpec().add(mesh, worlds)
    T = len(idx) * len(worlds)
    reps, t0 = 0, time.perf_counter()
    while True:
        fb, cb, _ = oracle_render(spec, WIDTH, HEIGHT, view, proj, attribute=crx.DISPLAY_VERTEX_COLORS,
                                  threads=threads)
        reps += 1
        el = time.perf_counter() - t0
        if el >= seconds or reps >= 50:
            break
    t1 = time.perf_counter()
    oracle_render(spec, WIDTH, HEIGHT, view, proj, attribute=crx.DISPLAY_VERTEX_COLORS, threads=1)
    one = time.perf_counter() - t1
    out = {"value": T * reps / el, "unit": "triangles/s", "cores": threads, "kind": "port",
           "sample": f"{reps} full frame(s) of the same workload (stages + resolve) on the CPU oracle on "
                     f"{threads} OpenMP threads in {el:.1f} s, then one frame on 1 thread in {one:.1f} s",
           "one_thread_value": T / one, "one_thread_ms_per_frame": 1e3 * one,
           "ms_per_frame": 1e3 * el / reps, "cpu": info}
    return out, fb, cb


def instances_valu_roofline(stage1_ms: float):
    """VALU roofline of the 3000-instance stage 1: wave64 VALU instructions
    per launch (SQ_INSTS_VALU, from the committed counter summary when it was
    taken on these kernels) over the live stage-1 time, against 1024 SIMDs
    issuing at most two wave64 VALU instructions per quad-cycle (4 clocks)
    at 2.4 GHz; issue_slots_busy is the counter pass's share of quad-cycles
    that issued any."""
    prof = load_profile("stage1_instances_3000x1M_1080p_valu")
    if not prof:
        return None
    peak = prof["simds"] * prof["clock_ghz"] * 1e9 * prof["valu_insts_per_simd_quad_cycle"] / 4.0
    single = prof["simds"] * prof["clock_ghz"] * 1e9 / 4.0
    achieved = prof["valu_insts_per_launch"] / (stage1_ms * 1e-3)
    return {"bound": "valu", "achieved": round(achieved / 1e9, 2), "peak": round(peak / 1e9, 1),
            "unit": "G wave64 VALU instructions/s", "frac": round(achieved / peak, 4),
            "single_issue_peak": round(single / 1e9, 1), "issue_slots_busy": prof["issue_slots_busy"],
            "kernel": prof["kernel"], "valu_insts_per_launch": prof["valu_insts_per_launch"],
            "stage1_ms_mean": round(stage1_ms, 4), "source": prof["source"]}


def bench_instances(ctx, steps: int, warmup: int, count: int = 3000):
    """3000 instances of the 1M-triangle sphere at 1920x1080 (one mesh read
    once, transformed per instance): frame time over `steps` frames, stage
    times from HIP events, and stage 1 against the VALU issue rate from the
    committed counter summary when it was taken on these kernels."""
    import crx
    import workloads as wl
    pos, idx, col = wl.sphere_mesh()
    worlds = list(wl.instance_transforms(count))
    view, proj = wl.instances_view()
    mesh = crx.Mesh(ctx, pos, idx, col)
    scene = crx.Scene(ctx)
    for w in worlds:
        scene.add(mesh, w)
    v = crx.make_view(view, proj, WIDTH, HEIGHT)
    fb, cb = ctx.alloc(8 * WIDTH * HEIGHT), ctx.alloc(8 * WIDTH * HEIGHT)
    settings = crx.render_settings()
    for _ in range(warmup):
        scene.render(v, fb, cb, settings)
    ctx.enable_timing(True)
    s1, pipe = [], []
    for _ in range(min(steps, 60)):
        scene.render(v, fb, cb, settings)
        t = ctx.timings()
        s1.append(t.stage1_ms)
        pipe.append(t.pipeline_ms)
    ctx.enable_timing(False)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        scene.render(v, fb, cb, settings)
    ctx.synchronize()
    el = time.perf_counter() - t0
    T = len(idx) * count
    out = {"workload": f"sphere-1M-x{count}-instances-1080p", "triangle_instances": T,
           "ms_per_frame": 1e3 * el / steps, "triangle_instances_per_s": T * steps / el,
           "stage1_ms_mean": round(float(np.mean(s1)), 4), "pipeline_ms_median": round(float(np.median(pipe)), 4)}
    rl = instances_valu_roofline(out["stage1_ms_mean"])
    if rl:
        out["roofline"] = rl
    for b in (fb, cb):
        b.free()
    scene.close()
    mesh.close()
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="heightfield", choices=["heightfield", "sphere", "instances"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-multiview", action="store_true",
                    help="skip the 64-view batched vs plain measurement on the heightfield")
    ap.add_argument("--multiview-only", action="store_true",
                    help="only the 64-view batched vs plain measurement (heightfield workload)")
    ap.add_argument("--compressed", action="store_true",
                    help="store the mesh IX_PU16-compressed (bit-packed indices, 3 x uint16 positions)")
    ap.add_argument("--serve", action="store_true",
                    help="only the out-of-core serving workload (Poisson requests, p50/p99 latency)")
    ap.add_argument("--serve-n", type=int, default=31623, help="serving terrain grid size (vertices per side)")
    ap.add_argument("--serve-tile", type=int, default=707, help="cells per node side (707 -> ~1M triangles)")
    ap.add_argument("--serve-seconds", type=float, default=10.0)
    ap.add_argument("--no-serve", action="store_true", help="skip the served out-of-core leg")
    ap.add_argument("--no-instances", action="store_true", help="skip the 3000-instance leg")
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--probe-copy", action="store_true",
                    help="also run the 1 GiB copy kernel that calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE")
    ap.add_argument("--probe-valu", action="store_true",
                    help="only the VALU issue-rate probes (crx_probe_valu), as one JSON object")
    ap.add_argument("--no-autotune", action="store_true",
                    help="multi-view: keep the default tuning instead of running crx_autotune first")
    ap.add_argument("--tuning", default="",
                    help="comma-separated crx_tuning overrides, e.g. stage1_order=0,stage1_block=256")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: --gpus {args.gpus} is not supported "
              "(the visibility-buffer path here runs on one GPU; several GPUs would split triangles or views "
              "and min-reduce the visibility buffers)",
              file=sys.stderr)
        return 2

    import crx
    if args.probe_valu:
        ctx = crx.Context(0)
        print(json.dumps({"valu_probe": bench_valu(ctx)}), flush=True)
        ctx.close()
        return 0
    if args.serve:
        ctx = crx.Context(0)
        sv = bench_serving(ctx, args.serve_n, args.serve_tile, args.serve_seconds, args.serve_rate)
        b = sv["batched"]
        print(json.dumps({
            "metric": "served views/sec with p50/p99 latency (out-of-core, Poisson arrivals)",
            "value": b["views_per_s"], "unit": "views/s", "n_gpus": 1, "steps": b["requests"], "warmup": 0,
            "ms_per_step": 1e3 * b["wall_s"] / max(1, b["requests"]), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": sv["workload"], "triangles": sv["triangles"], "pinned_gb": sv["pinned_gb"],
                       "rate_per_s": sv["rate_per_s"], "duration_s": sv["duration_s"]},
            "serving": sv}), flush=True)
        ctx.close()
        return 0
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
    t1 = time.perf_counter()
    mesh = crx.Mesh(ctx, pos, idx, col, compressed=args.compressed)
    if args.compressed:
        desc = dict(desc, workload=desc["workload"] + "-ix_pu16")
        log(f"[bench] IX_PU16 compression + upload in {time.perf_counter() - t1:.1f} s")
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

    if args.multiview_only:
        if args.workload != "heightfield":
            raise SystemExit("--multiview-only needs the heightfield workload")
        mv = bench_multiview(ctx, scene, pos, desc["grid"], args.steps, args.warmup, autotune=not args.no_autotune)
        log(f"[bench] multi-view {mv}")
        print(json.dumps({
            "metric": "batched views/sec (64 mixed-resolution views, full frames)",
            "value": mv["batched_views_per_s"], "unit": "views/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": mv["batched_ms_per_frame"], "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": dict(desc, workload=f"{desc['workload']}-64views"), "batched_multiview": mv}), flush=True)
        ctx.close()
        return 0

    draw_stats = None
    for _ in range(max(1, args.warmup)):
        draw_stats = scene.render(v, fb, cb, settings)
    ctx.synchronize()
    # the stage-1 kernel the draw took (crx_draw_stats::stage1_kernel)
    kernel = {0: ("k_stage1<256, 2, true, false>" if args.compressed else "k_stage1"), 1: "k_stage1_views",
              2: "k_stage1_pairs<256, %s>" % ("true" if args.compressed else "false")}.get(
        draw_stats.stage1_kernel, "k_stage1")

    atomics = bench_atomics(ctx)
    log(f"[bench] atomicMin probe: {atomics}")

    hbm = bench_hbm(ctx)
    log(f"[bench] HBM probe: {hbm}")
    if args.probe_copy:
        nbytes = 1 << 30
        src, dst, sink = ctx.alloc(nbytes), ctx.alloc(nbytes), ctx.alloc(16)
        for _ in range(3):
            ctx.copy(dst, src, nbytes)
            ctx.probe_read(src, nbytes, sink)
        ctx.synchronize()
        for b in (src, dst, sink):
            b.free()
        log(f"[bench] copy probe: 3 x k_copy and 3 x k_read of {nbytes} bytes")

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

    if args.compressed:
        bits = int(np.ceil(np.log2(np.float32(int(idx.max()) - int(idx.min()) + 1))))
        mesh_bytes = (3 * len(idx) * bits + 7) // 8 + 6 * V
    else:
        mesh_bytes = stage1_bytes(len(idx), V)
    algo = mesh_bytes * (len(worlds) if args.workload != "instances" else 1)
    if args.workload == "instances":
        algo = mesh_bytes + 64 * len(worlds)
    achieved = algo / (mean_s1 * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": kernel, "algorithmic_bytes_per_launch": algo,
                "stage1_ms_mean": round(mean_s1, 4)}
    # HBM bytes per launch from rocprofv3 FETCH_SIZE / WRITE_SIZE, used only
    # when they were measured on these kernel sources
    prof = load_profile(f"stage1_{args.workload}{'_ix_pu16' if args.compressed else ''}_traffic")
    if prof:
        roofline["traffic"] = prof["traffic_bytes"]
        roofline["traffic_source"] = prof["source"]
    else:
        roofline["traffic_note"] = "no counter summary measured on these kernel sources"
    if args.workload == "instances":
        roofline["note"] = ("the mesh is read once for all 3000 instances: this stage 1 is bound by the "
                            "per-instance arithmetic (VALU), not by HBM")
        valu = instances_valu_roofline(mean_s1)
        if valu:
            valu["hbm"] = roofline
            roofline = valu
    got_fb = fb.download(np.uint64, WIDTH * HEIGHT)
    got_cb = cb.download(np.uint64, WIDTH * HEIGHT)

    multiview = None
    if args.workload == "heightfield" and not args.no_multiview:
        log("[bench] batched multi-view: 64 seeded views, batched vs plain ...")
        multiview = bench_multiview(ctx, scene, pos, desc["grid"], args.steps, args.warmup,
                                    autotune=not args.no_autotune)
        log(f"[bench] multi-view {multiview}")

    cpu, parity = None, None
    if not args.no_cpu_baseline:
        log(f"[bench] CPU oracle baseline on {cpu_info()['threads']} threads and on 1 thread ...")
        cpu, want_fb, want_cb = cpu_baseline(pos, idx, col, worlds, view, proj, compressed=args.compressed)
        parity = {"fb_equal": bool(np.array_equal(got_fb, want_fb)), "cb_equal": bool(np.array_equal(got_cb, want_cb)),
                  "pixels": WIDTH * HEIGHT, "covered": int((got_fb != crx.CLEAR_PIXEL).sum()),
                  "fb_mismatches": int((got_fb != want_fb).sum()), "cb_mismatches": int((got_cb != want_cb).sum())}
        log(f"[bench] CPU baseline {cpu['value']:.3e} triangles/s ({cpu['sample']}); parity {parity}")
        del want_fb, want_cb

    scene.close()
    mesh.close()
    fb.free()
    cb.free()
    del pos, idx, col

    instances = None
    if args.workload == "heightfield" and not args.no_instances:
        log("[bench] 3000 instances of the 1M-triangle sphere ...")
        instances = bench_instances(ctx, args.steps, args.warmup)
        log(f"[bench] instances {instances}")

    serving = None
    if args.workload == "heightfield" and not args.no_serve:
        log("[bench] served out-of-core 2G-triangle heightfield at 200 requests/s ...")
        serving = bench_serving(ctx, args.serve_n, args.serve_tile, args.serve_seconds, args.serve_rate)

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
        "atomic_min_peak": atomics,
        "hbm_probe": hbm,
        "parity": parity,
        "batched_multiview": multiview,
        "instances": instances,
        "serving": serving,
    }
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
