"""The served out-of-core workload at its stated size: a 31623 x 31623
heightfield (1,999,901,768 triangles) cut into 2025 nodes of ~1M triangles
(707 x 707 cells) in 40 GB of page-locked host memory, half of the nodes
resident at start and a device node cache of half the scene.

* 48 Poisson requests (200 req/s, 8 moving cameras, resolutions from 256^2
  to 3840x2160) are drawn alone, one at a time, and then all submitted at
  once so that they batch: every result is identical in both (tolerance:
  none), and every 12th equals the oracle's draw of the nodes its view sees.
* An overview that sees more nodes than the cache holds is drawn in passes
  over groups of nodes and resolved in passes; it equals the oracle's
  one-pass draw of all those nodes.
"""
import os
import time

import numpy as np
import pytest

import crx
import oracle as orc
import workloads as wl
from crxtest import describe_mismatch

pytestmark = pytest.mark.gpu

N, TILE = 31623, 707
THREADS = int(os.environ.get("OMP_NUM_THREADS", "16"))


def say(msg):
    print(f"[served] {msg}", flush=True)


@pytest.fixture(scope="module")
def ctx():
    c = crx.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def terrain(ctx):
    t0 = time.perf_counter()
    nodes, bufs = wl.tiled_heightfield(ctx, N, TILE, threads=THREADS)
    tris = sum(nd["nt"] for nd in nodes)
    say(f"{len(nodes)} nodes, {tris} triangles in page-locked memory after {time.perf_counter() - t0:.1f} s")
    assert len(nodes) == 2025 and tris == 1_999_901_768
    yield nodes
    for b in bufs:
        b.free()


def _slot_bytes(nodes):
    al = lambda x: (x + 255) // 256 * 256
    return max(al(12 * nd["nv"]) + al(12 * nd["nt"]) + al(4 * nd["nv"]) for nd in nodes)


@pytest.fixture(scope="module")
def server(ctx, terrain):
    cfg = crx.server_config(node_cache_bytes=_slot_bytes(terrain) * (len(terrain) // 2), frame_bytes=4 << 30,
                            max_batch_views=64, lanes=2, resolve=1)
    srv = crx.Server(ctx, cfg)
    for k, nd in enumerate(terrain):
        srv.add_node(nd["pos"], nd["nv"], nd["idx"], nd["nt"], nd["col"], nd["aabb"], resident=(k % 2 == 0))
    t0 = time.perf_counter()
    srv.start()
    say(f"server started in {time.perf_counter() - t0:.1f} s ({srv.stats().cache_slots} cache slots)")
    yield srv
    srv.close()


def _fetch(srv, ticket, view):
    r = srv.wait(ticket)
    assert r.status == 0, crx.lib().crx_last_error()
    n = view.width * view.height
    fb = srv.download(r.framebuffer, n)
    cb = srv.download(r.colorbuffer, n)
    srv.release(ticket)
    return fb, cb, r


def _oracle_view(nodes, ids, v):
    """The oracle's draw and resolve of the nodes `ids` for view v, numbered
    with the scene-wide triangle prefix."""
    prefix = np.concatenate([[0], np.cumsum([nd["nt"] for nd in nodes])[:-1]]).astype(np.uint64)
    inst = []
    for j, i in enumerate(ids):
        nd = nodes[i]
        m = crx.CMesh()
        m.numTriangles = nd["nt"]
        m.positions = nd["pos"]
        m.indices = nd["idx"]
        m.colors = nd["col"]
        m.cummulativeTriangleCount = int(prefix[i])
        m.instances.offset = j
        m.instances.count = 1
        m.world = crx.identity()
        m.id = i
        m.numVertices = nd["nv"]
        m.isLoaded = 1
        inst.append(m)
    xf = [crx.mat4_mul(v.view, crx.identity()) for _ in ids]
    fb, _ = orc.draw(crx, inst, inst, xf, v.width, v.height, v.view, v.proj, False, threads=THREADS)
    cb = orc.resolve(crx, inst, prefix[ids], fb, v.width, v.height, v.view, v.proj, crx.DISPLAY_VERTEX_COLORS,
                     threads=THREADS)
    return fb, cb


def test_poisson_requests_alone_and_batched(server, terrain):
    ground = wl.ground_lookup(terrain, TILE)
    reqs = wl.serving_requests(N, ground, duration=1.0, rate=200.0)[:48]
    assert len(reqs) == 48
    views = [crx.make_view(v, p, w, h) for (_, v, p, w, h, _) in reqs]
    alone = [_fetch(server, server.submit(v), v)[:2] for v in views]
    say(f"48 requests drawn alone; submitting all at once ...")
    tickets = [server.submit(v) for v in views]
    batches = []
    for k, (t, v) in enumerate(zip(tickets, views)):
        fb, cb, r = _fetch(server, t, v)
        batches.append(r.batch_views)
        assert np.array_equal(fb, alone[k][0]), f"request {k}\n" + describe_mismatch(fb, alone[k][0], v.width)
        assert np.array_equal(cb, alone[k][1]), f"request {k} colours"
    st = server.stats()
    say(f"all 48 identical alone and batched (batches of up to {max(batches)} views; "
        f"{st.uploads} uploads, {st.evictions} evictions)")
    assert max(batches) > 1
    assert sum(int((fb != crx.CLEAR_PIXEL).sum()) for fb, _ in alone) > 1_000_000
    for k in range(0, 48, 12):
        v = views[k]
        ids = server.view_nodes(v)
        want_fb, want_cb = _oracle_view(terrain, ids, v)
        assert np.array_equal(alone[k][0], want_fb), f"request {k} vs oracle\n" + describe_mismatch(alone[k][0], want_fb, v.width)
        assert np.array_equal(alone[k][1], want_cb), f"request {k} colours vs oracle"
        say(f"request {k} ({v.width}x{v.height}, {len(ids)} nodes) equals the oracle")


def test_overview_larger_than_the_cache(server, terrain):
    W, H = 1920, 1080
    c = 0.5 * (N - 1)
    v = crx.make_view(crx.look_at((c, 20000.0, -8000.0), (c, 0.0, c)), wl.projection(W, H), W, H)
    ids = server.view_nodes(v)
    slots = server.stats().cache_slots
    say(f"overview sees {len(ids)} nodes with {slots} cache slots; drawing in passes ...")
    assert len(ids) > slots
    t0 = time.perf_counter()
    fb, cb, r = _fetch(server, server.submit(v), v)
    say(f"overview served in {time.perf_counter() - t0:.2f} s ({r.uploads} uploads); oracle ...")
    t0 = time.perf_counter()
    want_fb, want_cb = _oracle_view(terrain, ids, v)
    say(f"oracle overview in {time.perf_counter() - t0:.1f} s")
    assert np.array_equal(fb, want_fb), describe_mismatch(fb, want_fb, W)
    assert np.array_equal(cb, want_cb), "colours\n" + describe_mismatch(cb, want_cb, W)
    assert int((fb != crx.CLEAR_PIXEL).sum()) > W * H // 4
