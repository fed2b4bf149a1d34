"""The continuous-batching server over an out-of-core node store.

A tiled terrain lives in page-locked host memory with a device cache smaller
than the scene, so requests stream nodes in and evict others.  Every result
must be bit-identical whether its request is drawn alone or batched with the
others (tolerance: none), and equal to the oracle's draw of the nodes that
view sees, numbered with the scene-wide triangle prefix.
"""
import numpy as np
import pytest

import crx
import oracle as orc
import workloads as wl
from crxtest import describe_mismatch

pytestmark = pytest.mark.gpu

N, TILE = 513, 64
SIZES = ((128, 128), (256, 144), (200, 200), (96, 160))


@pytest.fixture(scope="module")
def ctx():
    c = crx.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def terrain(ctx):
    nodes, bufs = wl.tiled_heightfield(ctx, N, TILE, amplitude=30.0, scale=128.0, chunk_bytes=1 << 24)
    yield nodes
    for b in bufs:
        b.free()


def _server(ctx, nodes, slots, frame_mb, batch, lanes, **extra):
    slot = max((n["nv"] * 12 + 255) // 256 * 256 + (n["nt"] * 12 + 255) // 256 * 256 + (n["nv"] * 4 + 255) // 256 * 256
               for n in nodes)
    cfg = crx.server_config(node_cache_bytes=slot * slots, frame_bytes=frame_mb << 20, max_batch_views=batch,
                            lanes=lanes, resolve=1, **extra)
    srv = crx.Server(ctx, cfg)
    for k, nd in enumerate(nodes):
        srv.add_node(nd["pos"], nd["nv"], nd["idx"], nd["nt"], nd["col"], nd["aabb"], resident=(k % 2 == 0))
    srv.start()
    return srv


def _requests(nodes, count):
    ground = wl.ground_lookup(nodes, TILE)
    reqs = wl.serving_requests(N, ground, duration=count / 200.0 + 1.0, rate=200.0, sessions=5, sizes=SIZES,
                               seed=0x5E4E, speed=60.0, altitude=(15.0, 40.0))[:count]
    return [crx.make_view(v, p, W, H) for (_, v, p, W, H, _) in reqs]


def _fetch(srv, ticket, view):
    r = srv.wait(ticket)
    assert r.status == 0, crx.lib().crx_last_error()
    n = view.width * view.height
    fb = srv.download(r.framebuffer, n)
    cb = srv.download(r.colorbuffer, n)
    srv.release(ticket)
    return fb, cb, r


def _oracle_view(nodes, ids, v):
    """The oracle's draw and resolve of the nodes `ids` for view v, with the
    scene-wide triangle prefix."""
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
    fb, _ = orc.draw(crx, inst, inst, xf, v.width, v.height, v.view, v.proj, False)
    cb = orc.resolve(crx, inst, prefix[ids] if len(ids) else np.zeros(0, np.uint64), fb, v.width, v.height,
                     v.view, v.proj, crx.DISPLAY_VERTEX_COLORS)
    return fb, cb


def test_alone_vs_mix_and_oracle(ctx, terrain):
    views = _requests(terrain, 24)
    srv = _server(ctx, terrain, slots=20, frame_mb=2, batch=8, lanes=2)
    try:
        alone = []
        for v in views:
            alone.append(_fetch(srv, srv.submit(v), v)[:2])
        tickets = [srv.submit(v) for v in views]      # all at once: continuous batching
        mix, batches = [], []
        for t, v in zip(tickets, views):
            fb, cb, r = _fetch(srv, t, v)
            mix.append((fb, cb))
            batches.append(r.batch_views)
        st = srv.stats()
        print(f"[server] batches {st.batches} max_batch {st.max_batch} uploads {st.uploads} "
              f"evictions {st.evictions} resident {st.resident_nodes}/{st.cache_slots}", flush=True)
        assert max(batches) > 1, "requests submitted together should share batches"
        assert st.uploads > 0 and st.evictions > 0, "the cache is smaller than the scene"
        for k, v in enumerate(views):
            assert np.array_equal(mix[k][0], alone[k][0]), f"request {k}\n" + describe_mismatch(mix[k][0], alone[k][0], v.width)
            assert np.array_equal(mix[k][1], alone[k][1]), f"request {k} colours"
        assert sum(int((fb != crx.CLEAR_PIXEL).sum()) for fb, _ in alone) > 50_000
        # oracle: the nodes the view sees, numbered with the scene-wide prefix
        for k in range(0, len(views), 4):
            want_fb, want_cb = _oracle_view(terrain, srv.view_nodes(views[k]), views[k])
            assert np.array_equal(alone[k][0], want_fb), f"request {k} vs oracle\n" + describe_mismatch(alone[k][0], want_fb, views[k].width)
            assert np.array_equal(alone[k][1], want_cb), f"request {k} colours vs oracle"
    finally:
        srv.close()


def test_admission_waits_for_memory(ctx, terrain):
    """A frame budget that holds only a few results: requests are admitted as
    earlier results are released, and all complete."""
    views = _requests(terrain, 16)
    srv = _server(ctx, terrain, slots=24, frame_mb=1, batch=16, lanes=2)
    try:
        tickets = [srv.submit(v) for v in views]
        for t, v in zip(tickets, views):
            fb, _, r = _fetch(srv, t, v)
            assert r.done_time >= r.start_time >= r.submit_time
        st = srv.stats()
        assert st.requests == len(views)
        assert st.max_batch < len(views)
    finally:
        srv.close()


def test_view_seeing_more_nodes_than_the_cache(ctx, terrain):
    """An overview that sees all 64 nodes with a 16-slot cache: it is drawn
    in passes over groups of nodes into one visibility buffer and resolved
    in passes, each shading the pixels of its group's triangles.  Equal to
    the oracle's one-pass draw of all 64 nodes, and the same alone and
    between other requests on two lanes."""
    srv = _server(ctx, terrain, slots=16, frame_mb=8, batch=4, lanes=2)
    try:
        W, H = 320, 240
        far = crx.make_view(crx.look_at((256.0, 400.0, -200.0), (256.0, 0.0, 256.0)), wl.projection(W, H), W, H)
        ids = srv.view_nodes(far)
        assert len(ids) == len(terrain) == 64
        fb, cb, r = _fetch(srv, srv.submit(far), far)
        st = srv.stats()
        assert st.cache_slots == 16 and st.evictions > 0
        print(f"[server] overview: uploads {r.uploads} evictions {st.evictions}", flush=True)
        want_fb, want_cb = _oracle_view(terrain, ids, far)
        assert np.array_equal(fb, want_fb), describe_mismatch(fb, want_fb, W)
        assert np.array_equal(cb, want_cb), "colours\n" + describe_mismatch(cb, want_cb, W)
        assert int((fb != crx.CLEAR_PIXEL).sum()) > W * H // 4
        # between other requests: the overview and its neighbours as when alone
        views = _requests(terrain, 6)
        alone = [_fetch(srv, srv.submit(v), v)[:2] for v in views]
        mix_views = views[:3] + [far] + views[3:]
        tickets = [srv.submit(v) for v in mix_views]
        got = [_fetch(srv, t, v)[:2] for t, v in zip(tickets, mix_views)]
        want = alone[:3] + [(fb, cb)] + alone[3:]
        for k, v in enumerate(mix_views):
            assert np.array_equal(got[k][0], want[k][0]), f"request {k}\n" + describe_mismatch(got[k][0], want[k][0], v.width)
            assert np.array_equal(got[k][1], want[k][1]), f"request {k} colours"
    finally:
        srv.close()


def test_short_queues_are_not_merged(ctx, terrain):
    """merge_backlog: a backlog shorter than it is drawn one request per
    batch (the views of a batch finish together), a longer one is merged;
    every result equals the request drawn alone."""
    views = _requests(terrain, 12)
    for merge, lane_backlog in ((16, 2), (2, 1)):
        srv = _server(ctx, terrain, slots=40, frame_mb=4, batch=8, lanes=2, merge_backlog=merge,
                      lane_backlog=lane_backlog)
        try:
            alone = [_fetch(srv, srv.submit(v), v)[:2] for v in views]
            tickets = [srv.submit(v) for v in views]
            batches = []
            for k, (t, v) in enumerate(zip(tickets, views)):
                fb, cb, r = _fetch(srv, t, v)
                batches.append(r.batch_views)
                assert np.array_equal(fb, alone[k][0]) and np.array_equal(cb, alone[k][1]), f"request {k}"
            print(f"[server] merge_backlog {merge}: batches {batches}", flush=True)
            if merge > len(views):
                assert max(batches) == 1
            else:
                assert max(batches) > 1
        finally:
            srv.close()
