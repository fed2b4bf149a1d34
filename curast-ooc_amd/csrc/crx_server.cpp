// crx_server.cpp — serving with continuous batching over an out-of-core node
// store (include/crx.h, crx_server_*).
//
// Requests (one view each) queue in arrival order.  Each lane is a host
// thread with its own stream and crx_context (queues, counters, staging); a
// free lane takes the oldest requests whose result buffers fit the frame
// budget, makes the nodes they see resident, and draws them with one
// multi-view pass.  Lanes run concurrently, so batches overlap on the device
// and new requests are admitted as soon as a lane and memory free up.
//
// Node geometry lives in page-locked host memory; the device holds an LRU
// cache of fixed-size slots.  A missing node is copied into a free or evicted
// slot on the upload stream, and the drawing lane's stream waits for that
// copy's event (the async scheme of the reference's point-cloud octree,
// structureUpdate.cpp:256 and SNCOctree.h:102-119, here for triangle nodes).
// Nodes in use by a batch in flight are pinned and never evicted, and a
// slot is refilled only after the kernels that read its previous node: the
// lane that releases a node records its stream's release event, and the
// upload into that slot waits for the release events of the lanes that read
// it.
//
// With prefetch on, the nodes whose boxes touch those of a batch's nodes are
// queued after the batch, and a prefetch thread uploads them on its own
// stream, one at a time, into free slots or slots whose node no batch used
// recently: a camera that moves on finds the nodes coming into view already
// resident, so its requests do not wait for uploads.
//
// A view that sees more nodes than the cache holds is drawn in passes over
// groups of nodes into the same visibility buffer (the 64-bit atomic minimum
// makes the result independent of the grouping), then resolved in passes
// that each shade only the pixels whose triangle lies in the group's nodes.
//
// Triangle ids are global: node i's ids start at the triangle count of the
// nodes before it, and a view's per-node culling (SNTriangles.h:29-105) does
// not depend on the other views in its batch, so a result does not depend
// on how requests were batched.
#include "crx.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "crx_device.h"
#include "crx_host.h"

using namespace crx;

namespace {

struct Node {
    const float* hpos = nullptr;
    const uint32_t* hidx = nullptr;
    const uint32_t* hcol = nullptr;
    uint32_t numVertices = 0, numTriangles = 0;
    crx_box3 aabb{};
    crx_box3 wbox{};
    crx_mat4 world{};
    uint64_t cumm = 0;
    bool wantResident = false;
    // cache state (guarded by the cache mutex)
    int slot = -1;
    uint32_t pins = 0;
    uint64_t lastUse = 0;
    hipEvent_t ready = nullptr;   // recorded after the node's last upload
};

struct Request {
    crx_view view{};
    std::vector<uint32_t> nodes;     // the nodes the view sees (its node cull), ascending
    size_t offset = 0, bytes = 0;
    double submit = 0, start = 0, done = 0;
    uint32_t batchViews = 0, uploads = 0;
    crx_status status = CRX_OK;
    bool finished = false;
};

// first-fit allocator over [0, size) with coalescing (host bookkeeping of a
// device arena)
struct Arena {
    std::map<size_t, size_t> free;   // offset -> size
    void init(size_t size) { free.clear(); if (size) free[0] = size; }
    bool alloc(size_t bytes, size_t* off) {
        for (auto it = free.begin(); it != free.end(); ++it) {
            if (it->second >= bytes) {
                *off = it->first;
                const size_t rest = it->second - bytes;
                const size_t at = it->first + bytes;
                free.erase(it);
                if (rest) free[at] = rest;
                return true;
            }
        }
        return false;
    }
    void release(size_t off, size_t bytes) {
        auto it = free.emplace(off, bytes).first;
        auto next = std::next(it);
        if (next != free.end() && it->first + it->second == next->first) {
            it->second += next->second;
            free.erase(next);
        }
        if (it != free.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first) {
                prev->second += it->second;
                free.erase(it);
            }
        }
    }
};

constexpr size_t kAlign = 256;

}  // namespace

struct crx_server {
    crx_context* ctx = nullptr;
    crx_server_config cfg{};
    std::chrono::steady_clock::time_point t0;

    std::vector<Node> nodes;
    uint64_t totalTriangles = 0;
    bool started = false;
    // the nodes' world boxes and the range hierarchy over them (built at start)
    std::vector<crx_box3> wboxes;
    NodeTree tree;

    // device node cache
    char* cache = nullptr;
    size_t slotBytes = 0;
    uint32_t numSlots = 0;
    std::vector<int> slotNode;      // node in each slot, -1 = free
    std::vector<uint32_t> slotReaders;   // lanes (bit mask) whose reads of the slot's node may be in flight
    hipStream_t upload = nullptr;
    std::mutex cacheMu;
    std::condition_variable cacheCv;
    uint64_t useClock = 0;

    // frame arena and requests
    char* frames = nullptr;
    Arena arena;
    std::mutex mu;
    std::condition_variable workCv, doneCv;
    std::deque<uint64_t> pending;
    std::unordered_map<uint64_t, Request> reqs;
    uint64_t nextTicket = 1;
    bool stopping = false;
    uint32_t drawing = 0;   // lanes between admitting a batch and finishing it

    struct Lane {
        crx_context* ctx = nullptr;
        hipStream_t stream = nullptr;
        hipEvent_t released = nullptr;   // recorded on `stream` whenever the lane releases nodes
        hipEvent_t done = nullptr;       // recorded after a batch's last launch; the lane waits on it
        uint32_t index = 0;
        std::thread thread;
    };
    std::vector<Lane> lanes;
    hipStream_t sharedStream = nullptr;   // cfg.shared_stream: the one stream of all lanes

    // prefetch of the neighbours of drawn nodes (cfg.prefetch): queue and
    // flags guarded by cacheMu
    std::vector<std::vector<uint32_t>> neighbors;
    hipStream_t prefetchStream = nullptr;
    std::thread prefetcher;
    std::deque<uint32_t> prefetchQueue;
    std::vector<uint8_t> prefetchQueued;
    std::condition_variable prefetchCv;
    bool prefetchStop = false;

    // statistics (guarded by mu, except the cache counters: cacheMu)
    uint64_t statRequests = 0, statBatches = 0, statUploads = 0, statUploadBytes = 0, statEvictions = 0;
    uint64_t statPrefetches = 0;
    uint32_t statMaxBatch = 0;

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace {

size_t node_bytes(const Node& n) {
    return align_up(12ull * n.numVertices, kAlign) + align_up(12ull * n.numTriangles, kAlign) +
           align_up(4ull * n.numVertices, kAlign);
}

// device pointers of a node in its slot: positions | indices | colors
void slot_pointers(crx_server* s, const Node& n, float** pos, uint32_t** idx, uint32_t** col) {
    char* base = s->cache + (size_t)n.slot * s->slotBytes;
    *pos = reinterpret_cast<float*>(base);
    *idx = reinterpret_cast<uint32_t*>(base + align_up(12ull * n.numVertices, kAlign));
    *col = reinterpret_cast<uint32_t*>(base + align_up(12ull * n.numVertices, kAlign) +
                                       align_up(12ull * n.numTriangles, kAlign));
}

// Copies node `id` into `slot` on `stream` (the upload stream by default)
// and records its event.  Caller holds cacheMu.
crx_status upload_node(crx_server* s, uint32_t id, int slot, hipStream_t stream = nullptr) {
    if (!stream) stream = s->upload;
    Node& n = s->nodes[id];
    n.slot = slot;
    s->slotNode[slot] = (int)id;
    float* pos;
    uint32_t *idx, *col;
    slot_pointers(s, n, &pos, &idx, &col);
    // the copy waits for the kernels that read the slot's previous node
    for (uint32_t l = 0; l < s->lanes.size(); l++) {
        if (s->slotReaders[slot] & (1u << l)) CRX_HIP(hipStreamWaitEvent(stream, s->lanes[l].released, 0));
    }
    s->slotReaders[slot] = 0;
    if (n.numVertices) CRX_HIP(hipMemcpyAsync(pos, n.hpos, 12ull * n.numVertices, hipMemcpyHostToDevice, stream));
    if (n.hidx && n.numTriangles)
        CRX_HIP(hipMemcpyAsync(idx, n.hidx, 12ull * n.numTriangles, hipMemcpyHostToDevice, stream));
    if (n.hcol && n.numVertices)
        CRX_HIP(hipMemcpyAsync(col, n.hcol, 4ull * n.numVertices, hipMemcpyHostToDevice, stream));
    if (!n.ready) CRX_HIP(hipEventCreateWithFlags(&n.ready, hipEventDisableTiming));
    CRX_HIP(hipEventRecord(n.ready, stream));
    s->statUploads++;
    s->statUploadBytes += node_bytes(n);
    return CRX_OK;
}

// Makes every node of `need` resident and pins it; returns the number of
// uploads issued.  All or nothing: it waits, holding no pins, until the free
// and evictable slots cover the missing nodes, so two lanes can never block
// each other with partial sets.
crx_status acquire_nodes(crx_server* s, const std::vector<uint32_t>& need, uint32_t* uploads) {
    *uploads = 0;
    if (need.size() > s->numSlots)
        return fail(CRX_REJECTED, "a batch sees " + std::to_string(need.size()) + " nodes but the node cache has " +
                                      std::to_string(s->numSlots) + " slots");
    std::unique_lock<std::mutex> lk(s->cacheMu);
    std::vector<uint8_t> inNeed(s->nodes.size(), 0);
    for (uint32_t id : need) inNeed[id] = 1;
    for (;;) {
        uint32_t missing = 0, available = 0;
        for (uint32_t id : need) missing += s->nodes[id].slot < 0 ? 1u : 0u;
        for (uint32_t k = 0; k < s->numSlots; k++) {
            const int nd = s->slotNode[k];
            if (nd < 0 || (s->nodes[nd].pins == 0 && !inNeed[nd])) available++;
        }
        if (available >= missing) break;
        s->cacheCv.wait(lk);   // slots pinned by batches in flight
    }
    const uint64_t stamp = ++s->useClock;
    for (uint32_t id : need) {
        Node& n = s->nodes[id];
        if (n.slot >= 0) {
            n.pins++;
            n.lastUse = stamp;
        }
    }
    for (uint32_t id : need) {
        Node& n = s->nodes[id];
        if (n.slot >= 0) continue;
        int slot = -1;
        for (uint32_t k = 0; k < s->numSlots && slot < 0; k++)
            if (s->slotNode[k] < 0) slot = (int)k;
        if (slot < 0) {   // evict the least recently used unpinned node
            uint64_t best = UINT64_MAX;
            for (uint32_t k = 0; k < s->numSlots; k++) {
                const Node& c = s->nodes[s->slotNode[k]];
                if (c.pins == 0 && c.lastUse < best) { best = c.lastUse; slot = (int)k; }
            }
            s->nodes[s->slotNode[slot]].slot = -1;
            s->slotNode[slot] = -1;
            s->statEvictions++;
        }
        CRX_TRY(upload_node(s, id, slot));
        n.pins++;
        n.lastUse = stamp;
        (*uploads)++;
    }
    return CRX_OK;
}

// Unpins `need` after the lane has enqueued every kernel that reads those
// nodes: the lane's release event marks that point of its stream, and an
// upload into one of their slots waits for it.  Runs on every path, errors
// included.
crx_status release_nodes(crx_server* s, crx_server::Lane& lane, const std::vector<uint32_t>& need) {
    crx_status st = CRX_OK;
    {
        std::lock_guard<std::mutex> lk(s->cacheMu);
        hipError_t e = hipEventRecord(lane.released, lane.stream);
        if (e != hipSuccess) {
            // without the event the slots cannot be ordered after the reads
            st = fail(CRX_HIP_ERROR, std::string("hipEventRecord: ") + hipGetErrorString(e));
            (void)hipStreamSynchronize(lane.stream);
        }
        for (uint32_t id : need) {
            Node& n = s->nodes[id];
            n.pins--;
            if (e == hipSuccess) s->slotReaders[n.slot] |= 1u << lane.index;
        }
    }
    s->cacheCv.notify_all();
    return st;
}

// Batches a node may go unused and still not be evicted by a prefetch.
constexpr uint64_t kPrefetchKeep = 64;

// Queues the non-resident neighbours of the nodes a batch drew for the
// prefetch thread.
void note_drawn(crx_server* s, const std::vector<uint32_t>& ids) {
    if (!s->cfg.prefetch) return;
    bool any = false;
    {
        std::lock_guard<std::mutex> lk(s->cacheMu);
        for (uint32_t id : ids) {
            for (uint32_t nb : s->neighbors[id]) {
                if (s->nodes[nb].slot >= 0 || s->prefetchQueued[nb]) continue;
                s->prefetchQueued[nb] = 1;
                s->prefetchQueue.push_back(nb);
                any = true;
            }
        }
    }
    if (any) s->prefetchCv.notify_one();
}

// Uploads queued nodes one at a time on the prefetch stream, into a free
// slot or the slot of the least recently used unpinned node that no batch
// used in the last kPrefetchKeep batches; a node with no such slot is dropped.
void prefetch_main(crx_server* s) {
    (void)hipSetDevice(s->ctx->device);
    for (;;) {
        hipEvent_t ready = nullptr;
        {
            std::unique_lock<std::mutex> lk(s->cacheMu);
            s->prefetchCv.wait(lk, [&] { return s->prefetchStop || !s->prefetchQueue.empty(); });
            if (s->prefetchStop) return;
            const uint32_t id = s->prefetchQueue.front();
            s->prefetchQueue.pop_front();
            s->prefetchQueued[id] = 0;
            Node& n = s->nodes[id];
            if (n.slot >= 0) continue;
            int slot = -1;
            for (uint32_t k = 0; k < s->numSlots && slot < 0; k++)
                if (s->slotNode[k] < 0) slot = (int)k;
            if (slot < 0) {
                uint64_t best = UINT64_MAX;
                for (uint32_t k = 0; k < s->numSlots; k++) {
                    const Node& c = s->nodes[s->slotNode[k]];
                    if (c.pins == 0 && c.lastUse + kPrefetchKeep <= s->useClock && c.lastUse < best) {
                        best = c.lastUse;
                        slot = (int)k;
                    }
                }
                if (slot < 0) continue;
                s->nodes[s->slotNode[slot]].slot = -1;
                s->slotNode[slot] = -1;
                s->statEvictions++;
            }
            if (upload_node(s, id, slot, s->prefetchStream) != CRX_OK) {
                // the slot's contents are unknown: leave it free
                n.slot = -1;
                s->slotNode[slot] = -1;
                continue;
            }
            n.lastUse = s->useClock;
            s->statPrefetches++;
            ready = n.ready;
        }
        (void)hipEventSynchronize(ready);
    }
}

// Nodes whose world boxes touch (closed boxes that intersect), by a sweep
// over the boxes sorted by their lower x.
void find_neighbors(crx_server* s) {
    const uint32_t N = (uint32_t)s->nodes.size();
    s->neighbors.assign(N, {});
    std::vector<uint32_t> order(N);
    for (uint32_t i = 0; i < N; i++) order[i] = i;
    std::sort(order.begin(), order.end(),
              [&](uint32_t a, uint32_t b) { return s->nodes[a].wbox.min.x < s->nodes[b].wbox.min.x; });
    for (uint32_t a = 0; a < N; a++) {
        const crx_box3& A = s->nodes[order[a]].wbox;
        for (uint32_t b = a + 1; b < N; b++) {
            const crx_box3& B = s->nodes[order[b]].wbox;
            if (B.min.x > A.max.x) break;
            if (B.min.y <= A.max.y && A.min.y <= B.max.y && B.min.z <= A.max.z && A.min.z <= B.max.z) {
                s->neighbors[order[a]].push_back(order[b]);
                s->neighbors[order[b]].push_back(order[a]);
            }
        }
    }
}

// CMesh of node `id` in its slot (scene-wide triangle ids), listed at `k`.
crx_cmesh node_cmesh(crx_server* s, uint32_t id, uint32_t k) {
    const Node& n = s->nodes[id];
    float* pos;
    uint32_t *idx, *col;
    slot_pointers(s, n, &pos, &idx, &col);
    crx_cmesh m;
    std::memset(&m, 0, sizeof(m));
    m.numTriangles = n.numTriangles;
    m.indices = n.hidx ? idx : nullptr;
    m.positions = reinterpret_cast<crx_vec3*>(pos);
    m.colors = n.hcol ? col : nullptr;
    m.cummulativeTriangleCount = n.cumm;
    m.aabb = n.aabb;
    m.instances.offset = (int32_t)k;
    m.instances.count = 1;
    m.world = n.world;
    m.id = (int32_t)id;
    m.numVertices = n.numVertices;
    m.isLoaded = 1;
    m.flipTriangles = mirror_flag(n.world) ? 1 : 0;
    return m;
}

// CMesh records of the pinned nodes `ids`, and `stream` ordered after their uploads.
crx_status pinned_cmeshes(crx_server* s, const std::vector<uint32_t>& ids, hipStream_t stream,
                          std::vector<crx_cmesh>* inst) {
    std::lock_guard<std::mutex> lk(s->cacheMu);   // slot pointers are stable while pinned
    inst->resize(ids.size());
    for (uint32_t k = 0; k < ids.size(); k++) {
        const Node& n = s->nodes[ids[k]];
        (*inst)[k] = node_cmesh(s, ids[k], k);
        if (n.ready) CRX_HIP(hipStreamWaitEvent(stream, n.ready, 0));
    }
    return CRX_OK;
}

// Waits for the lane's last batch: an event after its last launch, so a
// lane sharing its stream does not wait for the batches other lanes
// enqueued after it.
crx_status lane_finish(crx_server::Lane& lane) {
    CRX_HIP(hipEventRecord(lane.done, lane.stream));
    CRX_HIP(hipEventSynchronize(lane.done));
    return CRX_OK;
}

// Draws a group of requests on a lane: one pass when the nodes the group
// sees fit the cache, else (a single view that sees more nodes than the
// cache holds) passes over groups of nodes.  Returns the status for all of them.
crx_status draw_group(crx_server* s, crx_server::Lane& lane, const std::vector<Request*>& batch, uint32_t* uploads) {
    const uint32_t V = (uint32_t)batch.size();
    const uint32_t N = (uint32_t)s->nodes.size();
    *uploads = 0;
    // per-view node cull (SNTriangles::update for each view, done by draw_batch)
    std::vector<uint8_t> vis((size_t)V * N, 0);
    std::vector<uint8_t> any(N, 0);
    for (uint32_t v = 0; v < V; v++) {
        for (uint32_t i : batch[v]->nodes) {
            vis[(size_t)v * N + i] = 1;
            any[i] = 1;
        }
    }
    std::vector<uint32_t> need;
    for (uint32_t i = 0; i < N; i++)
        if (any[i]) need.push_back(i);

    std::vector<crx_render_target> targets(V);
    std::vector<uint64_t*> bufs;
    std::vector<uint32_t> px;
    for (uint32_t v = 0; v < V; v++) {
        const crx_view& cv = batch[v]->view;
        crx_render_target& t = targets[v];
        std::memset(&t, 0, sizeof(t));
        const size_t pixels = (size_t)cv.width * (size_t)cv.height;
        t.framebuffer = reinterpret_cast<uint64_t*>(s->frames + batch[v]->offset);
        t.colorbuffer = s->cfg.resolve ? reinterpret_cast<uint64_t*>(s->frames + batch[v]->offset +
                                                                     align_up(8 * pixels, kAlign))
                                       : nullptr;
        t.width = cv.width;
        t.height = cv.height;
        t.view = cv.view;
        t.viewI = mat_inverse(cv.view);
        t.proj = cv.proj;
        t.f = cv.proj.m[5];
        t.aspect = (float)cv.width / (float)cv.height;
        bufs.push_back(t.framebuffer);
        px.push_back((uint32_t)pixels);
        if (t.colorbuffer) {
            bufs.push_back(t.colorbuffer);
            px.push_back((uint32_t)pixels);
        }
    }
    // view-major transforms and per-view node culls of the nodes `ids`
    std::vector<crx_mat4> xf;
    std::vector<uint8_t> skip;
    const auto lists = [&](const std::vector<uint32_t>& ids) -> bool {
        const uint32_t U = (uint32_t)ids.size();
        xf.assign((size_t)V * U, crx_mat4{});
        skip.assign((size_t)V * U, 0);
        bool anySkip = false;
        for (uint32_t v = 0; v < V; v++) {
            for (uint32_t k = 0; k < U; k++) {
                xf[(size_t)v * U + k] = mat_mul(batch[v]->view.view, s->nodes[ids[k]].world);
                if (!vis[(size_t)v * N + ids[k]]) {
                    skip[(size_t)v * U + k] = 1;
                    anySkip = true;
                }
            }
        }
        return anySkip && V > 1;
    };
    std::vector<crx_cmesh> inst;

    if (need.size() <= s->numSlots) {
        uint32_t up = 0;
        CRX_TRY(acquire_nodes(s, need, &up));
        *uploads += up;
        crx_status st = CRX_OK;
        do {
            st = pinned_cmeshes(s, need, lane.stream, &inst);
            if (st != CRX_OK) break;
            const uint32_t U = (uint32_t)need.size();
            const bool useSkip = lists(need);
            st = clear_buffers(lane.ctx, bufs.data(), px.data(), (uint32_t)bufs.size(), lane.stream);
            if (st != CRX_OK) break;
            crx_draw_stats ds;
            // the resolve reads the instance records and their triangle prefix
            // (cummulativeTriangleCount) from the draw's metadata
            const std::function<crx_status()> resolve = [&]() {
                const DrawPlan& D = lane.ctx->plan;
                const char* meta = static_cast<const char*>(lane.ctx->drawMeta.ptr);
                return resolve_views_device(lane.ctx, U ? reinterpret_cast<const InstanceRec*>(meta + D.offInst) : nullptr,
                                            U ? reinterpret_cast<const uint64_t*>(meta + D.offPrefix) : nullptr, U,
                                            targets.data(), V, CRX_DISPLAY_VERTEX_COLORS, lane.stream);
            };
            DrawExtras ex;
            ex.skip = useSkip ? skip.data() : nullptr;
            ex.afterStages = s->cfg.resolve ? &resolve : nullptr;
            ex.prefix = s->cfg.resolve != 0;
            ex.earlyReturn = true;   // lane_finish waits for the resolve
            st = draw_views(lane.ctx, inst.data(), U, inst.data(), U, xf.data(), targets.data(), V, 0, lane.stream, &ds,
                            ex);
        } while (false);
        const crx_status rs = release_nodes(s, lane, need);
        if (st == CRX_OK) st = rs;
        note_drawn(s, need);
        if (st == CRX_OK) st = lane_finish(lane);
        return st;
    }

    // Passes over groups of at most a lane's share of the cache.
    const size_t group = std::max<size_t>(1, s->numSlots / std::max<size_t>(1, s->lanes.size()));
    std::vector<std::vector<uint32_t>> groups;
    for (size_t g0 = 0; g0 < need.size(); g0 += group)
        groups.emplace_back(need.begin() + g0, need.begin() + std::min(need.size(), g0 + group));
    CRX_TRY(clear_buffers(lane.ctx, bufs.data(), px.data(), (uint32_t)bufs.size(), lane.stream));
    const auto pass = [&](const std::vector<uint32_t>& ids, bool resolvePass) -> crx_status {
        uint32_t up = 0;
        CRX_TRY(acquire_nodes(s, ids, &up));
        *uploads += up;
        crx_status st = pinned_cmeshes(s, ids, lane.stream, &inst);
        if (st == CRX_OK) {
            const uint32_t U = (uint32_t)ids.size();
            if (!resolvePass) {
                const bool useSkip = lists(ids);
                DrawExtras ex;
                ex.skip = useSkip ? skip.data() : nullptr;
                ex.earlyReturn = true;
                crx_draw_stats ds;
                st = draw_views(lane.ctx, inst.data(), U, inst.data(), U, xf.data(), targets.data(), V, 0, lane.stream,
                                &ds, ex);
            } else {
                std::vector<uint64_t> prefix(U);
                for (uint32_t k = 0; k < U; k++) prefix[k] = s->nodes[ids[k]].cumm;
                st = resolve_views(lane.ctx, inst.data(), U, prefix.data(), targets.data(), V,
                                   CRX_DISPLAY_VERTEX_COLORS, lane.stream, true);
            }
        }
        const crx_status rs = release_nodes(s, lane, ids);
        return st != CRX_OK ? st : rs;
    };
    for (const auto& ids : groups) CRX_TRY(pass(ids, false));
    // every pixel's triangle is final: shade each group's pixels, the last
    // drawn group (still resident) first
    if (s->cfg.resolve) {
        for (size_t g = groups.size(); g-- > 0;) CRX_TRY(pass(groups[g], true));
    }
    return lane_finish(lane);
}

// Draws one batch: views are grouped greedily (in order) so that the nodes
// each group sees fit the cache together; a view that alone sees more nodes
// than the cache holds is drawn on its own, in passes.  Sets every request's
// status.
void draw_batch(crx_server* s, crx_server::Lane& lane, const std::vector<Request*>& batch, uint32_t* uploads) {
    const uint32_t N = (uint32_t)s->nodes.size();
    *uploads = 0;
    std::vector<uint8_t> inGroup(N, 0);
    uint32_t groupNodes = 0;
    std::vector<Request*> group;
    auto flush = [&]() {
        if (group.empty()) return;
        uint32_t up = 0;
        const crx_status st = draw_group(s, lane, group, &up);
        *uploads += up;
        for (Request* r : group) r->status = st;
        group.clear();
        std::fill(inGroup.begin(), inGroup.end(), 0);
        groupNodes = 0;
    };
    for (Request* r : batch) {
        float planes[16];
        frustum_planes(r->view, planes);
        std::vector<uint32_t>& mine = r->nodes;
        s->tree.visible(planes, s->wboxes.data(), mine);
        if (mine.size() > s->numSlots) {
            flush();
            group.push_back(r);
            flush();
            continue;
        }
        uint32_t extra = 0;
        for (uint32_t i : mine) extra += inGroup[i] ? 0u : 1u;
        if (groupNodes + extra > s->numSlots) flush();
        for (uint32_t i : mine) {
            if (!inGroup[i]) { inGroup[i] = 1; groupNodes++; }
        }
        group.push_back(r);
    }
    flush();
}

size_t request_bytes(const crx_server* s, const crx_view& v) {
    const size_t pixels = (size_t)v.width * (size_t)v.height;
    return align_up(8 * pixels, kAlign) * (s->cfg.resolve ? 2 : 1);
}

void lane_main(crx_server* s, crx_server::Lane* lane) {
    (void)hipSetDevice(s->ctx->device);
    for (;;) {
        std::vector<uint64_t> tickets;
        std::vector<Request*> batch;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            for (;;) {
                // admit the oldest requests whose buffers fit (FIFO: stop at the first that does not);
                // with every lane idle the first lane takes the work (so light traffic keeps to one
                // lane, whose context keeps the node lists of the last frame); while another lane
                // draws, only a backlog of lane_backlog requests starts this one (a lone request
                // waits for the drawing lane instead of sharing the device with it); several
                // requests share one batch only from a backlog of merge_backlog (the views of one
                // batch finish together, so merging a short queue delays its head)
                const bool mayStart = s->stopping || (s->drawing == 0 ? lane->index == 0
                                                                        : s->pending.size() >= s->cfg.lane_backlog);
                const uint32_t cap = s->pending.size() >= s->cfg.merge_backlog ? s->cfg.max_batch_views : 1u;
                while (mayStart && !s->pending.empty() && batch.size() < cap) {
                    Request& r = s->reqs[s->pending.front()];
                    size_t off;
                    if (!s->arena.alloc(r.bytes, &off)) break;
                    r.offset = off;
                    tickets.push_back(s->pending.front());
                    batch.push_back(&r);
                    s->pending.pop_front();
                }
                if (!batch.empty()) break;
                if (s->stopping) {
                    // nothing admissible will free up any more: fail what is left
                    for (uint64_t t : s->pending) {
                        Request& r = s->reqs[t];
                        r.status = CRX_REJECTED;
                        r.finished = true;
                    }
                    s->pending.clear();
                    s->doneCv.notify_all();
                    return;
                }
                s->workCv.wait(lk);
            }
            const double t = s->now();
            for (Request* r : batch) r->start = t;
            s->drawing++;
        }
        uint32_t uploads = 0;
        draw_batch(s, *lane, batch, &uploads);
        {
            std::lock_guard<std::mutex> lk(s->mu);
            const double t = s->now();
            for (Request* r : batch) {
                r->done = t;
                r->batchViews = (uint32_t)batch.size();
                r->uploads = uploads;
                r->finished = true;
            }
            s->drawing--;
            s->statRequests += batch.size();
            s->statBatches++;
            s->statMaxBatch = std::max<uint32_t>(s->statMaxBatch, (uint32_t)batch.size());
        }
        s->doneCv.notify_all();
        s->workCv.notify_all();   // other lanes may admit what this lane left queued
    }
}

}  // namespace

extern "C" {

void crx_server_config_default(crx_server_config* out) {
    if (!out) return;
    out->node_cache_bytes = 8ull << 30;
    out->frame_bytes = 4ull << 30;
    out->max_batch_views = 64;
    out->lanes = 2;
    out->resolve = 1;
    out->prefetch = 1;
    out->shared_stream = 0;
    out->lane_backlog = 4;
    out->merge_backlog = 4;
    out->reserved = 0;
}

crx_status crx_host_alloc(crx_context* ctx, size_t bytes, void** out) {
    if (!ctx || !out) return fail(CRX_INVALID_ARG, "NULL argument");
    *out = nullptr;
    CRX_HIP(hipSetDevice(ctx->device));
    hipError_t e = hipHostMalloc(out, bytes ? bytes : 1, hipHostMallocDefault);
    if (e != hipSuccess) {
        *out = nullptr;
        return fail(CRX_OUT_OF_MEMORY, "hipHostMalloc(" + std::to_string(bytes) + "): " + hipGetErrorString(e));
    }
    return CRX_OK;
}

crx_status crx_host_free(crx_context* ctx, void* ptr) {
    if (!ctx) return fail(CRX_INVALID_ARG, "ctx is NULL");
    if (ptr) CRX_HIP(hipHostFree(ptr));
    return CRX_OK;
}

crx_status crx_server_create(crx_context* ctx, const crx_server_config* config, crx_server** out) {
    if (!ctx || !out) return fail(CRX_INVALID_ARG, "NULL argument");
    *out = nullptr;
    crx_server_config cfg;
    crx_server_config_default(&cfg);
    if (config) cfg = *config;
    if (cfg.max_batch_views == 0 || cfg.max_batch_views > kMaxViews)
        return fail(CRX_INVALID_ARG, "max_batch_views must be in [1, 256]");
    if (cfg.lanes == 0 || cfg.lanes > 8) return fail(CRX_INVALID_ARG, "lanes must be in [1, 8]");
    if (cfg.lane_backlog == 0 || cfg.merge_backlog == 0)
        return fail(CRX_INVALID_ARG, "lane_backlog and merge_backlog must be >= 1");
    if (cfg.frame_bytes == 0 || cfg.node_cache_bytes == 0) return fail(CRX_INVALID_ARG, "memory budgets must be > 0");
    crx_server* s = new crx_server();
    s->ctx = ctx;
    s->cfg = cfg;
    s->t0 = std::chrono::steady_clock::now();
    *out = s;
    return CRX_OK;
}

crx_status crx_server_add_node(crx_server* s, const float* positions, uint32_t numVertices, const uint32_t* indices,
                               uint32_t numTriangles, const uint32_t* colors, const crx_box3* aabb,
                               const crx_mat4* world, int resident, uint32_t* node_id) {
    if (!s || !aabb || !world) return fail(CRX_INVALID_ARG, "NULL argument");
    if (s->started) return fail(CRX_INVALID_ARG, "nodes cannot be added after crx_server_start");
    if (numVertices > 0 && !positions) return fail(CRX_INVALID_ARG, "positions is NULL");
    if (numTriangles > 1431655765u) return fail(CRX_INVALID_ARG, "more than 1431655765 triangles");
    if (!indices && (uint64_t)numVertices < 3ull * numTriangles)
        return fail(CRX_INVALID_ARG, "non-indexed node needs 3*numTriangles vertices");
    Node n;
    n.hpos = positions;
    n.hidx = indices;
    n.hcol = colors;
    n.numVertices = numVertices;
    n.numTriangles = numTriangles;
    n.aabb = *aabb;
    n.world = *world;
    n.wbox = world_box(*aabb, *world);
    n.cumm = s->totalTriangles;
    n.wantResident = resident != 0;
    s->totalTriangles += numTriangles;
    s->nodes.push_back(n);
    if (node_id) *node_id = (uint32_t)(s->nodes.size() - 1);
    return CRX_OK;
}

crx_status crx_server_start(crx_server* s) {
    if (!s) return fail(CRX_INVALID_ARG, "server is NULL");
    if (s->started) return fail(CRX_INVALID_ARG, "server already started");
    if (s->totalTriangles >= (1ull << 36)) return fail(CRX_INVALID_ARG, "more than 2^36 triangles in the scene");
    CRX_HIP(hipSetDevice(s->ctx->device));
    s->wboxes.resize(s->nodes.size());
    for (size_t i = 0; i < s->nodes.size(); i++) s->wboxes[i] = s->nodes[i].wbox;
    s->tree.build(s->wboxes.data(), (uint32_t)s->wboxes.size());
    size_t slot = kAlign;
    for (const Node& n : s->nodes) slot = std::max(slot, node_bytes(n));
    s->slotBytes = slot;
    s->numSlots = (uint32_t)std::min<uint64_t>(s->cfg.node_cache_bytes / slot, s->nodes.size());
    if (s->numSlots == 0 && !s->nodes.empty())
        return fail(CRX_INVALID_ARG, "node_cache_bytes holds no node (a slot takes " + std::to_string(slot) + " bytes)");
    if (s->numSlots) {
        hipError_t e = hipMalloc(&s->cache, (size_t)s->numSlots * slot);
        if (e != hipSuccess) { s->cache = nullptr; return fail(CRX_OUT_OF_MEMORY, std::string("node cache: ") + hipGetErrorString(e)); }
    }
    s->slotNode.assign(s->numSlots, -1);
    s->slotReaders.assign(s->numSlots, 0u);
    hipError_t e = hipMalloc(&s->frames, s->cfg.frame_bytes);
    if (e != hipSuccess) { s->frames = nullptr; return fail(CRX_OUT_OF_MEMORY, std::string("frame arena: ") + hipGetErrorString(e)); }
    s->arena.init(s->cfg.frame_bytes);
    CRX_HIP(hipStreamCreateWithFlags(&s->upload, hipStreamNonBlocking));
    {
        std::lock_guard<std::mutex> lk(s->cacheMu);
        uint32_t next = 0;
        for (uint32_t i = 0; i < s->nodes.size() && next < s->numSlots; i++) {
            if (!s->nodes[i].wantResident) continue;
            CRX_TRY(upload_node(s, i, (int)next++));
        }
        s->statUploads = 0;   // initial placement is not streaming
        s->statUploadBytes = 0;
    }
    CRX_HIP(hipStreamSynchronize(s->upload));
    s->lanes.resize(s->cfg.lanes);
    for (uint32_t l = 0; l < s->lanes.size(); l++) {
        auto& lane = s->lanes[l];
        lane.index = l;
        CRX_TRY(crx_context_create(s->ctx->device, &s->ctx->tuning, &lane.ctx));
        if (s->cfg.shared_stream) {
            if (!s->sharedStream) CRX_HIP(hipStreamCreateWithFlags(&s->sharedStream, hipStreamNonBlocking));
            lane.stream = s->sharedStream;
        } else {
            CRX_HIP(hipStreamCreateWithFlags(&lane.stream, hipStreamNonBlocking));
        }
        CRX_HIP(hipEventCreateWithFlags(&lane.released, hipEventDisableTiming));
        CRX_HIP(hipEventCreateWithFlags(&lane.done, hipEventDisableTiming));
    }
    s->started = true;
    if (s->cfg.prefetch) {
        find_neighbors(s);
        s->prefetchQueued.assign(s->nodes.size(), 0);
        CRX_HIP(hipStreamCreateWithFlags(&s->prefetchStream, hipStreamNonBlocking));
        s->prefetcher = std::thread(prefetch_main, s);
    }
    for (auto& lane : s->lanes) lane.thread = std::thread(lane_main, s, &lane);
    return CRX_OK;
}

crx_status crx_server_destroy(crx_server* s) {
    if (!s) return CRX_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stopping = true;
    }
    s->workCv.notify_all();
    for (auto& lane : s->lanes)
        if (lane.thread.joinable()) lane.thread.join();
    {
        std::lock_guard<std::mutex> lk(s->cacheMu);
        s->prefetchStop = true;
    }
    s->prefetchCv.notify_all();
    if (s->prefetcher.joinable()) s->prefetcher.join();
    (void)hipSetDevice(s->ctx->device);
    for (auto& lane : s->lanes) {
        if (lane.stream) (void)hipStreamSynchronize(lane.stream);
        if (lane.stream && lane.stream != s->sharedStream) (void)hipStreamDestroy(lane.stream);
        if (lane.released) (void)hipEventDestroy(lane.released);
        if (lane.done) (void)hipEventDestroy(lane.done);
        if (lane.ctx) crx_context_destroy(lane.ctx);
    }
    if (s->upload) {
        (void)hipStreamSynchronize(s->upload);
        (void)hipStreamDestroy(s->upload);
    }
    if (s->sharedStream) (void)hipStreamDestroy(s->sharedStream);
    if (s->prefetchStream) {
        (void)hipStreamSynchronize(s->prefetchStream);
        (void)hipStreamDestroy(s->prefetchStream);
    }
    for (Node& n : s->nodes)
        if (n.ready) (void)hipEventDestroy(n.ready);
    if (s->cache) (void)hipFree(s->cache);
    if (s->frames) (void)hipFree(s->frames);
    delete s;
    return CRX_OK;
}

crx_status crx_server_submit(crx_server* s, const crx_view* view, uint64_t* ticket) {
    if (!s || !view || !ticket) return fail(CRX_INVALID_ARG, "NULL argument");
    if (!s->started) return fail(CRX_INVALID_ARG, "crx_server_start has not been called");
    if (view->width <= 0 || view->height <= 0 || (int64_t)view->width * view->height > 0x7fffffffll)
        return fail(CRX_INVALID_ARG, "view size out of range");
    const size_t bytes = request_bytes(s, *view);
    if (bytes > s->cfg.frame_bytes) return fail(CRX_REJECTED, "request buffers exceed frame_bytes");
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stopping) return fail(CRX_REJECTED, "server is stopping");
        const uint64_t t = s->nextTicket++;
        Request& r = s->reqs[t];
        r.view = *view;
        r.bytes = bytes;
        r.submit = s->now();
        s->pending.push_back(t);
        *ticket = t;
    }
    s->workCv.notify_all();   // the lane allowed to start may not be the one woken first
    return CRX_OK;
}

crx_status crx_server_wait(crx_server* s, uint64_t ticket, crx_request_result* out) {
    if (!s || !out) return fail(CRX_INVALID_ARG, "NULL argument");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return fail(CRX_INVALID_ARG, "unknown ticket");
    const Request* rp = &it->second;   // element addresses survive rehashing
    s->doneCv.wait(lk, [&] { return rp->finished; });
    const Request& r = *rp;
    const size_t pixels = (size_t)r.view.width * (size_t)r.view.height;
    out->framebuffer = reinterpret_cast<const uint64_t*>(s->frames + r.offset);
    out->colorbuffer = s->cfg.resolve ? reinterpret_cast<const uint64_t*>(s->frames + r.offset + align_up(8 * pixels, kAlign))
                                      : nullptr;
    out->submit_time = r.submit;
    out->start_time = r.start;
    out->done_time = r.done;
    out->batch_views = r.batchViews;
    out->uploads = r.uploads;
    out->status = r.status;
    out->reserved = 0;
    return CRX_OK;
}

crx_status crx_server_release(crx_server* s, uint64_t ticket) {
    if (!s) return fail(CRX_INVALID_ARG, "server is NULL");
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->reqs.find(ticket);
        if (it == s->reqs.end()) return fail(CRX_INVALID_ARG, "unknown ticket");
        if (!it->second.finished) return fail(CRX_INVALID_ARG, "request still in flight");
        s->arena.release(it->second.offset, it->second.bytes);
        s->reqs.erase(it);
    }
    s->workCv.notify_all();
    return CRX_OK;
}

double crx_server_clock(const crx_server* s) { return s ? s->now() : 0.0; }

crx_status crx_server_get_stats(const crx_server* s_, crx_server_stats* out) {
    if (!s_ || !out) return fail(CRX_INVALID_ARG, "NULL argument");
    crx_server* s = const_cast<crx_server*>(s_);
    std::memset(out, 0, sizeof(*out));
    {
        std::lock_guard<std::mutex> lk(s->mu);
        out->requests = s->statRequests;
        out->batches = s->statBatches;
        out->max_batch = s->statMaxBatch;
    }
    std::lock_guard<std::mutex> lk(s->cacheMu);
    out->uploads = s->statUploads;
    out->upload_bytes = s->statUploadBytes;
    out->evictions = s->statEvictions;
    out->prefetches = (uint32_t)std::min<uint64_t>(s->statPrefetches, UINT32_MAX);
    out->cache_slots = s->numSlots;
    uint32_t res = 0;
    for (int k : s->slotNode) res += k >= 0 ? 1u : 0u;
    out->resident_nodes = res;
    return CRX_OK;
}

crx_status crx_server_view_nodes(const crx_server* s, const crx_view* view, uint32_t* ids, uint32_t capacity,
                                 uint32_t* count) {
    if (!s || !view || !count) return fail(CRX_INVALID_ARG, "NULL argument");
    float planes[16];
    frustum_planes(*view, planes);
    uint32_t c = 0;
    if (s->started) {
        std::vector<uint32_t> vis;
        s->tree.visible(planes, s->wboxes.data(), vis);
        for (uint32_t i : vis) {
            if (ids && c < capacity) ids[c] = i;
            c++;
        }
    } else {
        for (uint32_t i = 0; i < s->nodes.size(); i++) {
            if (!box_in_frustum(planes, s->nodes[i].wbox)) continue;
            if (ids && c < capacity) ids[c] = i;
            c++;
        }
    }
    *count = c;
    return CRX_OK;
}

}  // extern "C"
