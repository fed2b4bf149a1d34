// crx_host.h — host-side internals of the C-ABI shared by crx_api.cpp and
// crx_server.cpp: device / pinned buffers, the context, mesh and scene
// objects, and the draw / resolve / clear drivers behind the entry points.
#pragma once

#include "crx.h"

#include <hip/hip_runtime.h>

#include <functional>
#include <string>
#include <vector>

#include "crx_device.h"

namespace crx {

// Records the message of a failing call (crx_last_error) and returns s.
crx_status fail(crx_status s, const std::string& msg);

#define CRX_HIP(expr)                                                                         \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess) return fail(CRX_HIP_ERROR, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define CRX_TRY(expr)                      \
    do {                                   \
        crx_status s_ = (expr);            \
        if (s_ != CRX_OK) return s_;       \
    } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

struct DeviceBuffer {
    void* ptr = nullptr;
    size_t capacity = 0;
    // bumped by every allocation: contents kept under an older generation
    // are gone even when the new allocation has the same address
    uint64_t generation = 0;
    crx_status reserve(size_t bytes) {
        if (bytes <= capacity) return CRX_OK;
        generation++;
        if (ptr) {
            hipError_t e = hipFree(ptr);
            ptr = nullptr;
            capacity = 0;
            if (e != hipSuccess) return fail(CRX_HIP_ERROR, std::string("hipFree: ") + hipGetErrorString(e));
        }
        size_t want = std::max<size_t>(bytes, 4096);
        hipError_t e = hipMalloc(&ptr, want);
        if (e != hipSuccess) {
            ptr = nullptr;
            return fail(CRX_OUT_OF_MEMORY, "hipMalloc(" + std::to_string(want) + "): " + hipGetErrorString(e));
        }
        capacity = want;
        return CRX_OK;
    }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        capacity = 0;
    }
};

// Pinned host staging with an event guarding reuse: the buffer is rewritten only
// after the copy that last read it has completed.
struct Staging {
    void* ptr = nullptr;
    size_t capacity = 0;
    hipEvent_t done = nullptr;
    bool pending = false;
    crx_status acquire(size_t bytes) {
        if (pending) {
            hipError_t e = hipEventSynchronize(done);
            if (e != hipSuccess) return fail(CRX_HIP_ERROR, std::string("hipEventSynchronize: ") + hipGetErrorString(e));
            pending = false;
        }
        if (!done) {
            hipError_t e = hipEventCreateWithFlags(&done, hipEventDisableTiming);
            if (e != hipSuccess) return fail(CRX_HIP_ERROR, std::string("hipEventCreate: ") + hipGetErrorString(e));
        }
        if (bytes <= capacity) return CRX_OK;
        if (ptr) (void)hipHostFree(ptr);
        ptr = nullptr;
        capacity = 0;
        size_t want = std::max<size_t>(bytes * 3 / 2, 65536);
        hipError_t e = hipHostMalloc(&ptr, want, hipHostMallocDefault);
        if (e != hipSuccess) {
            ptr = nullptr;
            return fail(CRX_OUT_OF_MEMORY, std::string("hipHostMalloc: ") + hipGetErrorString(e));
        }
        capacity = want;
        return CRX_OK;
    }
    crx_status mark(hipStream_t stream) {
        hipError_t e = hipEventRecord(done, stream);
        if (e != hipSuccess) return fail(CRX_HIP_ERROR, std::string("hipEventRecord: ") + hipGetErrorString(e));
        pending = true;
        return CRX_OK;
    }
    void release() {
        if (pending && done) (void)hipEventSynchronize(done);
        if (ptr) (void)hipHostFree(ptr);
        if (done) (void)hipEventDestroy(done);
        ptr = nullptr;
        done = nullptr;
        capacity = 0;
        pending = false;
    }
};

inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

enum TimerSlot { T_CLEAR = 0, T_S1, T_S2, T_S3, T_S3_END, T_RESOLVE, T_RESOLVE_END, T_COUNT };

// Host side of one draw: records, metadata layout and grid sizes.
struct DrawPlan {
    std::vector<DrawItem> items;
    std::vector<InstanceRec> inst;
    std::vector<Transform> xf;                 // view-major
    std::vector<crx_render_target> targets;
    uint32_t numViews = 0, numInstances = 0, numItems = 0;
    uint64_t batches = 0, tris = 0;
    bool pu16 = false;
    uint32_t block = 256, trisPerLane = 1, batchShift = 8, g1 = 0, g2 = 0, g3 = 0;
    bool instanceBoxes = false;   // some item loops over >= 4 instances
    bool runs = false;            // k_stage1_runs (items carry the run layout; batches counts its units)
    bool prefix = false;          // the image carries the instances' triangle prefix (InstanceRec::cumm)
    size_t offItems = 0, offInst = 0, offPrefix = 0, offViews = 0, offXf = 0, metaBytes = 0;
};

// What a caller adds to a draw.  listsKey != 0 names the content of the
// instance list (equal keys, bitwise-equal lists): a draw whose key matches
// the records the context's metadata buffer already holds neither rebuilds
// nor uploads them again.  prefix adds the instances' triangle prefix to the
// metadata, for a resolve that reads the draw's records.  afterStages is
// enqueued after the stages (and again after a redraw); with earlyReturn the
// draw returns once the queue counters are back, while that work may still
// run on the stream.  countersZeroed: the caller's clear zeroed the counter
// block in stream order before the draw.
struct DrawExtras {
    const uint8_t* skip = nullptr;
    const std::function<crx_status()>* afterStages = nullptr;
    uint64_t listsKey = 0;
    bool prefix = false;
    bool earlyReturn = false;
    bool countersZeroed = false;
};

// A hierarchy over ranges of consecutive nodes (a binary tree whose leaves
// hold up to 4 nodes): visible() lists, in index order, exactly the nodes
// box_in_frustum keeps, without testing each one.  A range is dropped when
// its box fails a plane and kept whole when its box passes every plane with
// its nearest corner; both tests are monotone in the box (a node's box lies
// inside its range's box, and each product and sum of the plane test rounds
// monotonically), so every node gets the answer of its own test.  Ranges that
// hold a non-finite coordinate are always split down to their nodes.
struct NodeTree {
    struct Range {
        crx_box3 box;
        uint32_t lo, hi;        // nodes [lo, hi)
        uint32_t left, right;   // child ranges (0: a leaf)
        bool finite;
    };
    std::vector<Range> ranges;  // ranges[0]: all nodes
    void build(const crx_box3* boxes, uint32_t n);
    void visible(const float planes[16], const crx_box3* boxes, std::vector<uint32_t>& out) const;
};

}  // namespace crx

struct crx_context {
    crx::DrawPlan plan;        // draw records reused across draws (vectors keep their capacity)
    int device = 0;
    int numCUs = 0;
    crx_tuning tuning{};
    uint32_t* ctr = nullptr;          // device counters (CTR_COUNT)
    uint32_t* ctrHost = nullptr;      // pinned readback
    crx::DeviceBuffer nontrivial;          // uint64 entries
    crx::DeviceBuffer huge;                // int4 entries
    uint32_t ntCap = 0, hugeCap = 0;
    crx::DeviceBuffer drawMeta;            // items | instances | transforms
    crx::DeviceBuffer resolveMeta;         // instances | prefix | per-view records
    crx::DeviceBuffer clearMeta;           // per-buffer clear records
    crx::Staging drawStaging, resolveStaging, clearStaging;
    // instance records (and prefix) that drawMeta holds: the list key they
    // were uploaded under (0 = none), the buffer and their byte range
    uint64_t metaListsKey = 0;
    uint64_t metaListsGen = 0;        // drawMeta.generation the records were written under
    size_t metaListsBegin = 0, metaListsEnd = 0;
    bool metaListsPrefix = false;
    hipEvent_t countersBack = nullptr;   // the counter readback of the last draw
    hipEvent_t metaFree = nullptr;       // recorded after the last draw's work that reads the metadata buffers
    hipStream_t metaFreeStream = nullptr;
    bool metaFreePending = false;
    bool timing = false;
    hipEvent_t ev[crx::T_COUNT] = {};
    bool recorded[crx::T_COUNT] = {};
    uint64_t meshSeq = 0;
    uint64_t attrEpoch = 0;           // bumped by every crx_mesh_set_attributes on this context
    // kernel-level entry points: the records buffer every crx_raster_stage*
    // call rebuilds its records in, and the event after the last call's
    // launches (a call on another stream waits for it)
    crx::DeviceBuffer argsMeta;
    hipEvent_t argsUse = nullptr;
    hipStream_t argsUseStream = nullptr;
    bool argsUsePending = false;
};

struct crx_mesh {
    crx_context* ctx = nullptr;
    float* positions = nullptr;       // IX_PU16: 3 x uint16 per vertex
    uint32_t* indices = nullptr;      // IX_PU16: bit-packed
    uint32_t* colors = nullptr;
    bool owned = false;
    uint32_t numVertices = 0;
    uint32_t numTriangles = 0;
    crx_box3 aabb{};
    uint64_t seq = 0;
    // IX_PU16 description (crx_compress_mesh)
    bool compressed = false;
    uint32_t bitsPerIndex = 0, indexMin = 0, indexMax = 0;
    crx_vec3 compressionFactor{};
    uint32_t rowTris = 0;             // triangles per row of a mesh stored in rows (0 = none found)
    // resolve-only vertex attributes, always owned by the mesh (crx_mesh_set_attributes)
    float* uvs = nullptr;             // 2 floats per vertex
    float* normals = nullptr;         // 3 floats per vertex
};

struct crx_scene {
    crx_context* ctx = nullptr;
    struct Node {
        crx_mesh* mesh;
        crx_mat4 world;
        crx_box3 wbox;   // world-space box of the mesh's box (kept in step with world)
        crx_texture texture;   // SNTriangles::texture (zero: none)
    };
    std::vector<Node> nodes;
    // frame-preparation scratch, kept across frames like the reference's statics
    std::vector<crx_cmesh> unique, instances;
    std::vector<crx_mat4> transforms;
    std::vector<uint64_t> prefix;
    std::vector<uint8_t> skip;        // [view][instance]: dropped by that view's node cull
    std::vector<uint32_t> nodeOf;     // node of each listed instance
    // node-set version (every node change bumps it) and what the lists were
    // built from; a frame with the same version, instancing and kept nodes
    // keeps them and recomputes only the per-view transforms and skips
    uint64_t version = 0;
    uint64_t listsVersion = ~0ull;
    int listsInstanced = -1;
    std::vector<uint32_t> listsKept;
    uint64_t listsKey = 0;            // content key of the lists (new on every rebuild)
    uint64_t listsAttrEpoch = 0;      // the context's attrEpoch the lists were built at
    // the node cull's range hierarchy over the nodes' world boxes, as of treeVersion
    std::vector<crx_box3> wboxes;
    crx::NodeTree tree;
    uint64_t treeVersion = ~0ull;
};


namespace crx {

crx_mat4 mat_mul(const crx_mat4& A, const crx_mat4& B);
crx_mat4 mat_inverse(const crx_mat4& M);
bool mirror_flag(const crx_mat4& m);
ViewParams make_view_params(const crx_render_target& t, const Transform* transforms);
uint32_t grid_for(const crx_context* ctx, uint32_t wavesPerCU, uint32_t wavesPerBlock);
void set_counter_block(RasterParams& P, uint32_t* block);

// Stage-1 schedule of a draw: triangles per lane and grid size in waves per
// CU (0 = one workgroup per batch).  With stage1_tris_per_lane = 0 a draw
// that repeats a mesh over many instances takes one triangle per lane (its
// sweep keeps instance lists per wave and runs at higher occupancy) and 64
// waves per CU (its batches differ widely in cost, so more workgroups than
// fit at once let the dispatcher even out the load); an IX_PU16 draw takes
// two triangles per lane and 64 waves per CU (its index and vertex decode
// waits on LDS and loads, and more waves hide it: 1.89 ms against 2.10 ms at
// 32 on the 400 M-triangle heightfield); other draws take two triangles per
// lane and the tuning's grid; a draw of several views takes one triangle
// per lane and 64 waves per CU (its per-(triangle, view) work varies most:
// 11.07 ms against 12.25 ms at two per lane for 64 views of the 400 M-triangle
// heightfield).
struct S1Schedule {
    uint32_t trisPerLane, wavesPerCU;
};
inline S1Schedule s1_schedule(const crx_tuning& t, bool manyInstances, bool pu16, bool multiView = false) {
    if (t.stage1_tris_per_lane) return {t.stage1_tris_per_lane, t.stage1_waves_per_cu};
    if (manyInstances || multiView) return S1Schedule{1u, 64u};
    return pu16 ? S1Schedule{2u, 64u} : S1Schedule{2u, t.stage1_waves_per_cu};
}
crx_status check_tuning(const crx_tuning& t);
// The static scratch of crx_raster_visbuffer_noalloc on `device` (current):
// its address is looked up once per device.
crx_status noalloc_scratch_init(int device);
crx_status check_target(const crx_render_target* t);
void frustum_planes(const crx_view& view, float out[16]);
crx_box3 world_box(const crx_box3& b, const crx_mat4& world);
bool box_in_frustum(const float planes[16], const crx_box3& worldBox);


crx_status plan_draw(crx_context* ctx, const crx_cmesh* meshes, uint32_t numMeshes, const crx_cmesh* instances,
                     uint32_t numInstances, const crx_mat4* transforms, const crx_render_target* targets,
                     uint32_t numViews, int instanced, const uint8_t* skip, DrawPlan* plan,
                     bool instUnchanged = false, bool prefix = false);
RasterParams plan_write(crx_context* ctx, const DrawPlan& plan, char* host, char* dev, bool writeInst = true);
hipError_t enqueue_stages(crx_context* ctx, const DrawPlan& plan, const RasterParams& P, bool viewsKernel,
                          hipStream_t stream, bool countersZeroed = false);
crx_status ensure_queues(crx_context* ctx, uint32_t nt, uint32_t hg);

// drawTrianglesVisbuffer for numViews >= 1 targets sharing one instance list
// (transforms view-major).  extras.skip (nullable, numViews * numInstances
// bytes, view-major): a non-zero byte drops instance i from view v, as the
// host node cull of that view would; it needs numViews > 1.
crx_status draw_views(crx_context* ctx, const crx_cmesh* meshes, uint32_t numMeshes, const crx_cmesh* instances,
                      uint32_t numInstances, const crx_mat4* transforms, const crx_render_target* targets,
                      uint32_t numViews, int instanced, hipStream_t stream, crx_draw_stats* stats,
                      const DrawExtras& extras = DrawExtras{});
// The resolve of numViews targets over instance records and prefix already in
// device memory (a draw's metadata); uploads at most the per-view records.
crx_status resolve_views_device(crx_context* ctx, const InstanceRec* instDev, const uint64_t* prefixDev,
                                uint32_t numInstances, const crx_render_target* targets, uint32_t numViews,
                                crx_display_attribute attribute, hipStream_t stream);
// ownedOnly: pixels whose triangle lies outside the listed instances are
// left as they are (a resolve in passes over groups of instances).
crx_status resolve_views(crx_context* ctx, const crx_cmesh* instances, uint32_t numInstances, const uint64_t* prefix,
                         const crx_render_target* targets, uint32_t numViews, crx_display_attribute attribute,
                         hipStream_t stream, bool ownedOnly = false);
// Clears n buffers in one launch; a non-null ctr counter block is zeroed by
// the same launch (*zeroed tells whether one ran).
crx_status clear_buffers(crx_context* ctx, uint64_t* const* buffers, const uint32_t* numPixels, uint32_t n,
                         hipStream_t stream, uint32_t* ctr = nullptr, bool* zeroed = nullptr);
uint32_t scene_lists(crx_scene* scene, const crx_view* views, uint32_t numViews, const crx_render_settings& settings,
                     bool keepAll);
void view_targets(const crx_view* views, uint32_t numViews, uint64_t* const* framebuffers,
                  uint64_t* const* colorbuffers, bool resolve, std::vector<crx_render_target>* targets,
                  std::vector<uint64_t*>* clearBufs, std::vector<uint32_t>* clearPx);
size_t resolve_image_layout(uint32_t numInstances, uint32_t numViews, size_t* offPrefix, size_t* offViews);
crx_status resolve_image(const crx_cmesh* instances, uint32_t numInstances, const uint64_t* prefix,
                         const crx_render_target* targets, uint32_t numViews, crx_display_attribute attribute,
                         char* dev, std::vector<char>* image, size_t* offViews, uint32_t* blocks);
crx_status clear_records(uint64_t* const* buffers, const uint32_t* numPixels, uint32_t n, std::vector<ClearView>* cv,
                         uint32_t* blocks);

}  // namespace crx
