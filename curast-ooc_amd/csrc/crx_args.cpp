// crx_args.cpp — kernel-level entry points: the three stages of
// drawTrianglesVisbuffer as separate calls over caller-owned device data.
//
// The reference's host code launches kernel_stage{1,2,3}_* by name with one
// by-value RasterArgs (CuRast_render.h:235-264, triangles_visbuffer.cu:755-835)
// whose CMesh / mat4 arrays, counters and queues all live in device memory,
// and each stage kernel depends on nothing but its RasterArgs.  Here each
// call does the same: it builds the compact draw records its kernel reads
// from the args' device CMesh arrays on the device (k_plan_items /
// k_plan_instances, and the view record stored by a kernel from its
// argument, so a call never reads device memory from the host and stays
// asynchronous), then launches its stage.  Stage 1 also zeroes the five
// counters, as the reference's stage 1 does before its grid barrier.
//
// The records of every call on a context go to the context's one records
// buffer.  A call on another stream than the previous call waits (on the
// device) for that call's launches, so no call rewrites records that a
// kernel of an earlier call may still read.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <string>

#include "crx.h"
#include "crx_device.h"
#include "crx_host.h"
#include "crx_internal.h"

namespace crx {
namespace {

crx_status check_args(const crx_context* ctx, const crx_raster_args* a) {
    if (!ctx) return fail(CRX_INVALID_ARG, "ctx is NULL");
    if (!a) return fail(CRX_INVALID_ARG, "args is NULL");
    CRX_TRY(check_target(&a->target));
    if (!a->numProcessedBatches || !a->numProcessedBatches_nontrivial || !a->hugeTrianglesCounter ||
        !a->numProcessedHugeTriangles || !a->nontrivialTrianglesCounter)
        return fail(CRX_INVALID_ARG, "args counters must all be device addresses");
    if (!a->nontrivialTrianglesList || !a->hugeTriangles)
        return fail(CRX_INVALID_ARG, "args queues must be device addresses");
    if (a->nontrivialCapacity == 0 || a->hugeCapacity == 0)
        return fail(CRX_INVALID_ARG, "args queue capacities must be > 0");
    if (a->numInstances > kInstanceMask + 1u) return fail(CRX_INVALID_ARG, "more than 2^24 instances");
    if (a->numInstances > 0 && (!a->instances || !a->transforms))
        return fail(CRX_INVALID_ARG, "instances / transforms is NULL");
    return CRX_OK;
}

// Byte layout of one call's records: items | instances | transforms | view | batch count.
struct ArgsImage {
    size_t offItems = 0, offInst = 0, offXf = 0, offView = 0, offCount = 0, bytes = 0;
};

ArgsImage args_image(uint32_t numItems, uint32_t numInstances) {
    ArgsImage L;
    L.offItems = 0;
    L.offInst = align_up(L.offItems + (size_t)numItems * sizeof(DrawItem), 256);
    L.offXf = align_up(L.offInst + (size_t)numInstances * sizeof(InstanceRec), 256);
    L.offView = align_up(L.offXf + (size_t)numInstances * sizeof(Transform), 256);
    L.offCount = align_up(L.offView + sizeof(ViewParams), 256);
    L.bytes = L.offCount + 256;
    return L;
}

// What a stage launch of args `a` needs: the records at `dev` (layout L),
// the kernel parameters and the grids.  `items` = numItems draw items are
// built (stage 1); stages 2 and 3 only read instances, transforms and the view.
struct ArgsLaunch {
    RasterParams P{};
    uint32_t perLane = 1, g1 = 0, g2 = 0, g3 = 0;
};

ArgsLaunch args_launch(const crx_context* ctx, const crx_raster_args* a, bool instanced, bool pu16, uint32_t numItems,
                       char* dev, const ArgsImage& L) {
    const crx_tuning& tu = ctx->tuning;
    ArgsLaunch r;
    RasterParams& P = r.P;
    P.ctrNontrivial = a->nontrivialTrianglesCounter;
    P.ctrHuge = a->hugeTrianglesCounter;
    P.cursor2 = a->numProcessedBatches_nontrivial;
    P.cursor3 = a->numProcessedHugeTriangles;
    P.nontrivial = a->nontrivialTrianglesList;
    P.huge = reinterpret_cast<int4*>(a->hugeTriangles);
    P.ntCap = a->nontrivialCapacity;
    P.hugeCap = a->hugeCapacity;
    // instance counts live in device memory: an instanced draw takes the
    // many-instances choice
    const S1Schedule sched = s1_schedule(tu, instanced, pu16);
    r.perLane = sched.trisPerLane;
    uint32_t batchShift = 0;
    while ((1u << batchShift) < tu.stage1_block * r.perLane) batchShift++;
    P.items = reinterpret_cast<const DrawItem*>(dev + L.offItems);
    P.inst = reinterpret_cast<const InstanceRec*>(dev + L.offInst);
    P.views = reinterpret_cast<const ViewParams*>(dev + L.offView);
    P.numViews = 1;
    P.numItems = numItems;
    P.numBatches = 0;
    P.batchCount = reinterpret_cast<const uint32_t*>(dev + L.offCount);
    P.smallThreshold = tu.small_threshold;
    P.tile = (int32_t)tu.tile_size;
    P.batchShift = batchShift;
    P.order = tu.stage1_order;
    P.cullOrder = tu.stage1_cull_order;
    P.fastTiny = tu.stage1_fast_tiny;
    r.g1 = (grid_for(ctx, sched.wavesPerCU ? sched.wavesPerCU : 32u, tu.stage1_block / 64) + 7u) & ~7u;
    r.g2 = grid_for(ctx, tu.stage2_waves_per_cu, 4);
    r.g3 = grid_for(ctx, tu.stage3_waves_per_cu, 4);
    P.numWaves2 = r.g2 * 4u;
    P.numWaves3 = r.g3 * 4u;
    return r;
}

// Builds the records of args `a` at `dev` on stream s: draw items (stage 1
// only, numItems > 0), instance records, transforms and the view record.
hipError_t build_records(const crx_raster_args* a, bool instanced, bool pu16, uint32_t numItems, uint32_t batchShift,
                         char* dev, const ArgsImage& L, hipStream_t s) {
    const ViewParams view = make_view_params(a->target, reinterpret_cast<const Transform*>(dev + L.offXf));
    hipError_t e = launch_store_view(view, reinterpret_cast<ViewParams*>(dev + L.offView), s);
    if (e != hipSuccess) return e;
    return launch_plan(numItems ? (instanced ? a->meshes : a->instances) : nullptr, numItems, a->instances,
                       a->transforms, a->numInstances, instanced ? 1 : 0, pu16 ? 1 : 0, batchShift,
                       reinterpret_cast<DrawItem*>(dev + L.offItems), reinterpret_cast<InstanceRec*>(dev + L.offInst),
                       reinterpret_cast<Transform*>(dev + L.offXf), reinterpret_cast<uint32_t*>(dev + L.offCount), s);
}

// The context's records buffer, at least `bytes`, ordered on `stream` after
// every launch of the previous kernel-level call.
crx_status args_buffer(crx_context* ctx, size_t bytes, hipStream_t stream, char** dev) {
    if (!ctx->argsUse) CRX_HIP(hipEventCreateWithFlags(&ctx->argsUse, hipEventDisableTiming));
    if (ctx->argsUsePending && stream != ctx->argsUseStream) CRX_HIP(hipStreamWaitEvent(stream, ctx->argsUse, 0));
    if (bytes > ctx->argsMeta.capacity) {
        // kernels of earlier calls may still read the old buffer
        if (ctx->argsUsePending) CRX_HIP(hipEventSynchronize(ctx->argsUse));
        CRX_TRY(ctx->argsMeta.reserve(bytes));
    }
    *dev = static_cast<char*>(ctx->argsMeta.ptr);
    return CRX_OK;
}

crx_status args_done(crx_context* ctx, hipStream_t stream) {
    CRX_HIP(hipEventRecord(ctx->argsUse, stream));
    ctx->argsUsePending = true;
    ctx->argsUseStream = stream;
    return CRX_OK;
}

crx_status check_flags(const crx_raster_args* a, uint32_t flags, uint32_t* numItems) {
    if (flags & ~(CRX_RASTER_INSTANCED | CRX_RASTER_COMPRESSED)) return fail(CRX_INVALID_ARG, "unknown flags");
    const bool instanced = (flags & CRX_RASTER_INSTANCED) != 0;
    *numItems = instanced ? a->numMeshes : a->numInstances;
    if (*numItems > 0 && !(instanced ? a->meshes : a->instances)) return fail(CRX_INVALID_ARG, "mesh list is NULL");
    return CRX_OK;
}

// Stage 2 or 3 of args `a` on stream s with the records at dev.
crx_status run_late_stage(crx_context* ctx, const crx_raster_args* a, int stage, char* dev, hipStream_t s) {
    const ArgsImage L = args_image(0, a->numInstances);
    const ArgsLaunch r = args_launch(ctx, a, false, false, 0, dev, L);
    CRX_HIP(build_records(a, false, false, 0, r.P.batchShift, dev, L, s));
    if (stage == 2) CRX_HIP(launch_stage2(r.P, r.g2, s));
    else CRX_HIP(launch_stage3(r.P, r.g3, s));
    return CRX_OK;
}

// Scratch of crx_raster_visbuffer_noalloc: the draw records of one call
// (items | instances | transforms | view | batch count), in device memory the
// library never allocates: a file-scope array that each GPU holds until the
// process ends.  Host code uses the address hipGetSymbolAddress returns for
// the current device, never the array's name.
constexpr size_t kNoAllocScratchBytes = size_t(4) << 20;
static __device__ __align__(256) char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratchMu;
void* g_scratch[kMaxDevices] = {};

}  // namespace

// Looks up this device's copy of the scratch array once (the device must be
// current); crx_context_create calls it at set-up.
crx_status noalloc_scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return fail(CRX_INVALID_ARG, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratchMu);
    if (g_scratch[device]) return CRX_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(CRX_HIP_ERROR, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[device] = p;
    return CRX_OK;
}

}  // namespace crx

using namespace crx;

extern "C" {

crx_status crx_raster_stage1(crx_context* ctx, const crx_raster_args* a, uint32_t flags, void* stream) {
    CRX_TRY(check_args(ctx, a));
    uint32_t numItems = 0;
    CRX_TRY(check_flags(a, flags, &numItems));
    const bool instanced = (flags & CRX_RASTER_INSTANCED) != 0;
    const bool pu16 = (flags & CRX_RASTER_COMPRESSED) != 0;
    hipStream_t s = as_stream(stream);
    CRX_HIP(hipSetDevice(ctx->device));
    CounterSet cs{{a->numProcessedBatches, a->numProcessedBatches_nontrivial, a->hugeTrianglesCounter,
                   a->nontrivialTrianglesCounter, a->numProcessedHugeTriangles}};
    CRX_HIP(launch_reset_counter_set(cs, s));
    if (numItems == 0 || a->numInstances == 0) return CRX_OK;
    const ArgsImage L = args_image(numItems, a->numInstances);
    char* dev = nullptr;
    CRX_TRY(args_buffer(ctx, L.bytes, s, &dev));
    const ArgsLaunch r = args_launch(ctx, a, instanced, pu16, numItems, dev, L);
    CRX_HIP(build_records(a, instanced, pu16, numItems, r.P.batchShift, dev, L, s));
    CRX_HIP(launch_stage1(r.P, ctx->tuning.stage1_block, r.perLane, r.g1, pu16, instanced, s));
    return args_done(ctx, s);
}

crx_status crx_raster_stage2(crx_context* ctx, const crx_raster_args* a, void* stream) {
    CRX_TRY(check_args(ctx, a));
    if (a->numInstances == 0) return CRX_OK;
    hipStream_t s = as_stream(stream);
    CRX_HIP(hipSetDevice(ctx->device));
    char* dev = nullptr;
    CRX_TRY(args_buffer(ctx, args_image(0, a->numInstances).bytes, s, &dev));
    CRX_TRY(run_late_stage(ctx, a, 2, dev, s));
    return args_done(ctx, s);
}

crx_status crx_raster_stage3(crx_context* ctx, const crx_raster_args* a, void* stream) {
    CRX_TRY(check_args(ctx, a));
    if (a->numInstances == 0) return CRX_OK;
    hipStream_t s = as_stream(stream);
    CRX_HIP(hipSetDevice(ctx->device));
    char* dev = nullptr;
    CRX_TRY(args_buffer(ctx, args_image(0, a->numInstances).bytes, s, &dev));
    CRX_TRY(run_late_stage(ctx, a, 3, dev, s));
    return args_done(ctx, s);
}

crx_status crx_raster_visbuffer_noalloc(crx_context* ctx, const crx_raster_args* a, uint32_t flags, void* stream) {
    CRX_TRY(check_args(ctx, a));
    uint32_t numItems = 0;
    CRX_TRY(check_flags(a, flags, &numItems));
    const bool instanced = (flags & CRX_RASTER_INSTANCED) != 0;
    const bool pu16 = (flags & CRX_RASTER_COMPRESSED) != 0;
    const ArgsImage L = args_image(numItems, a->numInstances);
    if (L.bytes > kNoAllocScratchBytes)
        return fail(CRX_INVALID_ARG, "crx_raster_visbuffer_noalloc: the draw records take " + std::to_string(L.bytes) +
                                         " bytes, more than the " + std::to_string(kNoAllocScratchBytes) +
                                         "-byte static scratch holds (fewer meshes / instances per call)");
    char* dev = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratchMu);
        if (ctx->device >= 0 && ctx->device < kMaxDevices) dev = static_cast<char*>(g_scratch[ctx->device]);
    }
    if (!dev) return fail(CRX_INVALID_ARG, "crx_raster_visbuffer_noalloc: no scratch address for the context's device");
    hipStream_t s = as_stream(stream);
    CRX_HIP(hipSetDevice(ctx->device));
    CounterSet cs{{a->numProcessedBatches, a->numProcessedBatches_nontrivial, a->hugeTrianglesCounter,
                   a->nontrivialTrianglesCounter, a->numProcessedHugeTriangles}};
    CRX_HIP(launch_reset_counter_set(cs, s));
    if (numItems == 0 || a->numInstances == 0) return CRX_OK;
    // one image for all three stages: stages 2 and 3 read its instances,
    // transforms and view, which stage 1's records include
    const ArgsLaunch r = args_launch(ctx, a, instanced, pu16, numItems, dev, L);
    CRX_HIP(build_records(a, instanced, pu16, numItems, r.P.batchShift, dev, L, s));
    CRX_HIP(launch_stage1(r.P, ctx->tuning.stage1_block, r.perLane, r.g1, pu16, instanced, s));
    CRX_HIP(launch_stage2(r.P, r.g2, s));
    CRX_HIP(launch_stage3(r.P, r.g3, s));
    return CRX_OK;
}

}  // extern "C"
