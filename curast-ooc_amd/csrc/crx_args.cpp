// crx_args.cpp — kernel-level entry points: the three stages of
// drawTrianglesVisbuffer as separate calls over caller-owned device data.
//
// The reference's host code launches kernel_stage{1,2,3}_* by name with one
// by-value RasterArgs (CuRast_render.h:235-264, triangles_visbuffer.cu:755-835)
// whose CMesh / mat4 arrays, counters and queues all live in device memory.
// crx_raster_stage1 takes the same fields (crx_raster_args), builds the
// compact draw records on the device (k_plan_items / k_plan_instances, so the
// call never reads device memory from the host and stays asynchronous),
// zeroes the five counters as the reference's stage 1 does before its grid
// barrier, and launches the stage-1 kernel; stages 2 and 3 reuse that plan.
#include <hip/hip_runtime.h>

#include <cstring>
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
    return CRX_OK;
}

bool same_draw(const crx_context* ctx, const crx_raster_args* a) {
    return ctx->args.valid && ctx->args.instances == a->instances && ctx->args.transforms == a->transforms &&
           ctx->args.numInstances == a->numInstances && ctx->args.framebuffer == a->target.framebuffer;
}

// Queue and counter addresses of the args (stage 2 and 3 read them again).
void bind_queues(RasterParams& P, const crx_raster_args* a) {
    P.ctrNontrivial = a->nontrivialTrianglesCounter;
    P.ctrHuge = a->hugeTrianglesCounter;
    P.cursor2 = a->numProcessedBatches_nontrivial;
    P.cursor3 = a->numProcessedHugeTriangles;
    P.nontrivial = a->nontrivialTrianglesList;
    P.huge = reinterpret_cast<int4*>(a->hugeTriangles);
    P.ntCap = a->nontrivialCapacity;
    P.hugeCap = a->hugeCapacity;
}

}  // namespace
}  // namespace crx

using namespace crx;

extern "C" {

crx_status crx_raster_stage1(crx_context* ctx, const crx_raster_args* a, uint32_t flags, void* stream) {
    CRX_TRY(check_args(ctx, a));
    if (flags & ~(CRX_RASTER_INSTANCED | CRX_RASTER_COMPRESSED)) return fail(CRX_INVALID_ARG, "unknown flags");
    const bool instanced = (flags & CRX_RASTER_INSTANCED) != 0;
    const bool pu16 = (flags & CRX_RASTER_COMPRESSED) != 0;
    const uint32_t numItems = instanced ? a->numMeshes : a->numInstances;
    if (numItems > 0 && !(instanced ? a->meshes : a->instances)) return fail(CRX_INVALID_ARG, "mesh list is NULL");
    if (a->numInstances > 0 && (!a->instances || !a->transforms))
        return fail(CRX_INVALID_ARG, "instances / transforms is NULL");
    hipStream_t s = as_stream(stream);
    CRX_HIP(hipSetDevice(ctx->device));
    ctx->args.valid = false;
    CounterSet cs{{a->numProcessedBatches, a->numProcessedBatches_nontrivial, a->hugeTrianglesCounter,
                   a->nontrivialTrianglesCounter, a->numProcessedHugeTriangles}};
    CRX_HIP(launch_reset_counter_set(cs, s));

    const crx_tuning& tu = ctx->tuning;
    RasterParams P{};
    bind_queues(P, a);
    // instance counts live in device memory: an instanced draw takes the
    // many-instances choice
    const S1Schedule sched = s1_schedule(tu, instanced, pu16);
    const uint32_t perLane = sched.trisPerLane;
    uint32_t batchShift = 0;
    while ((1u << batchShift) < tu.stage1_block * perLane) batchShift++;
    // device image: items | instances | transforms | view | batch count
    const size_t offItems = 0;
    const size_t offInst = align_up(offItems + (size_t)numItems * sizeof(DrawItem), 256);
    const size_t offXf = align_up(offInst + (size_t)a->numInstances * sizeof(InstanceRec), 256);
    const size_t offView = align_up(offXf + (size_t)a->numInstances * sizeof(Transform), 256);
    const size_t offCount = align_up(offView + sizeof(ViewParams), 256);
    const size_t bytes = offCount + 256;
    CRX_TRY(ctx->argsMeta.reserve(bytes));
    char* dev = static_cast<char*>(ctx->argsMeta.ptr);
    P.items = reinterpret_cast<const DrawItem*>(dev + offItems);
    P.inst = reinterpret_cast<const InstanceRec*>(dev + offInst);
    P.views = reinterpret_cast<const ViewParams*>(dev + offView);
    const ViewParams view = make_view_params(a->target, reinterpret_cast<const Transform*>(dev + offXf));
    P.numViews = 1;
    P.numItems = numItems;
    P.numBatches = 0;
    P.batchCount = reinterpret_cast<const uint32_t*>(dev + offCount);
    P.smallThreshold = tu.small_threshold;
    P.tile = (int32_t)tu.tile_size;
    P.batchShift = batchShift;
    P.order = tu.stage1_order;
    P.cullOrder = tu.stage1_cull_order;
    P.fastTiny = tu.stage1_fast_tiny;
    const uint32_t g1 = (grid_for(ctx, sched.wavesPerCU ? sched.wavesPerCU : 32u, tu.stage1_block / 64) + 7u) & ~7u;
    const uint32_t g2 = grid_for(ctx, tu.stage2_waves_per_cu, 4), g3 = grid_for(ctx, tu.stage3_waves_per_cu, 4);
    P.numWaves2 = g2 * 4u;
    P.numWaves3 = g3 * 4u;

    if (numItems > 0 && a->numInstances > 0) {
        // the view record goes up through page-locked staging, the rest is built on the device
        CRX_TRY(ctx->argsStaging.acquire(sizeof(ViewParams)));
        std::memcpy(ctx->argsStaging.ptr, &view, sizeof(ViewParams));
        CRX_HIP(hipMemcpyAsync(dev + offView, ctx->argsStaging.ptr, sizeof(ViewParams), hipMemcpyHostToDevice, s));
        CRX_TRY(ctx->argsStaging.mark(s));
        CRX_HIP(launch_plan(instanced ? a->meshes : a->instances, numItems, a->instances, a->transforms,
                            a->numInstances, instanced ? 1 : 0, pu16 ? 1 : 0, batchShift,
                            reinterpret_cast<DrawItem*>(dev + offItems), reinterpret_cast<InstanceRec*>(dev + offInst),
                            reinterpret_cast<Transform*>(dev + offXf), reinterpret_cast<uint32_t*>(dev + offCount), s));
        CRX_HIP(launch_stage1(P, tu.stage1_block, perLane, g1, pu16, instanced, s));
    }
    ctx->args.valid = true;
    ctx->args.instances = a->instances;
    ctx->args.transforms = a->transforms;
    ctx->args.numInstances = a->numInstances;
    ctx->args.framebuffer = a->target.framebuffer;
    ctx->args.P = P;
    ctx->args.g2 = (numItems > 0 && a->numInstances > 0) ? g2 : 0u;
    ctx->args.g3 = (numItems > 0 && a->numInstances > 0) ? g3 : 0u;
    return CRX_OK;
}

crx_status crx_raster_stage2(crx_context* ctx, const crx_raster_args* a, void* stream) {
    CRX_TRY(check_args(ctx, a));
    if (!same_draw(ctx, a)) return fail(CRX_INVALID_ARG, "crx_raster_stage2 needs crx_raster_stage1 of the same args first");
    if (ctx->args.g2 == 0) return CRX_OK;
    CRX_HIP(hipSetDevice(ctx->device));
    RasterParams P = ctx->args.P;
    bind_queues(P, a);
    CRX_HIP(launch_stage2(P, ctx->args.g2, as_stream(stream)));
    return CRX_OK;
}

crx_status crx_raster_stage3(crx_context* ctx, const crx_raster_args* a, void* stream) {
    CRX_TRY(check_args(ctx, a));
    if (!same_draw(ctx, a)) return fail(CRX_INVALID_ARG, "crx_raster_stage3 needs crx_raster_stage1 of the same args first");
    if (ctx->args.g3 == 0) return CRX_OK;
    CRX_HIP(hipSetDevice(ctx->device));
    RasterParams P = ctx->args.P;
    bind_queues(P, a);
    CRX_HIP(launch_stage3(P, ctx->args.g3, as_stream(stream)));
    return CRX_OK;
}

}  // extern "C"
