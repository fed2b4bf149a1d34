// crx_frame.cpp — a scene frame recorded once into a hipGraph
// (include/crx.h, crx_frame_*).
//
// The graph holds three host-to-device copies of the frame's metadata from
// page-locked staging buffers, the clear of every target, the counter reset,
// the three stages (every stage reads the views' records from device memory,
// so new cameras only need new staging contents), a copy
// of the counters back to page-locked memory and the resolve.  Launching the
// frame is one hipGraphLaunch; new cameras rewrite the staging buffers.  The
// clear and the counter reset are kernels: the graph records no memset node.
#include "crx.h"

#include <hip/hip_runtime.h>

#include <cstring>
#include <string>
#include <vector>

#include "crx_device.h"
#include "crx_host.h"
#include "crx_internal.h"

using namespace crx;

struct crx_frame {
    crx_context* ctx = nullptr;
    crx_scene* scene = nullptr;
    crx_render_settings settings{};
    uint32_t numViews = 0;
    std::vector<crx_view> views;
    std::vector<uint64_t*> fbs, cbs;

    DrawPlan plan;
    size_t nodes = 0;                 // scene node count at creation
    DeviceBuffer drawMeta, resolveMeta, clearMeta, nontrivial, huge;
    uint32_t ntCap = 0, hugeCap = 0;
    uint32_t* ctr = nullptr;          // device counters
    uint32_t* ctrHost = nullptr;      // page-locked copy written by the graph
    void* stageDraw = nullptr;        // page-locked sources of the graph's copies
    void* stageResolve = nullptr;
    void* stageClear = nullptr;
    size_t drawBytes = 0, resolveBytes = 0, clearBytes = 0;
    size_t resolveOffViews = 0;
    uint32_t resolveBlocks = 0, clearBlocks = 0, numClear = 0;
    RasterParams P{};

    hipStream_t capture = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t done = nullptr;
    bool launched = false;
};

namespace {

// Host records of the frame for its current views; with `write`, the
// metadata images are written into the staging buffers for the frame's
// device buffers.  Without it, only the sizes are computed.
crx_status frame_fill(crx_frame* F, bool write) {
    crx_scene* sc = F->scene;
    if (sc->nodes.size() != F->nodes) return fail(CRX_INVALID_ARG, "the scene's node set changed: recreate the frame");
    const uint32_t u = scene_lists(sc, F->views.data(), F->numViews, F->settings, true);
    const uint32_t n = (uint32_t)sc->instances.size();
    std::vector<crx_render_target> targets;
    std::vector<uint64_t*> clearBufs;
    std::vector<uint32_t> clearPx;
    view_targets(F->views.data(), F->numViews, F->fbs.data(), F->cbs.data(), F->settings.resolve != 0, &targets,
                 &clearBufs, &clearPx);
    DrawPlan D;
    CRX_TRY(plan_draw(F->ctx, sc->unique.data(), u, sc->instances.data(), n, sc->transforms.data(), targets.data(),
                      F->numViews, F->settings.instanced, F->settings.frustum_culling ? sc->skip.data() : nullptr,
                      &D));
    std::vector<char> rimage;
    size_t offViews = 0;
    uint32_t rblocks = 0;
    if (F->settings.resolve) {
        CRX_TRY(resolve_image(sc->instances.data(), n, sc->prefix.data(), targets.data(), F->numViews,
                              (crx_display_attribute)F->settings.attribute,
                              static_cast<char*>(F->resolveMeta.ptr), &rimage, &offViews, &rblocks));
    }
    std::vector<ClearView> cv;
    uint32_t cblocks = 0;
    CRX_TRY(clear_records(clearBufs.data(), clearPx.data(), (uint32_t)clearBufs.size(), &cv, &cblocks));
    if (!write) {
        F->plan = D;
        F->drawBytes = D.metaBytes;
        F->resolveBytes = rimage.size();
        F->resolveOffViews = offViews;
        F->resolveBlocks = rblocks;
        F->clearBytes = cv.size() * sizeof(ClearView);
        F->clearBlocks = cblocks;
        F->numClear = (uint32_t)cv.size();
        return CRX_OK;
    }
    if (D.metaBytes != F->drawBytes || D.batches != F->plan.batches || rimage.size() != F->resolveBytes)
        return fail(CRX_INVALID_ARG, "the scene's draw layout changed: recreate the frame");
    F->plan = D;
    F->P = plan_write(F->ctx, D, static_cast<char*>(F->stageDraw), static_cast<char*>(F->drawMeta.ptr));
    set_counter_block(F->P, F->ctr);
    F->P.nontrivial = static_cast<uint64_t*>(F->nontrivial.ptr);
    F->P.huge = static_cast<int4*>(F->huge.ptr);
    F->P.ntCap = F->ntCap;
    F->P.hugeCap = F->hugeCap;
    if (!rimage.empty()) std::memcpy(F->stageResolve, rimage.data(), rimage.size());
    if (!cv.empty()) std::memcpy(F->stageClear, cv.data(), F->clearBytes);
    return CRX_OK;
}

crx_status frame_capture(crx_frame* F) {
    crx_context* ctx = F->ctx;
    hipStream_t s = F->capture;
    const bool timing = ctx->timing;
    ctx->timing = false;   // no event records inside the graph
    crx_status st = CRX_OK;
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess && F->drawBytes)
        e = hipMemcpyAsync(F->drawMeta.ptr, F->stageDraw, F->drawBytes, hipMemcpyHostToDevice, s);
    if (e == hipSuccess && F->resolveBytes)
        e = hipMemcpyAsync(F->resolveMeta.ptr, F->stageResolve, F->resolveBytes, hipMemcpyHostToDevice, s);
    if (e == hipSuccess && F->clearBytes)
        e = hipMemcpyAsync(F->clearMeta.ptr, F->stageClear, F->clearBytes, hipMemcpyHostToDevice, s);
    // the clear zeroes the frame's counters as well (else stage 1's reset does)
    if (e == hipSuccess && F->numClear)
        e = launch_clear_views(static_cast<const ClearView*>(F->clearMeta.ptr), F->numClear, F->clearBlocks, F->ctr, s);
    if (e == hipSuccess && F->plan.batches) e = enqueue_stages(ctx, F->plan, F->P, false, s, F->numClear > 0);
    if (e == hipSuccess)
        e = hipMemcpyAsync(F->ctrHost, F->ctr, CTR_COUNT * sizeof(uint32_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && F->settings.resolve)
        e = launch_resolve_views(reinterpret_cast<const ResolveParams*>(static_cast<char*>(F->resolveMeta.ptr) +
                                                                        F->resolveOffViews),
                                 F->numViews, F->resolveBlocks, full_resolve((crx_display_attribute)F->settings.attribute), s);
    hipGraph_t g = nullptr;
    const hipError_t e2 = hipStreamEndCapture(s, &g);
    if (e == hipSuccess) e = e2;
    ctx->timing = timing;
    if (e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return fail(CRX_HIP_ERROR, std::string("frame capture: ") + hipGetErrorString(e));
    }
    F->graph = g;
    e = hipGraphInstantiate(&F->exec, F->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) st = fail(CRX_HIP_ERROR, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    return st;
}

}  // namespace

extern "C" {

crx_status crx_frame_destroy(crx_frame* F) {
    if (!F) return CRX_OK;
    (void)hipSetDevice(F->ctx->device);
    if (F->launched && F->done) (void)hipEventSynchronize(F->done);
    if (F->exec) (void)hipGraphExecDestroy(F->exec);
    if (F->graph) (void)hipGraphDestroy(F->graph);
    if (F->capture) (void)hipStreamDestroy(F->capture);
    if (F->done) (void)hipEventDestroy(F->done);
    F->drawMeta.release();
    F->resolveMeta.release();
    F->clearMeta.release();
    F->nontrivial.release();
    F->huge.release();
    if (F->ctr) (void)hipFree(F->ctr);
    if (F->ctrHost) (void)hipHostFree(F->ctrHost);
    if (F->stageDraw) (void)hipHostFree(F->stageDraw);
    if (F->stageResolve) (void)hipHostFree(F->stageResolve);
    if (F->stageClear) (void)hipHostFree(F->stageClear);
    delete F;
    return CRX_OK;
}

crx_status crx_frame_create(crx_context* ctx, crx_scene* scene, const crx_view* views, uint32_t numViews,
                            const crx_render_settings* settings, uint64_t* const* framebuffers,
                            uint64_t* const* colorbuffers, crx_frame** out) {
    if (!ctx || !scene || !views || !framebuffers || !out) return fail(CRX_INVALID_ARG, "NULL argument");
    *out = nullptr;
    if (numViews == 0 || numViews > kMaxViews)
        return fail(CRX_INVALID_ARG, "numViews must be in [1, " + std::to_string(kMaxViews) + "]");
    crx_render_settings rs;
    crx_render_settings_default(&rs);
    if (settings) rs = *settings;
    if (rs.resolve && !colorbuffers) return fail(CRX_INVALID_ARG, "resolve needs colorbuffers");
    for (uint32_t v = 0; v < numViews; v++) {
        if (views[v].width <= 0 || views[v].height <= 0) return fail(CRX_INVALID_ARG, "view size must be positive");
        if (!framebuffers[v]) return fail(CRX_INVALID_ARG, "framebuffer is NULL");
        if (rs.resolve && !colorbuffers[v]) return fail(CRX_INVALID_ARG, "resolve needs a colorbuffer");
    }
    CRX_HIP(hipSetDevice(ctx->device));
    crx_frame* F = new crx_frame();
    F->ctx = ctx;
    F->scene = scene;
    F->settings = rs;
    F->numViews = numViews;
    F->views.assign(views, views + numViews);
    F->fbs.assign(framebuffers, framebuffers + numViews);
    F->cbs.assign(numViews, nullptr);
    if (rs.resolve) F->cbs.assign(colorbuffers, colorbuffers + numViews);
    F->nodes = scene->nodes.size();
    crx_status st = frame_fill(F, false);
    auto hip = [&](hipError_t e, const char* what) {
        if (st == CRX_OK && e != hipSuccess) st = fail(CRX_HIP_ERROR, std::string(what) + ": " + hipGetErrorString(e));
    };
    if (st == CRX_OK) st = F->drawMeta.reserve(std::max<size_t>(F->drawBytes, 256));
    if (st == CRX_OK) st = F->resolveMeta.reserve(std::max<size_t>(F->resolveBytes, 256));
    if (st == CRX_OK) st = F->clearMeta.reserve(std::max<size_t>(F->clearBytes, 256));
    if (st == CRX_OK) {
        F->ntCap = ctx->tuning.max_nontrivial;
        F->hugeCap = ctx->tuning.max_huge;
        st = F->nontrivial.reserve((size_t)F->ntCap * sizeof(uint64_t));
        if (st == CRX_OK) st = F->huge.reserve((size_t)F->hugeCap * sizeof(int4));
    }
    hip(hipMalloc(&F->ctr, CTR_COUNT * sizeof(uint32_t)), "hipMalloc");
    hip(hipHostMalloc((void**)&F->ctrHost, CTR_COUNT * sizeof(uint32_t), hipHostMallocDefault), "hipHostMalloc");
    hip(hipHostMalloc(&F->stageDraw, std::max<size_t>(F->drawBytes, 256), hipHostMallocDefault), "hipHostMalloc");
    hip(hipHostMalloc(&F->stageResolve, std::max<size_t>(F->resolveBytes, 256), hipHostMallocDefault), "hipHostMalloc");
    hip(hipHostMalloc(&F->stageClear, std::max<size_t>(F->clearBytes, 256), hipHostMallocDefault), "hipHostMalloc");
    hip(hipStreamCreateWithFlags(&F->capture, hipStreamNonBlocking), "hipStreamCreate");
    hip(hipEventCreateWithFlags(&F->done, hipEventDisableTiming), "hipEventCreate");
    if (st == CRX_OK) st = frame_fill(F, true);
    if (st == CRX_OK) st = frame_capture(F);
    if (st != CRX_OK) {
        crx_frame_destroy(F);
        return st;
    }
    *out = F;
    return CRX_OK;
}

crx_status crx_frame_set_views(crx_frame* F, const crx_view* views) {
    if (!F || !views) return fail(CRX_INVALID_ARG, "NULL argument");
    for (uint32_t v = 0; v < F->numViews; v++) {
        if (views[v].width != F->views[v].width || views[v].height != F->views[v].height)
            return fail(CRX_INVALID_ARG, "a frame's view sizes are fixed");
    }
    CRX_HIP(hipSetDevice(F->ctx->device));
    if (F->launched) CRX_HIP(hipEventSynchronize(F->done));   // the graph reads the staging buffers
    F->views.assign(views, views + F->numViews);
    return frame_fill(F, true);
}

crx_status crx_frame_launch(crx_frame* F, void* stream) {
    if (!F) return fail(CRX_INVALID_ARG, "frame is NULL");
    CRX_HIP(hipSetDevice(F->ctx->device));
    CRX_HIP(hipGraphLaunch(F->exec, as_stream(stream)));
    CRX_HIP(hipEventRecord(F->done, as_stream(stream)));
    F->launched = true;
    return CRX_OK;
}

crx_status crx_frame_wait(crx_frame* F, crx_draw_stats* stats) {
    if (!F) return fail(CRX_INVALID_ARG, "frame is NULL");
    if (stats) std::memset(stats, 0, sizeof(*stats));
    if (!F->launched) return CRX_OK;
    CRX_HIP(hipEventSynchronize(F->done));
    const uint32_t nNt = F->ctrHost[CTR_NONTRIVIAL];
    const uint32_t nHuge = F->ctrHost[CTR_HUGE];
    if (stats) {
        stats->num_triangles = F->plan.tris * F->numViews;
        stats->num_nontrivial = nNt;
        stats->num_huge = nHuge;
    }
    if (nNt > F->ntCap || nHuge > F->hugeCap)
        return fail(CRX_QUEUE_OVERFLOW, "a stage queue of the frame overflowed (" + std::to_string(nNt) + " / " +
                                             std::to_string(nHuge) + " entries)");
    return CRX_OK;
}

}  // extern "C"
