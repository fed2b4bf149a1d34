// crx_internal.h — launch wrappers shared by the kernels and the host glue.
#pragma once

#include <hip/hip_runtime.h>

#include "crx.h"
#include "crx_device.h"

namespace crx {

// Clears fb (and fb2, when given) to the clear pixel; a non-null ctr counter
// block is zeroed too.
hipError_t launch_clear(uint64_t* fb, uint32_t numPixels, uint32_t* ctr, int numCUs, hipStream_t stream,
                        uint64_t* fb2 = nullptr, uint32_t numPixels2 = 0);
hipError_t launch_reset_counters(uint32_t* ctr, hipStream_t stream);
hipError_t launch_reset_counter_set(const CounterSet& c, hipStream_t stream);
// device-built draw plan from caller-owned CMesh arrays (kernel-level entry points)
hipError_t launch_plan(const crx_cmesh* list, uint32_t numItems, const crx_cmesh* instances,
                       const crx_mat4* transforms, uint32_t numInstances, int instanced, int pu16,
                       uint32_t batchShift, DrawItem* items, InstanceRec* inst, Transform* xf, uint32_t* batchCount,
                       hipStream_t stream);
hipError_t launch_store_view(const ViewParams& v, ViewParams* dst, hipStream_t stream);
// trisPerLane: 1, 2 or 4 (P.batchShift = log2(block * trisPerLane));
// pu16: the draw's meshes are IX_PU16-compressed
// instanceBoxes: the draw repeats meshes over many instances (wave boxes
// let groups of triangles skip instances that cull them all as tiny)
hipError_t launch_stage1(const RasterParams& P, uint32_t block, uint32_t trisPerLane, uint32_t grid, bool pu16,
                         bool instanceBoxes, hipStream_t stream);
// Pairs of consecutive triangles per lane, block 256 (k_stage1_pairs): the
// items carry the strip layout (DrawItem::rowTris ...).
hipError_t launch_stage1_pairs(const RasterParams& P, uint32_t block, uint32_t grid, bool pu16, hipStream_t stream);
hipError_t launch_stage1_views(const RasterParams& P, uint32_t block, uint32_t trisPerLane, uint32_t grid, bool pu16,
                               hipStream_t stream);
hipError_t launch_stage2(const RasterParams& P, uint32_t grid, hipStream_t stream);
hipError_t launch_stage3(const RasterParams& P, uint32_t grid, hipStream_t stream);
hipError_t launch_resolve(const ResolveParams& R, hipStream_t stream);
// views: device array of numViews records; numBlocks = total blocks over all views
// full: the variant with the TEXTURE and STAGE attributes (full_resolve)
hipError_t launch_resolve_views(const ResolveParams* views, uint32_t numViews, uint32_t numBlocks, bool full,
                                hipStream_t stream);
// The display attributes the resolve's full variant shades (the others run
// without their code and its registers).
inline bool full_resolve(crx_display_attribute a) { return a == CRX_DISPLAY_TEXTURE || a == CRX_DISPLAY_STAGE; }
hipError_t launch_clear_views(const ClearView* views, uint32_t numViews, uint32_t numBlocks, uint32_t* ctr,
                              hipStream_t stream);
hipError_t launch_copy(void* dst, const void* src, size_t bytes, int numCUs, hipStream_t stream);
hipError_t launch_read(const void* src, size_t bytes, int numCUs, uint32_t* sink, hipStream_t stream);
hipError_t launch_check_division(uint64_t pairs, int numCUs, unsigned long long* out, hipStream_t stream);
// One mip level: dst (tw x th, tw * th < 2^32) from src (sw x sh), 2x2 box sums / 4
hipError_t launch_mip_level(const uint32_t* src, uint32_t* dst, uint32_t sw, uint32_t sh, uint32_t tw, uint32_t th,
                            uint32_t alpha, int numCUs, hipStream_t stream);
hipError_t launch_valu_probe(uint32_t iters, int mode, uint32_t blocks, uint32_t* sink, hipStream_t stream);
hipError_t launch_atomic_probe(uint64_t* buf, uint32_t logEntries, uint32_t perLane, int scattered, uint32_t blocks,
                               hipStream_t stream);

}  // namespace crx
