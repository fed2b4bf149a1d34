// crx_api.cpp — host side of the C-ABI (include/crx.h).
//
// Builds the compact device records of crx_device.h from the reference's CMesh
// arrays, owns queues / counters / staging per context, launches the kernels of
// crx_kernels.hip and restates the frame preparation of CuRast::draw
// (CuRast_render.h:266-519) for the scene API.  Host float arithmetic that
// feeds the kernels (view*world, inverse(view), mirror flags) follows GLM's
// evaluation order and is compiled without contraction, so it produces the
// same bits as the device or the reference's GLM would under IEEE rules.
#include "crx.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "crx_device.h"
#include "crx_host.h"
#include "crx_internal.h"

using namespace crx;

namespace crx {

namespace {
thread_local std::string g_error;
}  // namespace

crx_status fail(crx_status s, const std::string& msg) {
    g_error = msg;
    return s;
}

}  // namespace crx

namespace crx {

// ---- host math in GLM 0.9.9 order --------------------------------------------

// tmat4x4 * tmat4x4, type_mat4x4.inl:601-616
crx_mat4 mat_mul(const crx_mat4& A, const crx_mat4& B) {
    crx_mat4 R;
    for (int j = 0; j < 4; j++) {
        const float b0 = B.m[4 * j], b1 = B.m[4 * j + 1], b2 = B.m[4 * j + 2], b3 = B.m[4 * j + 3];
        for (int r = 0; r < 4; r++) {
            float acc = A.m[r] * b0 + A.m[4 + r] * b1;
            acc = acc + A.m[8 + r] * b2;
            acc = acc + A.m[12 + r] * b3;
            R.m[4 * j + r] = acc;
        }
    }
    return R;
}

// compute_inverse<tmat4x4>, func_matrix.inl:298-395
crx_mat4 mat_inverse(const crx_mat4& M) {
#define E(c, r) M.m[4 * (c) + (r)]
    const float c00 = E(2,2) * E(3,3) - E(3,2) * E(2,3);
    const float c02 = E(1,2) * E(3,3) - E(3,2) * E(1,3);
    const float c03 = E(1,2) * E(2,3) - E(2,2) * E(1,3);
    const float c04 = E(2,1) * E(3,3) - E(3,1) * E(2,3);
    const float c06 = E(1,1) * E(3,3) - E(3,1) * E(1,3);
    const float c07 = E(1,1) * E(2,3) - E(2,1) * E(1,3);
    const float c08 = E(2,1) * E(3,2) - E(3,1) * E(2,2);
    const float c10 = E(1,1) * E(3,2) - E(3,1) * E(1,2);
    const float c11 = E(1,1) * E(2,2) - E(2,1) * E(1,2);
    const float c12 = E(2,0) * E(3,3) - E(3,0) * E(2,3);
    const float c14 = E(1,0) * E(3,3) - E(3,0) * E(1,3);
    const float c15 = E(1,0) * E(2,3) - E(2,0) * E(1,3);
    const float c16 = E(2,0) * E(3,2) - E(3,0) * E(2,2);
    const float c18 = E(1,0) * E(3,2) - E(3,0) * E(1,2);
    const float c19 = E(1,0) * E(2,2) - E(2,0) * E(1,2);
    const float c20 = E(2,0) * E(3,1) - E(3,0) * E(2,1);
    const float c22 = E(1,0) * E(3,1) - E(3,0) * E(1,1);
    const float c23 = E(1,0) * E(2,1) - E(2,0) * E(1,1);
    const float fac[6][4] = {{c00, c00, c02, c03}, {c04, c04, c06, c07}, {c08, c08, c10, c11},
                             {c12, c12, c14, c15}, {c16, c16, c18, c19}, {c20, c20, c22, c23}};
    const float vec[4][4] = {{E(1,0), E(0,0), E(0,0), E(0,0)}, {E(1,1), E(0,1), E(0,1), E(0,1)},
                             {E(1,2), E(0,2), E(0,2), E(0,2)}, {E(1,3), E(0,3), E(0,3), E(0,3)}};
    crx_mat4 I;
    for (int i = 0; i < 4; i++) {
        const float inv0 = (vec[1][i] * fac[0][i] - vec[2][i] * fac[1][i]) + vec[3][i] * fac[2][i];
        const float inv1 = (vec[0][i] * fac[0][i] - vec[2][i] * fac[3][i]) + vec[3][i] * fac[4][i];
        const float inv2 = (vec[0][i] * fac[1][i] - vec[1][i] * fac[3][i]) + vec[3][i] * fac[5][i];
        const float inv3 = (vec[0][i] * fac[2][i] - vec[1][i] * fac[4][i]) + vec[2][i] * fac[5][i];
        const float sa = (i & 1) ? -1.0f : 1.0f;   // SignA = (+,-,+,-)
        const float sb = -sa;                       // SignB = (-,+,-,+)
        I.m[i] = inv0 * sa;
        I.m[4 + i] = inv1 * sb;
        I.m[8 + i] = inv2 * sa;
        I.m[12 + i] = inv3 * sb;
    }
    const float d0 = E(0,0) * I.m[0], d1 = E(0,1) * I.m[4], d2 = E(0,2) * I.m[8], d3 = E(0,3) * I.m[12];
    const float det = (d0 + d1) + (d2 + d3);
    const float oneOver = 1.0f / det;
    for (int i = 0; i < 16; i++) I.m[i] = I.m[i] * oneOver;
#undef E
    return I;
}

// det(worldView[0..2]) = dot(cross(c0, c1), c2), triangles_visbuffer.cu:179-183
bool mirror_flag(const crx_mat4& m) {
    const f3 c0{m.m[0], m.m[1], m.m[2]};
    const f3 c1{m.m[4], m.m[5], m.m[6]};
    const f3 c2{m.m[8], m.m[9], m.m[10]};
    return dot3(cross3(c0, c1), c2) < 0.0f;
}

crx_status check_tuning(const crx_tuning& t) {
    if (t.small_threshold > 1u << 24) return fail(CRX_INVALID_ARG, "tuning.small_threshold too large");
    if (t.tile_size < 16 || t.tile_size > 256 || (t.tile_size % 16) != 0)
        return fail(CRX_INVALID_ARG, "tuning.tile_size must be a multiple of 16 in [16, 256]");
    if (t.tile_size > 64 && (t.tile_size % 64) != 0)
        return fail(CRX_INVALID_ARG, "tuning.tile_size above 64 must be a multiple of 64");
    if (t.tile_size < 64 && (64 % t.tile_size) != 0)
        return fail(CRX_INVALID_ARG, "tuning.tile_size below 64 must divide 64");
    const uint32_t b = t.stage1_block;
    if (b != 128 && b != 256 && b != 512)
        return fail(CRX_INVALID_ARG, "tuning.stage1_block must be 128, 256 or 512");
    if (t.stage1_cull_order > 3) return fail(CRX_INVALID_ARG, "tuning.stage1_cull_order must be 0, 1, 2 or 3");
    if (t.stage1_fast_tiny > 1) return fail(CRX_INVALID_ARG, "tuning.stage1_fast_tiny must be 0 or 1");
    if (t.stage1_order > 2) return fail(CRX_INVALID_ARG, "tuning.stage1_order must be 0, 1 or 2");
    if (t.stage1_tris_per_lane > 4 || t.stage1_tris_per_lane == 3)
        return fail(CRX_INVALID_ARG, "tuning.stage1_tris_per_lane must be 0, 1, 2 or 4");
    if (t.stage1_layout > 2) return fail(CRX_INVALID_ARG, "tuning.stage1_layout must be 0, 1 or 2");
    if (t.stage1_strip_rows == 0 || t.stage1_strip_rows > 256)
        return fail(CRX_INVALID_ARG, "tuning.stage1_strip_rows must be in [1, 256]");
    if (t.stage1_strip_batches == 0 || t.stage1_strip_batches > 8)
        return fail(CRX_INVALID_ARG, "tuning.stage1_strip_batches must be in [1, 8]");
    if (t.max_nontrivial == 0 || t.max_huge == 0) return fail(CRX_INVALID_ARG, "tuning queue capacities must be > 0");
    if (t.stage2_waves_per_cu == 0 || t.stage3_waves_per_cu == 0 || t.stage2_waves_per_cu > 64 ||
        t.stage3_waves_per_cu > 64 || t.stage1_waves_per_cu > 64)
        return fail(CRX_INVALID_ARG, "tuning waves per CU out of range");
    return CRX_OK;
}

crx_status check_target(const crx_render_target* t) {
    if (!t) return fail(CRX_INVALID_ARG, "target is NULL");
    if (!t->framebuffer) return fail(CRX_INVALID_ARG, "target->framebuffer is NULL");
    if (t->width <= 0 || t->height <= 0) return fail(CRX_INVALID_ARG, "target size must be positive");
    if ((int64_t)t->width * (int64_t)t->height > 0x7fffffffll)
        return fail(CRX_INVALID_ARG, "target has more than 2^31-1 pixels");
    return CRX_OK;
}

crx_status check_mesh(const crx_cmesh& m, const char* what, uint32_t k) {
    // the message is built only on failure: draws check thousands of instances
    const auto name = [&]() { return std::string(what) + "[" + std::to_string(k) + "]"; };
    if (m.numTriangles > 1431655765u) return fail(CRX_INVALID_ARG, name() + " has more than 1431655765 triangles");
    if (m.numTriangles > 0 && !m.positions) return fail(CRX_INVALID_ARG, name() + " has no positions");
    if (m.compressed) {
        // IX_PU16 is an index-buffer format (the _compressed kernels are _indexbuffer ones)
        if (m.numTriangles > 0 && !m.indices) return fail(CRX_INVALID_ARG, name() + " is compressed without indices");
        if (m.bitsPerIndex > 32) return fail(CRX_INVALID_ARG, name() + " has bitsPerIndex > 32");
    }
    // a texture that TEXTURE would sample needs a level 0 of at least one texel
    if (m.texture.data && !m.texture.huffmanTables && (m.texture.width <= 0 || m.texture.height <= 0))
        return fail(CRX_INVALID_ARG, name() + " has a texture of width or height <= 0");
    return CRX_OK;
}

InstanceRec make_instance(const crx_cmesh& m) {
    InstanceRec r;
    r.indices = m.indices;
    r.positions = reinterpret_cast<const float*>(m.positions);
    r.colors = m.colors;
    r.normals = reinterpret_cast<const float*>(m.normals);
    r.uvs = reinterpret_cast<const float*>(m.uvs);
    r.cumm = m.cummulativeTriangleCount;
    r.numTriangles = m.numTriangles;
    r.transformIndex = (uint32_t)m.instances.offset;
    r.id = m.id;
    r.isLoaded = m.isLoaded ? 1u : 0u;
    std::memcpy(r.world, m.world.m, sizeof(r.world));
    r.compressed = m.compressed ? 1u : 0u;
    r.bitsPerIndex = m.bitsPerIndex;
    r.indexMin = m.index_min;
    r.cf[0] = m.compressionFactor.x; r.cf[1] = m.compressionFactor.y; r.cf[2] = m.compressionFactor.z;
    r.amin[0] = m.aabb.min.x; r.amin[1] = m.aabb.min.y; r.amin[2] = m.aabb.min.z;
    r.amax[0] = m.aabb.max.x; r.amax[1] = m.aabb.max.y; r.amax[2] = m.aabb.max.z;
    r.texData = m.texture.data;
    r.texW = m.texture.width;
    r.texH = m.texture.height;
    r.texJpeg = m.texture.huffmanTables ? 1u : 0u;
    return r;
}

crx_status ensure_queues(crx_context* ctx, uint32_t nt, uint32_t hg) {
    if (nt > ctx->ntCap) {
        CRX_TRY(ctx->nontrivial.reserve((size_t)nt * sizeof(uint64_t)));
        ctx->ntCap = nt;
    }
    if (hg > ctx->hugeCap) {
        CRX_TRY(ctx->huge.reserve((size_t)hg * sizeof(int4)));
        ctx->hugeCap = hg;
    }
    return CRX_OK;
}

void record(crx_context* ctx, int slot, hipStream_t stream) {
    if (!ctx->timing) return;
    if (hipEventRecord(ctx->ev[slot], stream) == hipSuccess) ctx->recorded[slot] = true;
}

uint32_t grid_for(const crx_context* ctx, uint32_t wavesPerCU, uint32_t wavesPerBlock) {
    uint64_t g = (uint64_t)ctx->numCUs * wavesPerCU / wavesPerBlock;
    if (g < 8) g = 8;
    return (uint32_t)g;
}

// Points the four stage counters of P into one CTR_COUNT block.
void set_counter_block(RasterParams& P, uint32_t* block) {
    P.ctrNontrivial = block + CTR_NONTRIVIAL;
    P.ctrHuge = block + CTR_HUGE;
    P.cursor2 = block + CTR_CURSOR2;
    P.cursor3 = block + CTR_CURSOR3;
    P.claims = block + CTR_CLAIM;
}

// Per-view record of a draw: size, projection and the view's transforms.
ViewParams make_view_params(const crx_render_target& t, const Transform* transforms) {
    ViewParams V{};
    V.fb = t.framebuffer;
    V.transforms = transforms;
    V.W = t.width;
    V.H = t.height;
    V.Wf = (float)t.width;
    V.Hf = (float)t.height;
    V.Wm1 = (float)(t.width - 1);
    V.Hm1 = (float)(t.height - 1);
    V.f = t.proj.m[5];
    V.aspect = (float)t.width / (float)t.height;
    V.fa = V.f / V.aspect;
    V.faI = 1.0f / (V.f / V.aspect);
    V.fI = 1.0f / V.f;
    // |v_rcp_f32 screen coordinate - IEEE one| <= 0.75*max(W,H)*2^-20 for |ndc| <= 2;
    // the margin is 5x that plus an absolute 2^-6
    V.tinyMargin = (float)std::max(t.width, t.height) * 0x1p-18f + 0x1p-6f;
    return V;
}

// Row length of meshes stored in rows, by their device index buffer: filled
// by crx_mesh_create, read when a draw plans its items (a draw of CMesh
// arrays whose index buffer is not a crx mesh walks contiguously).  The hint
// only orders the stage-1 walk; no result depends on it.
std::mutex g_rowHintsMu;
std::unordered_map<const void*, uint32_t> g_rowHints;

uint32_t row_hint_of(const void* indices) {
    if (!indices) return 0;
    std::lock_guard<std::mutex> lk(g_rowHintsMu);
    auto it = g_rowHints.find(indices);
    return it == g_rowHints.end() ? 0u : it->second;
}

// Triangles per row of a mesh stored in rows (a grid, a lat-long sphere):
// over a window of triangles in the middle of the index buffer, the most
// common distance from a vertex's first use to its first use more than 64
// triangles later (the same vertex in the next row), rounded down to even;
// 0 when no distance covers a quarter of the vertices seen.  In a row of
// quads cut along either diagonal a vertex first appears one triangle
// further along in its second row than in its first, so the distance is the
// row's length plus one; a row of quads holds an even number of triangles,
// and rows of even length keep every row's pairs of consecutive triangles
// on quads.
uint32_t detect_row_triangles(const uint32_t* idx, uint64_t T) {
    const uint64_t W = std::min<uint64_t>(T, 1ull << 18);
    if (W < 1024) return 0;
    const uint64_t t0 = T / 2 - W / 2;
    std::unordered_map<uint32_t, uint64_t> first;
    first.reserve(2 * W);
    std::unordered_map<uint64_t, uint32_t> gaps;
    std::unordered_map<uint32_t, uint8_t> done;
    for (uint64_t t = t0; t < t0 + W; t++) {
        for (int k = 0; k < 3; k++) {
            const uint32_t v = idx[3 * t + k];
            auto f = first.find(v);
            if (f == first.end()) {
                first.emplace(v, t);
            } else if (t - f->second > 64 && !done[v]) {
                done[v] = 1;
                gaps[t - f->second]++;
            }
        }
    }
    uint64_t best = 0;
    uint32_t count = 0;
    for (const auto& g : gaps)
        if (g.second > count || (g.second == count && g.first < best)) { best = g.first; count = g.second; }
    if (count * 4ull < first.size() || best > 0xFFFFFFFFull) return 0;
    return (uint32_t)best & ~1u;
}

// Host records and metadata layout of a draw of numViews >= 1 views sharing
// one instance list (transforms view-major).  No device work.
crx_status plan_draw(crx_context* ctx, const crx_cmesh* meshes, uint32_t numMeshes, const crx_cmesh* instances,
                     uint32_t numInstances, const crx_mat4* transforms, const crx_render_target* targets,
                     uint32_t numViews, int instanced, const uint8_t* skip, DrawPlan* plan, bool instUnchanged,
                     bool prefix) {
    if (!ctx) return fail(CRX_INVALID_ARG, "ctx is NULL");
    if (numViews == 0 || numViews > kMaxViews)
        return fail(CRX_INVALID_ARG, "numViews must be in [1, " + std::to_string(kMaxViews) + "]");
    if (!targets) return fail(CRX_INVALID_ARG, "targets is NULL");
    for (uint32_t v = 0; v < numViews; v++) CRX_TRY(check_target(targets + v));
    DrawPlan& D = *plan;
    {
        // scalars back to their defaults; the record vectors keep their
        // storage (a context plans every draw into the same DrawPlan, and
        // fresh multi-hundred-KB vectors per draw cost page faults)
        DrawPlan fresh;
        fresh.items.swap(D.items);
        fresh.inst.swap(D.inst);
        fresh.xf.swap(D.xf);
        fresh.targets.swap(D.targets);
        D = std::move(fresh);
    }
    D.numViews = numViews;
    D.numInstances = numInstances;
    D.prefix = prefix;
    D.targets.assign(targets, targets + numViews);
    // drawTrianglesVisbuffer returns early for an empty instance list (CuRast_render.h:231)
    if (numInstances == 0) return CRX_OK;
    if (!instances || !transforms) return fail(CRX_INVALID_ARG, "instances / transforms is NULL");
    if (instanced && (!meshes || numMeshes == 0)) return fail(CRX_INVALID_ARG, "instanced draw without meshes");
    if (numViews > 1 && numInstances > kInstanceMask + 1u)
        return fail(CRX_INVALID_ARG, "a multi-view draw takes at most 2^24 instances");
    const crx_tuning& tu = ctx->tuning;
    const crx_cmesh* list = instanced ? meshes : instances;
    D.numItems = instanced ? numMeshes : numInstances;
    bool manyInstances = false;
    for (uint32_t k = 0; instanced && k < D.numItems; k++)
        manyInstances = manyInstances || list[k].instances.count >= (int32_t)kBoxInstances;
    // the draw's stage-1 kernel is chosen by its first mesh, as the reference
    // picks the _compressed / _uncompressed kernels (CuRast_render.h:235-241)
    D.pu16 = D.numItems > 0 && list[0].compressed;
    const S1Schedule sched = s1_schedule(tu, manyInstances, D.pu16, numViews > 1);
    D.block = tu.stage1_block;
    D.trisPerLane = sched.trisPerLane;
    D.batchShift = 0;   // log2(triangles per batch)
    while ((1u << D.batchShift) < D.block * D.trisPerLane) D.batchShift++;
    const uint64_t batchTris = 1ull << D.batchShift;
    // pairs of consecutive triangles per lane (k_stage1_pairs), for draws
    // large enough that every wave of the grid claims many 128-triangle
    // batches (a small draw keeps the short per-batch walk of k_stage1).  An
    // IX_PU16 draw takes them only with stage1_tris_per_lane set to 2: its
    // sweep is bound by the decode's VALU work, which the pairs do not cut
    // (1.98 ms against 1.91 ms for k_stage1 on the 400 M-triangle heightfield)
    D.runs = tu.stage1_layout != 0 && numViews == 1 && !manyInstances && D.block == 256 && D.trisPerLane == 2 &&
             (!D.pu16 || tu.stage1_tris_per_lane == 2);
    if (D.runs) {
        uint64_t tris = 0;
        for (uint32_t k = 0; k < D.numItems; k++) tris += list[k].numTriangles;
        const uint64_t waves = (uint64_t)grid_for(ctx, sched.wavesPerCU ? sched.wavesPerCU : 32u, D.block / 64) *
                               (D.block / 64);
        D.runs = tris >= 16ull * 128ull * waves;
    }

    D.items.resize(D.numItems);
    for (uint32_t k = 0; k < D.numItems; k++) {
        const crx_cmesh& m = list[k];
        CRX_TRY(check_mesh(m, instanced ? "meshes" : "instances", k));
        if ((m.compressed != 0) != D.pu16)
            return fail(CRX_UNSUPPORTED, "a draw mixes compressed and uncompressed meshes");
        const int64_t off = m.instances.offset;
        const int64_t cnt = instanced ? (int64_t)m.instances.count : 1;
        if (off < 0 || cnt < 0 || off + cnt > (int64_t)numInstances)
            return fail(CRX_INVALID_ARG, "instances{offset,count} of item " + std::to_string(k) + " out of range");
        DrawItem& it = D.items[k];
        it = DrawItem{};
        // a non-indexed item reads its positions in place of an index buffer
        // (a valid address of at least 36 bytes per triangle) and masks the
        // words off, so stage 1 issues the same loads for both kinds
        it.indices = m.indices ? m.indices : reinterpret_cast<const uint32_t*>(m.positions);
        it.indexMask = m.indices ? 0xFFFFFFFFu : 0u;
        it.positions = reinterpret_cast<const float*>(m.positions);
        it.cumm = m.cummulativeTriangleCount;
        it.numTriangles = m.numTriangles;
        it.instOffset = (uint32_t)off;
        it.instCount = (uint32_t)cnt;
        if (cnt >= (int64_t)kBoxInstances) D.instanceBoxes = true;
        it.firstBatch = (uint32_t)D.batches;
        it.bitsPerIndex = m.bitsPerIndex;
        it.indexMin = m.index_min;
        it.cf[0] = m.compressionFactor.x; it.cf[1] = m.compressionFactor.y; it.cf[2] = m.compressionFactor.z;
        it.amin[0] = m.aabb.min.x; it.amin[1] = m.aabb.min.y; it.amin[2] = m.aabb.min.z;
        it.numVertices = m.numVertices;
        // pair layout (batches of one wave's 128 triangles): units of
        // stage1_strip_rows rows of a strip of stage1_strip_batches batches
        // of the mesh's rows (a mesh stored in rows of at least eight
        // batches), else of rows of 128 triangles (a contiguous walk)
        constexpr uint64_t waveTris = 128;
        const uint32_t rowHint = (D.runs && tu.stage1_layout == 2) ? row_hint_of(m.indices) : 0u;
        const bool strips = rowHint >= 8 * waveTris;
        const uint64_t stripTris = waveTris * (strips ? tu.stage1_strip_batches : 1u);
        it.rowTris = strips ? rowHint : (uint32_t)waveTris;
        it.unitCols = strips ? tu.stage1_strip_batches : 1u;
        it.strips = strips ? (uint32_t)((rowHint + stripTris - 1) / stripTris) : 1u;
        it.unitRows = tu.stage1_strip_rows;
        const uint64_t rows = ((uint64_t)m.numTriangles + it.rowTris - 1) / it.rowTris;
        it.unitSegs = (uint32_t)((rows + it.unitRows - 1) / it.unitRows);
        // the groups of a strip rounded up to a multiple of 8 (the units past
        // the last row are empty): unit s * unitSegs + g then runs on the XCD
        // of its group g for every strip s, so the waves that draw the two
        // sides of a strip edge share their XCD's L2
        if (strips) it.unitSegs = (it.unitSegs + 7u) & ~7u;
        const uint64_t units = (uint64_t)it.strips * it.unitSegs;
        D.batches += D.runs ? units : ((uint64_t)m.numTriangles + batchTris - 1) >> D.batchShift;
        D.tris += (uint64_t)m.numTriangles * (uint64_t)cnt;
        if (D.batches >= (1ull << 32)) return fail(CRX_INVALID_ARG, "too many triangle batches for one draw");
    }
    // instUnchanged: the list is the one the plan's records were built (and
    // checked) from, so they stand
    if (!instUnchanged || D.inst.size() != numInstances) {
        D.inst.resize(numInstances);
        for (uint32_t i = 0; i < numInstances; i++) {
            CRX_TRY(check_mesh(instances[i], "instances", i));
            const int64_t ti = instances[i].instances.offset;
            if (ti < 0 || ti >= (int64_t)numInstances)
                return fail(CRX_INVALID_ARG, "instances[" + std::to_string(i) + "].instances.offset out of range");
            D.inst[i] = make_instance(instances[i]);
        }
    }
    const size_t numXf = (size_t)numViews * numInstances;
    D.xf.resize(numXf);
    for (size_t i = 0; i < numXf; i++) {
        Transform& x = D.xf[i];
        std::memcpy(x.m, transforms[i].m, sizeof(x.m));
        x.flip = mirror_flag(transforms[i]) ? 1u : 0u;
        x.skip = (skip && skip[i]) ? 1u : 0u;
        x.pad[0] = x.pad[1] = 0u;
    }
    D.offItems = 0;
    D.offInst = align_up(D.offItems + D.items.size() * sizeof(DrawItem), 256);
    D.offPrefix = align_up(D.offInst + D.inst.size() * sizeof(InstanceRec), 256);
    D.offViews = align_up(D.offPrefix + (prefix ? D.inst.size() * sizeof(uint64_t) : 0), 256);
    D.offXf = align_up(D.offViews + (size_t)numViews * sizeof(ViewParams), 256);
    D.metaBytes = D.offXf + D.xf.size() * sizeof(Transform);
    if (sched.wavesPerCU == 0) {
        D.g1 = (uint32_t)std::min<uint64_t>(D.batches, 1ull << 31);
    } else {
        D.g1 = grid_for(ctx, sched.wavesPerCU, D.block / 64);
        if (D.g1 > D.batches) D.g1 = (uint32_t)D.batches;
    }
    D.g1 = (D.g1 + 7u) & ~7u;   // the stage-1 XCD split assumes a multiple of 8 workgroups
    D.g2 = grid_for(ctx, tu.stage2_waves_per_cu, 4);
    D.g3 = grid_for(ctx, tu.stage3_waves_per_cu, 4);
    return CRX_OK;
}

// Writes the plan's metadata image (to be copied to device address `dev`)
// into `host` and returns the kernel parameters; the queues are the
// context's.  Without writeInst the instance records and prefix are left out
// of the image (the device already holds them).
RasterParams plan_write(crx_context* ctx, const DrawPlan& D, char* host, char* dev, bool writeInst) {
    const Transform* xfDev = reinterpret_cast<const Transform*>(dev + D.offXf);
    std::vector<ViewParams> views(D.numViews);
    for (uint32_t v = 0; v < D.numViews; v++)
        views[v] = make_view_params(D.targets[v], xfDev + (size_t)v * D.numInstances);
    std::memcpy(host + D.offItems, D.items.data(), D.items.size() * sizeof(DrawItem));
    if (writeInst) {
        std::memcpy(host + D.offInst, D.inst.data(), D.inst.size() * sizeof(InstanceRec));
        if (D.prefix) {
            uint64_t* prefix = reinterpret_cast<uint64_t*>(host + D.offPrefix);
            for (size_t i = 0; i < D.inst.size(); i++) prefix[i] = D.inst[i].cumm;
        }
    }
    std::memcpy(host + D.offViews, views.data(), views.size() * sizeof(ViewParams));
    std::memcpy(host + D.offXf, D.xf.data(), D.xf.size() * sizeof(Transform));
    const crx_tuning& tu = ctx->tuning;
    RasterParams P{};
    P.items = reinterpret_cast<const DrawItem*>(dev + D.offItems);
    P.inst = reinterpret_cast<const InstanceRec*>(dev + D.offInst);
    P.views = reinterpret_cast<const ViewParams*>(dev + D.offViews);
    P.numViews = D.numViews;
    set_counter_block(P, ctx->ctr);
    P.numItems = D.numItems;
    P.numBatches = (uint32_t)D.batches;
    P.smallThreshold = tu.small_threshold;
    P.tile = (int32_t)tu.tile_size;
    P.batchShift = D.batchShift;
    P.order = tu.stage1_order;
    // cull order 3 decides per item; a draw of several views tests backfaces
    // first (10.70 ms against 11.07 ms for frustum first on 64 views)
    P.cullOrder = (tu.stage1_cull_order == 3u && D.numViews > 1) ? 0u : tu.stage1_cull_order;
    P.fastTiny = tu.stage1_fast_tiny;
    P.numWaves2 = D.g2 * 4u;
    P.numWaves3 = D.g3 * 4u;
    P.nontrivial = static_cast<uint64_t*>(ctx->nontrivial.ptr);
    P.huge = static_cast<int4*>(ctx->huge.ptr);
    P.ntCap = ctx->ntCap;
    P.hugeCap = ctx->hugeCap;
    return P;
}

// Counter reset and the three stages.  viewsKernel forces the multi-view
// stage 1 (which reads every view record from memory) for a single view.
hipError_t enqueue_stages(crx_context* ctx, const DrawPlan& D, const RasterParams& P, bool viewsKernel,
                          hipStream_t stream, bool countersZeroed) {
    hipError_t e = hipSuccess;
    if (!countersZeroed) {
        e = launch_reset_counters(P.ctrNontrivial - CTR_NONTRIVIAL, stream);   // the plan's counter block
        if (e != hipSuccess) return e;
    }
    record(ctx, T_S1, stream);
    if (D.numViews == 1 && !viewsKernel && D.runs)
        e = launch_stage1_pairs(P, D.block, D.g1, D.pu16, stream);
    else if (D.numViews == 1 && !viewsKernel)
        e = launch_stage1(P, D.block, D.trisPerLane, D.g1, D.pu16, D.instanceBoxes, stream);
    else e = launch_stage1_views(P, D.block, D.trisPerLane, D.g1, D.pu16, stream);
    if (e != hipSuccess) return e;
    record(ctx, T_S2, stream);
    e = launch_stage2(P, D.g2, stream);
    if (e != hipSuccess) return e;
    record(ctx, T_S3, stream);
    e = launch_stage3(P, D.g3, stream);
    if (e != hipSuccess) return e;
    record(ctx, T_S3_END, stream);
    return hipSuccess;
}

// Instance records (and prefix) from which a draw that finds them already in
// the metadata buffer copies the rest in two pieces rather than all in one.
constexpr size_t kSplitUploadBytes = 64 * 1024;

// Orders `stream` after the last draw's work that may still read the
// context's metadata buffers (a resolve enqueued before an early return on
// another stream).
crx_status wait_meta_free(crx_context* ctx, hipStream_t stream) {
    if (ctx->metaFreePending && stream != ctx->metaFreeStream) CRX_HIP(hipStreamWaitEvent(stream, ctx->metaFree, 0));
    return CRX_OK;
}

// drawTrianglesVisbuffer (CuRast_render.h:211-264) for numViews >= 1 views
// sharing one instance list: transforms[v * numInstances + i] is view v's
// view * world of instance i.  One view runs the plain stage 1; several views
// run the batched stage 1, which fetches each triangle once for all views.
crx_status draw_views(crx_context* ctx, const crx_cmesh* meshes, uint32_t numMeshes, const crx_cmesh* instances,
                      uint32_t numInstances, const crx_mat4* transforms, const crx_render_target* targets,
                      uint32_t numViews, int instanced, hipStream_t stream, crx_draw_stats* stats,
                      const DrawExtras& ex) {
    if (stats) std::memset(stats, 0, sizeof(*stats));
    if (!ctx) return fail(CRX_INVALID_ARG, "ctx is NULL");
    if (ex.skip && numViews < 2) return fail(CRX_INVALID_ARG, "per-view instance culling needs a multi-view draw");
    DrawPlan& D = ctx->plan;
    // the plan's instance records stand when the list is the one they were
    // last built from; the key is void until this draw's records are in place
    const bool sameLists = ex.listsKey != 0 && ex.listsKey == ctx->metaListsKey;
    ctx->metaListsKey = 0;
    CRX_TRY(plan_draw(ctx, meshes, numMeshes, instances, numInstances, transforms, targets, numViews, instanced,
                      ex.skip, &D, sameLists, ex.prefix));
    if (stats) stats->num_triangles = D.tris * numViews;
    if (D.batches == 0 && !ex.afterStages) return CRX_OK;
    CRX_HIP(hipSetDevice(ctx->device));
    CRX_TRY(wait_meta_free(ctx, stream));

    RasterParams P{};
    if (D.numInstances > 0) {
        CRX_TRY(ctx->drawMeta.reserve(D.metaBytes));
        CRX_TRY(ctx->drawStaging.acquire(D.metaBytes));
        char* meta = static_cast<char*>(ctx->drawMeta.ptr);
        char* image = static_cast<char*>(ctx->drawStaging.ptr);
        // records (and prefix) already in the buffer at the same place: copy
        // only the draw items and the per-view records and transforms (two
        // copies, so only when the records skipped outweigh a copy launch)
        const bool held = sameLists && ctx->metaListsGen == ctx->drawMeta.generation && ctx->metaListsBegin == D.offInst &&
                          ctx->metaListsEnd == D.offViews && ctx->metaListsPrefix == D.prefix &&
                          D.offViews - D.offInst >= kSplitUploadBytes;
        P = plan_write(ctx, D, image, meta, !held);
        if (held) {
            CRX_HIP(hipMemcpyAsync(meta, image, D.offInst, hipMemcpyHostToDevice, stream));
            CRX_HIP(hipMemcpyAsync(meta + D.offViews, image + D.offViews, D.metaBytes - D.offViews,
                                   hipMemcpyHostToDevice, stream));
        } else {
            CRX_HIP(hipMemcpyAsync(meta, image, D.metaBytes, hipMemcpyHostToDevice, stream));
        }
        CRX_TRY(ctx->drawStaging.mark(stream));
        ctx->metaListsKey = ex.listsKey;
        ctx->metaListsGen = ctx->drawMeta.generation;
        ctx->metaListsBegin = D.offInst;
        ctx->metaListsEnd = D.offViews;
        ctx->metaListsPrefix = D.prefix;
    }
    const auto finish = [&]() -> crx_status {
        if (ex.earlyReturn && ex.afterStages) {
            CRX_HIP(hipEventRecord(ctx->metaFree, stream));
            ctx->metaFreePending = true;
            ctx->metaFreeStream = stream;
        }
        return CRX_OK;
    };
    if (D.batches == 0) {
        // nothing to draw; the work that follows still runs (a resolve
        // writes every pixel)
        CRX_TRY((*ex.afterStages)());
        if (!ex.earlyReturn) CRX_HIP(hipStreamSynchronize(stream));
        return finish();
    }
    CRX_TRY(ensure_queues(ctx, ctx->tuning.max_nontrivial, ctx->tuning.max_huge));

    uint32_t redraws = 0;
    for (;;) {
        P.nontrivial = static_cast<uint64_t*>(ctx->nontrivial.ptr);
        P.huge = static_cast<int4*>(ctx->huge.ptr);
        P.ntCap = ctx->ntCap;
        P.hugeCap = ctx->hugeCap;
        // the caller's clear zeroed the counters for the first pass only
        CRX_HIP(enqueue_stages(ctx, D, P, false, stream, ex.countersZeroed && redraws == 0));
        CRX_HIP(hipMemcpyAsync(ctx->ctrHost, ctx->ctr, CTR_COUNT * sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
        CRX_HIP(hipEventRecord(ctx->countersBack, stream));
        // work that follows the stages (the resolve) is prepared and enqueued
        // while they run; after a redraw it is enqueued again
        if (ex.afterStages) CRX_TRY((*ex.afterStages)());
        if (ex.earlyReturn) CRX_HIP(hipEventSynchronize(ctx->countersBack));
        else CRX_HIP(hipStreamSynchronize(stream));
        const uint32_t nNt = ctx->ctrHost[CTR_NONTRIVIAL];
        const uint32_t nHuge = ctx->ctrHost[CTR_HUGE];
        if (stats) {
            stats->num_nontrivial = nNt;
            stats->num_huge = nHuge;
            stats->redraws = redraws;
            stats->stage1_kernel = D.numViews > 1 ? 1u : D.runs ? 2u : 0u;
        }
        if (nNt <= ctx->ntCap && nHuge <= ctx->hugeCap) break;
        // Queue overflow: grow and draw again.  Every value the truncated pass
        // wrote is also written by the full pass, so the minimum is unchanged.
        if (++redraws > 8) return fail(CRX_QUEUE_OVERFLOW, "stage queues kept overflowing");
        const uint64_t ntWant = std::max<uint64_t>(ctx->ntCap, (uint64_t)nNt + nNt / 4 + 1024);
        uint64_t hgWant = ctx->hugeCap;
        if (nHuge > ctx->hugeCap || nNt > ctx->ntCap) hgWant = std::max<uint64_t>(hgWant, (uint64_t)nHuge * 2 + 1024);
        if (ntWant > 0xffffffffull || hgWant > 0xffffffffull)
            return fail(CRX_QUEUE_OVERFLOW, "stage queue would exceed 2^32 entries");
        CRX_TRY(ensure_queues(ctx, (uint32_t)ntWant, (uint32_t)hgWant));
    }
    return finish();
}

// Per-view resolve record (the instance list and prefix are shared).
ResolveParams make_resolve_params(const crx_render_target& t, const InstanceRec* inst, const uint64_t* prefix,
                                  uint32_t numInstances, crx_display_attribute attribute) {
    ResolveParams R{};
    R.inst = inst;
    R.prefix = prefix;
    R.fb = t.framebuffer;
    R.cb = t.colorbuffer;
    R.numInstances = numInstances;
    R.W = t.width;
    R.H = t.height;
    R.attribute = (int32_t)attribute;
    R.Wf = (float)t.width;
    R.Hf = (float)t.height;
    R.p00I = 1.0f / t.proj.m[0];
    R.p11I = 1.0f / t.proj.m[5];
    const crx_mat4 viewI = mat_inverse(t.view);
    std::memcpy(R.viewI, viewI.m, sizeof(R.viewI));
    const f3 o = xform(viewI.m, f3{0.0f, 0.0f, 0.0f}, 1.0f);
    R.origin[0] = o.x;
    R.origin[1] = o.y;
    R.origin[2] = o.z;
    std::memcpy(R.view, t.view.m, sizeof(R.view));
    R.f = t.proj.m[5];
    R.aspect = (float)t.width / (float)t.height;
    return R;
}

// kernel_resolve_visbuffer_to_colorbuffer2D (resolve.cu:640-1141) for
// numViews >= 1 targets sharing one instance list, in one launch.
// Metadata image of a resolve: InstanceRec[n] | prefix[n] | ResolveParams[numViews],
// with device pointers for base address `dev`.
// Layout of a resolve image: instance records | prefix | per-view records.
size_t resolve_image_layout(uint32_t numInstances, uint32_t numViews, size_t* offPrefix, size_t* offViews) {
    *offPrefix = align_up((size_t)numInstances * sizeof(InstanceRec), 256);
    *offViews = align_up(*offPrefix + (size_t)numInstances * sizeof(uint64_t) + 8, 256);
    return *offViews + (size_t)numViews * sizeof(ResolveParams);
}

// Writes the resolve image of device base `dev` into `out` (host memory of
// resolve_image_layout bytes): the records go straight to their places, with
// no intermediate copies (a scene of thousands of instances makes this a
// memory-bound loop on the host).
crx_status check_resolve(const crx_render_target* targets, uint32_t numViews, crx_display_attribute attribute) {
    if (numViews == 0 || numViews > kMaxViews)
        return fail(CRX_INVALID_ARG, "numViews must be in [1, " + std::to_string(kMaxViews) + "]");
    if (!targets) return fail(CRX_INVALID_ARG, "targets is NULL");
    for (uint32_t v = 0; v < numViews; v++) {
        CRX_TRY(check_target(targets + v));
        if (!targets[v].colorbuffer) return fail(CRX_INVALID_ARG, "target colorbuffer is NULL");
    }
    if ((int)attribute < 0 || (int)attribute > 7) return fail(CRX_INVALID_ARG, "unknown display attribute");
    return CRX_OK;
}

// First 64x4 block of each view in a multi-view resolve grid; returns the total.
crx_status resolve_blocks(const crx_render_target* targets, uint32_t numViews, ResolveParams* rv, uint32_t* blocks) {
    uint64_t nb = 0;
    for (uint32_t v = 0; v < numViews; v++) {
        rv[v].firstBlock = (uint32_t)nb;
        nb += (uint64_t)((targets[v].width + 63) / 64) * (uint64_t)((targets[v].height + 3) / 4);
    }
    if (nb >= (1ull << 31)) return fail(CRX_INVALID_ARG, "too many pixels for one resolve");
    *blocks = (uint32_t)nb;
    return CRX_OK;
}

crx_status resolve_image_into(const crx_cmesh* instances, uint32_t numInstances, const uint64_t* prefix,
                              const crx_render_target* targets, uint32_t numViews, crx_display_attribute attribute,
                              char* dev, char* out, size_t* offViews, uint32_t* blocks, bool ownedOnly = false) {
    CRX_TRY(check_resolve(targets, numViews, attribute));
    if (numInstances > 0 && (!instances || !prefix)) return fail(CRX_INVALID_ARG, "instances / prefix is NULL");
    size_t offPrefix = 0;
    resolve_image_layout(numInstances, numViews, &offPrefix, offViews);
    InstanceRec* inst = reinterpret_cast<InstanceRec*>(out);
    for (uint32_t i = 0; i < numInstances; i++) {
        CRX_TRY(check_mesh(instances[i], "instances", i));
        inst[i] = make_instance(instances[i]);
    }
    if (numInstances) std::memcpy(out + offPrefix, prefix, (size_t)numInstances * sizeof(uint64_t));
    const InstanceRec* instDev = reinterpret_cast<const InstanceRec*>(dev);
    const uint64_t* prefixDev = reinterpret_cast<const uint64_t*>(dev + offPrefix);
    ResolveParams* rv = reinterpret_cast<ResolveParams*>(out + *offViews);
    for (uint32_t v = 0; v < numViews; v++) {
        rv[v] = make_resolve_params(targets[v], instDev, prefixDev, numInstances, attribute);
        rv[v].ownedOnly = ownedOnly ? 1u : 0u;
    }
    return resolve_blocks(targets, numViews, rv, blocks);
}

crx_status resolve_image(const crx_cmesh* instances, uint32_t numInstances, const uint64_t* prefix,
                         const crx_render_target* targets, uint32_t numViews, crx_display_attribute attribute,
                         char* dev, std::vector<char>* image, size_t* offViews, uint32_t* blocks) {
    size_t offPrefix = 0;
    image->resize(resolve_image_layout(numInstances, numViews, &offPrefix, offViews));
    return resolve_image_into(instances, numInstances, prefix, targets, numViews, attribute, dev, image->data(),
                              offViews, blocks);
}

// kernel_resolve_visbuffer_to_colorbuffer2D (resolve.cu:640-1141) for
// numViews >= 1 targets sharing one instance list, in one launch.
crx_status resolve_views(crx_context* ctx, const crx_cmesh* instances, uint32_t numInstances, const uint64_t* prefix,
                         const crx_render_target* targets, uint32_t numViews, crx_display_attribute attribute,
                         hipStream_t stream, bool ownedOnly) {
    if (!ctx) return fail(CRX_INVALID_ARG, "ctx is NULL");
    CRX_HIP(hipSetDevice(ctx->device));
    CRX_TRY(wait_meta_free(ctx, stream));
    size_t offPrefix = 0, offViews = 0;
    uint32_t blocks = 0;
    const size_t bytes = resolve_image_layout(numInstances, numViews, &offPrefix, &offViews);
    CRX_TRY(ctx->resolveMeta.reserve(bytes));
    char* meta = static_cast<char*>(ctx->resolveMeta.ptr);
    // the image is written straight into the page-locked staging buffer
    CRX_TRY(ctx->resolveStaging.acquire(bytes));
    char* image = static_cast<char*>(ctx->resolveStaging.ptr);
    CRX_TRY(resolve_image_into(instances, numInstances, prefix, targets, numViews, attribute, meta, image, &offViews,
                               &blocks, ownedOnly));
    const ResolveParams* rv = reinterpret_cast<const ResolveParams*>(image + offViews);
    if (numInstances || numViews > 1) {
        CRX_HIP(hipMemcpyAsync(meta, image, bytes, hipMemcpyHostToDevice, stream));
        CRX_TRY(ctx->resolveStaging.mark(stream));
    }
    record(ctx, T_RESOLVE, stream);
    if (numViews == 1) {
        CRX_HIP(launch_resolve(rv[0], stream));
    } else {
        CRX_HIP(launch_resolve_views(reinterpret_cast<const ResolveParams*>(meta + offViews), numViews, blocks,
                                     full_resolve(attribute), stream));
    }
    record(ctx, T_RESOLVE_END, stream);
    return CRX_OK;
}

crx_status resolve_views_device(crx_context* ctx, const InstanceRec* instDev, const uint64_t* prefixDev,
                                uint32_t numInstances, const crx_render_target* targets, uint32_t numViews,
                                crx_display_attribute attribute, hipStream_t stream) {
    if (!ctx) return fail(CRX_INVALID_ARG, "ctx is NULL");
    CRX_TRY(check_resolve(targets, numViews, attribute));
    if (numInstances > 0 && (!instDev || !prefixDev)) return fail(CRX_INVALID_ARG, "instances / prefix is NULL");
    CRX_HIP(hipSetDevice(ctx->device));
    if (numViews == 1) {
        // one view: its record is the kernel argument, nothing to upload
        const ResolveParams R = make_resolve_params(targets[0], instDev, prefixDev, numInstances, attribute);
        record(ctx, T_RESOLVE, stream);
        CRX_HIP(launch_resolve(R, stream));
        record(ctx, T_RESOLVE_END, stream);
        return CRX_OK;
    }
    CRX_TRY(wait_meta_free(ctx, stream));
    const size_t bytes = (size_t)numViews * sizeof(ResolveParams);
    CRX_TRY(ctx->resolveMeta.reserve(bytes));
    CRX_TRY(ctx->resolveStaging.acquire(bytes));
    ResolveParams* rv = static_cast<ResolveParams*>(ctx->resolveStaging.ptr);
    for (uint32_t v = 0; v < numViews; v++)
        rv[v] = make_resolve_params(targets[v], instDev, prefixDev, numInstances, attribute);
    uint32_t blocks = 0;
    CRX_TRY(resolve_blocks(targets, numViews, rv, &blocks));
    CRX_HIP(hipMemcpyAsync(ctx->resolveMeta.ptr, rv, bytes, hipMemcpyHostToDevice, stream));
    CRX_TRY(ctx->resolveStaging.mark(stream));
    record(ctx, T_RESOLVE, stream);
    CRX_HIP(launch_resolve_views(static_cast<const ResolveParams*>(ctx->resolveMeta.ptr), numViews, blocks,
                                 full_resolve(attribute), stream));
    record(ctx, T_RESOLVE_END, stream);
    return CRX_OK;
}

// Records of a multi-buffer clear (empty buffers skipped) and its block count.
crx_status clear_records(uint64_t* const* buffers, const uint32_t* numPixels, uint32_t n, std::vector<ClearView>* cv,
                         uint32_t* blocks) {
    cv->clear();
    uint64_t nb = 0;
    for (uint32_t k = 0; k < n; k++) {
        if (numPixels[k] == 0) continue;
        if (!buffers[k]) return fail(CRX_INVALID_ARG, "buffer to clear is NULL");
        ClearView c;
        c.fb = buffers[k];
        c.numPixels = numPixels[k];
        c.firstBlock = (uint32_t)nb;
        cv->push_back(c);
        nb += (numPixels[k] + kClearChunk - 1) / kClearChunk;
    }
    if (nb >= (1ull << 31)) return fail(CRX_INVALID_ARG, "too many pixels for one clear");
    *blocks = (uint32_t)nb;
    return CRX_OK;
}

// kernel_clearFramebuffer (resolve.cu:213-232) over several buffers in one launch.
crx_status clear_buffers(crx_context* ctx, uint64_t* const* buffers, const uint32_t* numPixels, uint32_t n,
                         hipStream_t stream, uint32_t* ctr, bool* zeroed) {
    if (zeroed) *zeroed = false;
    std::vector<ClearView> cv;
    uint32_t blocks = 0;
    CRX_TRY(clear_records(buffers, numPixels, n, &cv, &blocks));
    if (cv.empty()) return CRX_OK;
    if (cv.size() <= 2) {
        // one or two buffers (a view's visibility and colour buffers): one
        // launch, with no metadata to stage and copy
        const bool two = cv.size() == 2;
        CRX_HIP(launch_clear(cv[0].fb, cv[0].numPixels, ctr, ctx->numCUs, stream, two ? cv[1].fb : nullptr,
                             two ? cv[1].numPixels : 0u));
        if (zeroed) *zeroed = ctr != nullptr;
        return CRX_OK;
    }
    const size_t bytes = cv.size() * sizeof(ClearView);
    CRX_TRY(ctx->clearMeta.reserve(bytes));
    CRX_TRY(ctx->clearStaging.acquire(bytes));
    std::memcpy(ctx->clearStaging.ptr, cv.data(), bytes);
    CRX_HIP(hipMemcpyAsync(ctx->clearMeta.ptr, ctx->clearStaging.ptr, bytes, hipMemcpyHostToDevice, stream));
    CRX_TRY(ctx->clearStaging.mark(stream));
    CRX_HIP(launch_clear_views(static_cast<const ClearView*>(ctx->clearMeta.ptr), (uint32_t)cv.size(), blocks, ctr,
                               stream));
    if (zeroed) *zeroed = ctr != nullptr;
    return CRX_OK;
}

// SNTriangles::update (SNTriangles.h:29-105): the four side planes of
// proj * view, normalised, as (nx, ny, nz, d) x 4.
void frustum_planes(const crx_view& view, float out[16]) {
    const crx_mat4 viewProj = mat_mul(view.proj, view.view);
    for (int p = 0; p < 4; p++) {
        const int r = p < 2 ? 0 : 1;
        const float sg = (p & 1) ? -1.0f : 1.0f;
        float q[4];
        for (int c = 0; c < 4; c++) {
            const float r3 = viewProj.m[4 * c + 3], rr = viewProj.m[4 * c + r];
            q[c] = sg > 0.0f ? r3 + rr : r3 - rr;
        }
        const f3 n{q[0], q[1], q[2]};
        const float len = sqrtf(dot3(n, n));
        out[4 * p + 0] = n.x / len;
        out[4 * p + 1] = n.y / len;
        out[4 * p + 2] = n.z / len;
        out[4 * p + 3] = q[3] / len;
    }
}

// Box3::transform: world-space AABB of the 8 transformed corners.
crx_box3 world_box(const crx_box3& b, const crx_mat4& world) {
    const float cx[2] = {b.min.x, b.max.x}, cy[2] = {b.min.y, b.max.y}, cz[2] = {b.min.z, b.max.z};
    f3 wmin{INFINITY, INFINITY, INFINITY}, wmax{-INFINITY, -INFINITY, -INFINITY};
    for (int k = 0; k < 8; k++) {
        const f3 w = xform(world.m, f3{cx[k & 1], cy[(k >> 1) & 1], cz[(k >> 2) & 1]}, 1.0f);
        wmin.x = std::fmin(wmin.x, w.x); wmin.y = std::fmin(wmin.y, w.y); wmin.z = std::fmin(wmin.z, w.z);
        wmax.x = std::fmax(wmax.x, w.x); wmax.y = std::fmax(wmax.y, w.y); wmax.z = std::fmax(wmax.z, w.z);
    }
    return crx_box3{{wmin.x, wmin.y, wmin.z}, {wmax.x, wmax.y, wmax.z}};
}

// Node frustum test of SNTriangles::update: visible unless the box lies
// entirely on the negative side of one plane (positive-vertex test).
bool box_in_frustum(const float planes[16], const crx_box3& wb) {
    for (int p = 0; p < 4; p++) {
        const float* pl = planes + 4 * p;
        const f3 n{pl[0], pl[1], pl[2]};
        const f3 c{n.x >= 0.0f ? wb.max.x : wb.min.x, n.y >= 0.0f ? wb.max.y : wb.min.y,
                   n.z >= 0.0f ? wb.max.z : wb.min.z};
        if (dot3(n, c) + pl[3] < 0.0f) return false;
    }
    return true;
}

// A content key for a scene's instance list, new on every rebuild and unique
// across scenes.
uint64_t next_lists_key() {
    static std::atomic<uint64_t> key{0};
    return ++key;
}

// CuRast::draw's frame preparation (CuRast_render.h:338-435) for numViews
// views: the node cull of each view (SNTriangles::update, SNTriangles.h:29-105),
// the nodes kept (visible in any view, or all with keepAll), sorted by
// triangle count then mesh with node order kept inside a mesh, grouped into
// unique meshes, with triangle prefix, view-major transforms and per-view
// skip flags for the nodes a view's own cull dropped.  Fills the scene's
// scratch vectors; returns the number of unique meshes.
uint32_t scene_lists(crx_scene* scene, const crx_view* views, uint32_t numViews, const crx_render_settings& settings,
                     bool keepAll) {
    std::vector<float> planes((size_t)numViews * 16);
    for (uint32_t v = 0; v < numViews; v++) frustum_planes(views[v], &planes[(size_t)v * 16]);
    const uint32_t N = (uint32_t)scene->nodes.size();
    std::vector<uint8_t> vis((size_t)N * numViews, settings.frustum_culling ? 0 : 1);
    std::vector<uint8_t> any(N, settings.frustum_culling ? 0 : 1);
    if (settings.frustum_culling) {
        // each view's cull through the range hierarchy (each node gets its
        // own box test's answer), rebuilt when the nodes changed
        if (scene->treeVersion != scene->version) {
            scene->wboxes.resize(N);
            for (uint32_t i = 0; i < N; i++) scene->wboxes[i] = scene->nodes[i].wbox;
            scene->tree.build(scene->wboxes.data(), N);
            scene->treeVersion = scene->version;
        }
        std::vector<uint32_t> in;
        for (uint32_t v = 0; v < numViews; v++) {
            scene->tree.visible(&planes[(size_t)v * 16], scene->wboxes.data(), in);
            for (uint32_t i : in) {
                vis[(size_t)i * numViews + v] = 1;
                any[i] = 1;
            }
        }
    }
    std::vector<uint32_t> kept;
    kept.reserve(N);
    for (uint32_t i = 0; i < N; i++)
        if (any[i] || keepAll) kept.push_back(i);
    // CuRast_render.h:338-345: numTriangles descending, then mesh identity;
    // instances of one mesh keep node order.
    const auto before = [&](uint32_t a, uint32_t b) {
        const crx_mesh* A = scene->nodes[a].mesh;
        const crx_mesh* B = scene->nodes[b].mesh;
        if (A->numTriangles == B->numTriangles) return A->seq < B->seq;
        return A->numTriangles > B->numTriangles;
    };
    // scenes are usually built mesh by mesh: then the order already holds
    if (!std::is_sorted(kept.begin(), kept.end(), before)) std::stable_sort(kept.begin(), kept.end(), before);
    const uint32_t n = (uint32_t)kept.size();
    const int instancedFlag = settings.instanced ? 1 : 0;
    // the per-node records depend on the node set, the instancing switch and
    // the kept nodes only: a frame that changes none of them keeps them
    if (scene->listsKey == 0 || scene->listsVersion != scene->version || scene->listsInstanced != instancedFlag ||
        scene->listsAttrEpoch != scene->ctx->attrEpoch || scene->listsKept != kept) {
        scene->unique.resize(n);
        scene->instances.resize(n);
        scene->prefix.resize(n);
        scene->nodeOf.resize(n);
        uint64_t sum = 0;
        uint32_t u = 0;
        for (uint32_t i = 0; i < n; i++) {
            const crx_scene::Node& nd = scene->nodes[kept[i]];
            crx_cmesh cm;
            std::memset(&cm, 0, sizeof(cm));
            cm.numTriangles = nd.mesh->numTriangles;
            cm.indices = nd.mesh->indices;
            cm.positions = reinterpret_cast<crx_vec3*>(nd.mesh->positions);
            cm.colors = nd.mesh->colors;
            cm.uvs = reinterpret_cast<crx_vec2*>(nd.mesh->uvs);
            cm.normals = reinterpret_cast<crx_vec3*>(nd.mesh->normals);
            cm.texture = nd.texture;
            cm.aabb = nd.mesh->aabb;
            cm.world = nd.world;
            cm.id = (int32_t)kept[i];
            cm.numVertices = nd.mesh->numVertices;
            cm.address = nd.mesh->seq;
            cm.isLoaded = 1;
            cm.flipTriangles = mirror_flag(nd.world) ? 1 : 0;
            cm.compressed = nd.mesh->compressed ? 1 : 0;
            cm.bitsPerIndex = nd.mesh->bitsPerIndex;
            cm.index_min = nd.mesh->indexMin;
            cm.index_max = nd.mesh->indexMax;
            cm.compressionFactor = nd.mesh->compressionFactor;
            cm.cummulativeTriangleCount = sum;
            cm.instances.offset = (int32_t)i;
            cm.instances.count = 0;
            scene->prefix[i] = sum;
            scene->nodeOf[i] = kept[i];
            sum += cm.numTriangles;
            scene->instances[i] = cm;
            if (u == 0 || scene->unique[u - 1].address != cm.address || !settings.instanced) {
                scene->unique[u] = cm;
                u++;
            }
            scene->unique[u - 1].instances.count++;
        }
        scene->unique.resize(u);
        scene->listsVersion = scene->version;
        scene->listsAttrEpoch = scene->ctx->attrEpoch;
        scene->listsInstanced = instancedFlag;
        scene->listsKept.swap(kept);
        scene->listsKey = next_lists_key();
    }
    // per view: view * world of each listed node and the view's own cull
    scene->transforms.resize((size_t)n * numViews);
    scene->skip.assign((size_t)n * numViews, 0);
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t node = scene->nodeOf[i];
        const crx_mat4& world = scene->nodes[node].world;
        for (uint32_t v = 0; v < numViews; v++) {
            scene->transforms[(size_t)v * n + i] = mat_mul(views[v].view, world);
            scene->skip[(size_t)v * n + i] = vis[(size_t)node * numViews + v] ? 0 : 1;
        }
    }
    return (uint32_t)scene->unique.size();
}

// Render targets for views (framebuffers / colourbuffers per view) and the
// list of buffers a frame clears.
void view_targets(const crx_view* views, uint32_t numViews, uint64_t* const* framebuffers,
                  uint64_t* const* colorbuffers, bool resolve, std::vector<crx_render_target>* targets,
                  std::vector<uint64_t*>* clearBufs, std::vector<uint32_t>* clearPx) {
    targets->resize(numViews);
    clearBufs->clear();
    clearPx->clear();
    for (uint32_t v = 0; v < numViews; v++) {
        crx_render_target& t = (*targets)[v];
        std::memset(&t, 0, sizeof(t));
        t.framebuffer = framebuffers[v];
        t.colorbuffer = resolve ? colorbuffers[v] : nullptr;
        t.width = views[v].width;
        t.height = views[v].height;
        t.view = views[v].view;
        t.viewI = mat_inverse(views[v].view);
        t.proj = views[v].proj;
        const f3 cam = xform(t.viewI.m, f3{0.0f, 0.0f, 0.0f}, 1.0f);
        t.cameraPos = crx_vec3{cam.x, cam.y, cam.z};
        t.f = views[v].proj.m[5];
        t.aspect = (float)views[v].width / (float)views[v].height;
        const uint32_t px = (uint32_t)views[v].width * (uint32_t)views[v].height;
        clearBufs->push_back(t.framebuffer);
        clearPx->push_back(px);
        if (resolve) {
            clearBufs->push_back(t.colorbuffer);
            clearPx->push_back(px);
        }
    }
}

// CuRast::render / CuRast::draw (CuRast_render.h:266-519, :873-904) for
// numViews >= 1 views: node culling (SNTriangles::update, SNTriangles.h:29-105)
// keeps a node visible in any view, the visible nodes are sorted and grouped
// once, and every view is drawn from that one instance list.
crx_status render_views(crx_context* ctx, crx_scene* scene, const crx_view* views, uint32_t numViews,
                        const crx_render_settings* settings_, uint64_t* const* framebuffers,
                        uint64_t* const* colorbuffers, hipStream_t stream, crx_draw_stats* stats) {
    if (!ctx || !scene || !views || !framebuffers) return fail(CRX_INVALID_ARG, "NULL argument");
    if (numViews == 0 || numViews > kMaxViews)
        return fail(CRX_INVALID_ARG, "numViews must be in [1, " + std::to_string(kMaxViews) + "]");
    crx_render_settings settings;
    crx_render_settings_default(&settings);
    if (settings_) settings = *settings_;
    if (settings.resolve && !colorbuffers) return fail(CRX_INVALID_ARG, "resolve needs colorbuffers");
    for (uint32_t v = 0; v < numViews; v++) {
        if (views[v].width <= 0 || views[v].height <= 0) return fail(CRX_INVALID_ARG, "view size must be positive");
        if (!framebuffers[v]) return fail(CRX_INVALID_ARG, "framebuffer is NULL");
        if (settings.resolve && !colorbuffers[v]) return fail(CRX_INVALID_ARG, "resolve needs a colorbuffer");
    }

    const uint32_t u = scene_lists(scene, views, numViews, settings, false);
    const uint32_t n = (uint32_t)scene->instances.size();

    std::vector<crx_render_target> targets;
    std::vector<uint64_t*> clearBufs;
    std::vector<uint32_t> clearPx;
    view_targets(views, numViews, framebuffers, colorbuffers, settings.resolve != 0, &targets, &clearBufs, &clearPx);

    CRX_HIP(hipSetDevice(ctx->device));
    record(ctx, T_CLEAR, stream);
    // the clear also zeroes the draw counters (stage 1 then needs no reset)
    bool zeroed = false;
    CRX_TRY(clear_buffers(ctx, clearBufs.data(), clearPx.data(), (uint32_t)clearBufs.size(), stream, ctx->ctr,
                          &zeroed));
    bool anySkip = false;
    for (uint8_t k : scene->skip) anySkip = anySkip || k != 0;
    // the resolve reads the instance records and prefix the draw uploaded
    const std::function<crx_status()> resolve = [&]() {
        const DrawPlan& D = ctx->plan;
        const char* meta = static_cast<const char*>(ctx->drawMeta.ptr);
        return resolve_views_device(ctx, n ? reinterpret_cast<const InstanceRec*>(meta + D.offInst) : nullptr,
                                    n ? reinterpret_cast<const uint64_t*>(meta + D.offPrefix) : nullptr, n,
                                    targets.data(), numViews, (crx_display_attribute)settings.attribute, stream);
    };
    DrawExtras ex;
    ex.skip = (anySkip && numViews > 1) ? scene->skip.data() : nullptr;
    ex.afterStages = settings.resolve ? &resolve : nullptr;
    ex.listsKey = scene->listsKey;
    ex.prefix = settings.resolve != 0;
    ex.earlyReturn = true;
    ex.countersZeroed = zeroed;
    CRX_TRY(draw_views(ctx, scene->unique.data(), u, scene->instances.data(), n, scene->transforms.data(),
                       targets.data(), numViews, settings.instanced, stream, stats, ex));
    return CRX_OK;
}

}  // namespace crx

extern "C" {

int crx_abi_version(void) { return CRX_ABI_VERSION; }

const char* crx_status_string(crx_status s) {
    switch (s) {
        case CRX_OK: return "CRX_OK";
        case CRX_INVALID_ARG: return "CRX_INVALID_ARG";
        case CRX_OUT_OF_MEMORY: return "CRX_OUT_OF_MEMORY";
        case CRX_QUEUE_OVERFLOW: return "CRX_QUEUE_OVERFLOW";
        case CRX_HIP_ERROR: return "CRX_HIP_ERROR";
        case CRX_NO_DEVICE: return "CRX_NO_DEVICE";
        case CRX_UNSUPPORTED: return "CRX_UNSUPPORTED";
        case CRX_REJECTED: return "CRX_REJECTED";
    }
    return "CRX_UNKNOWN";
}

const char* crx_last_error(void) { return g_error.c_str(); }

void crx_tuning_default(crx_tuning* out) {
    if (!out) return;
    out->small_threshold = 128;       // THRESHOLD_SMALL
    out->tile_size = 64;              // TILE_SIZE
    out->max_nontrivial = 5000000;    // MAX_NONTRIVIAL_TRIANGLES
    out->max_huge = 5000000;          // MAX_HUGE_TRIANGLES
    out->stage1_block = 256;          // TRIANGLES_PER_SWEEP
    out->stage1_waves_per_cu = 32;
    out->stage2_waves_per_cu = 16;
    out->stage3_waves_per_cu = 16;
    out->stage1_cull_order = 3;
    out->stage1_order = 1;
    out->stage1_fast_tiny = 0;          // the exact path divides cheaply (short sequences)
    out->stage1_tris_per_lane = 0;
    out->stage1_layout = 2;
    out->stage1_strip_rows = 8;
    out->stage1_strip_batches = 1;
}

crx_status crx_context_create(int device, const crx_tuning* tuning, crx_context** out) {
    if (!out) return fail(CRX_INVALID_ARG, "out is NULL");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(CRX_NO_DEVICE, "no HIP device available");
    if (device < 0 || device >= n) return fail(CRX_INVALID_ARG, "device index out of range");
    crx_tuning t;
    crx_tuning_default(&t);
    if (tuning) t = *tuning;
    CRX_TRY(check_tuning(t));
    CRX_HIP(hipSetDevice(device));
    int cus = 0;
    CRX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    crx_context* ctx = new crx_context();
    ctx->device = device;
    ctx->numCUs = cus;
    ctx->tuning = t;
    hipError_t e = hipMalloc(&ctx->ctr, CTR_COUNT * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemset(ctx->ctr, 0, CTR_COUNT * sizeof(uint32_t));
    if (e == hipSuccess) e = hipHostMalloc((void**)&ctx->ctrHost, CTR_COUNT * sizeof(uint32_t), hipHostMallocDefault);
    for (int i = 0; e == hipSuccess && i < T_COUNT; i++) e = hipEventCreate(&ctx->ev[i]);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->countersBack, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->metaFree, hipEventDisableTiming);
    if (e != hipSuccess) {
        crx_context_destroy(ctx);
        return fail(CRX_HIP_ERROR, std::string("context allocation: ") + hipGetErrorString(e));
    }
    const crx_status ss = noalloc_scratch_init(device);
    if (ss != CRX_OK) {
        crx_context_destroy(ctx);
        return ss;
    }
    *out = ctx;
    return CRX_OK;
}

crx_status crx_context_destroy(crx_context* ctx) {
    if (!ctx) return CRX_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipDeviceSynchronize();
    ctx->drawStaging.release();
    ctx->resolveStaging.release();
    ctx->clearStaging.release();
    ctx->nontrivial.release();
    ctx->huge.release();
    ctx->drawMeta.release();
    ctx->resolveMeta.release();
    ctx->clearMeta.release();
    ctx->argsMeta.release();
    if (ctx->argsUse) (void)hipEventDestroy(ctx->argsUse);
    if (ctx->ctr) (void)hipFree(ctx->ctr);
    if (ctx->ctrHost) (void)hipHostFree(ctx->ctrHost);
    for (int i = 0; i < T_COUNT; i++)
        if (ctx->ev[i]) (void)hipEventDestroy(ctx->ev[i]);
    if (ctx->countersBack) (void)hipEventDestroy(ctx->countersBack);
    if (ctx->metaFree) (void)hipEventDestroy(ctx->metaFree);
    delete ctx;
    return CRX_OK;
}

crx_status crx_context_set_tuning(crx_context* ctx, const crx_tuning* tuning) {
    if (!ctx || !tuning) return fail(CRX_INVALID_ARG, "NULL argument");
    CRX_TRY(check_tuning(*tuning));
    ctx->tuning = *tuning;
    return CRX_OK;
}

crx_status crx_context_get_tuning(const crx_context* ctx, crx_tuning* out) {
    if (!ctx || !out) return fail(CRX_INVALID_ARG, "NULL argument");
    *out = ctx->tuning;
    return CRX_OK;
}

crx_status crx_context_num_cus(const crx_context* ctx, int* out) {
    if (!ctx || !out) return fail(CRX_INVALID_ARG, "NULL argument");
    *out = ctx->numCUs;
    return CRX_OK;
}

crx_status crx_context_enable_timing(crx_context* ctx, int enable) {
    if (!ctx) return fail(CRX_INVALID_ARG, "ctx is NULL");
    ctx->timing = enable != 0;
    for (bool& r : ctx->recorded) r = false;
    return CRX_OK;
}

crx_status crx_context_timings(crx_context* ctx, crx_timings* out) {
    if (!ctx || !out) return fail(CRX_INVALID_ARG, "NULL argument");
    auto span = [&](int a, int b) -> float {
        if (!ctx->recorded[a] || !ctx->recorded[b]) return -1.0f;
        if (hipEventSynchronize(ctx->ev[b]) != hipSuccess) return -1.0f;
        float ms = -1.0f;
        if (hipEventElapsedTime(&ms, ctx->ev[a], ctx->ev[b]) != hipSuccess) return -1.0f;
        return ms;
    };
    out->clear_ms = span(T_CLEAR, T_S1);
    out->stage1_ms = span(T_S1, T_S2);
    out->stage2_ms = span(T_S2, T_S3);
    out->stage3_ms = span(T_S3, T_S3_END);
    out->pipeline_ms = span(T_S1, T_S3_END);
    out->resolve_ms = span(T_RESOLVE, T_RESOLVE_END);
    return CRX_OK;
}

crx_status crx_device_alloc(crx_context* ctx, size_t bytes, void** out) {
    if (!ctx || !out) return fail(CRX_INVALID_ARG, "NULL argument");
    *out = nullptr;
    CRX_HIP(hipSetDevice(ctx->device));
    hipError_t e = hipMalloc(out, bytes ? bytes : 1);
    if (e != hipSuccess) {
        *out = nullptr;
        return fail(CRX_OUT_OF_MEMORY, "hipMalloc(" + std::to_string(bytes) + "): " + hipGetErrorString(e));
    }
    return CRX_OK;
}

crx_status crx_device_free(crx_context* ctx, void* ptr) {
    if (!ctx) return fail(CRX_INVALID_ARG, "ctx is NULL");
    if (ptr) CRX_HIP(hipFree(ptr));
    return CRX_OK;
}

crx_status crx_memcpy_htod(crx_context* ctx, void* dst, const void* src, size_t bytes, void* stream) {
    if (!ctx || (!dst && bytes) || (!src && bytes)) return fail(CRX_INVALID_ARG, "NULL argument");
    if (bytes == 0) return CRX_OK;
    CRX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, as_stream(stream)));
    CRX_HIP(hipStreamSynchronize(as_stream(stream)));
    return CRX_OK;
}

crx_status crx_memcpy_dtoh(crx_context* ctx, void* dst, const void* src, size_t bytes, void* stream) {
    if (!ctx || (!dst && bytes) || (!src && bytes)) return fail(CRX_INVALID_ARG, "NULL argument");
    if (bytes == 0) return CRX_OK;
    CRX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, as_stream(stream)));
    CRX_HIP(hipStreamSynchronize(as_stream(stream)));
    return CRX_OK;
}

crx_status crx_stream_synchronize(crx_context* ctx, void* stream) {
    if (!ctx) return fail(CRX_INVALID_ARG, "ctx is NULL");
    CRX_HIP(hipStreamSynchronize(as_stream(stream)));
    return CRX_OK;
}

crx_status crx_copy_buffer(crx_context* ctx, void* dst, const void* src, size_t bytes, void* stream) {
    if (!ctx || !dst || !src) return fail(CRX_INVALID_ARG, "NULL argument");
    if ((bytes % 16) != 0 || ((uintptr_t)dst % 16) != 0 || ((uintptr_t)src % 16) != 0)
        return fail(CRX_INVALID_ARG, "crx_copy_buffer needs 16-byte aligned pointers and size");
    CRX_HIP(hipSetDevice(ctx->device));
    CRX_HIP(launch_copy(dst, src, bytes, ctx->numCUs, as_stream(stream)));
    return CRX_OK;
}

crx_status crx_check_division(crx_context* ctx, uint64_t pairs, uint64_t* rcpMismatches, uint64_t* divMismatches,
                              uint64_t* smallNumeratorMismatches) {
    if (!ctx || !rcpMismatches || !divMismatches || !smallNumeratorMismatches)
        return fail(CRX_INVALID_ARG, "NULL argument");
    CRX_HIP(hipSetDevice(ctx->device));
    unsigned long long* d = nullptr;
    CRX_HIP(hipMalloc(&d, 3 * sizeof(unsigned long long)));
    unsigned long long h[3] = {0ull, 0ull, 0ull};
    hipError_t e = hipMemcpy(d, h, sizeof(h), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_check_division(pairs, ctx->numCUs, d, nullptr);
    if (e == hipSuccess) e = hipMemcpy(h, d, sizeof(h), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(CRX_HIP_ERROR, std::string("crx_check_division: ") + hipGetErrorString(e));
    *rcpMismatches = h[0];
    *divMismatches = h[1];
    *smallNumeratorMismatches = h[2];
    return CRX_OK;
}

crx_status crx_probe_read(crx_context* ctx, const void* src, size_t bytes, uint32_t* sink, void* stream) {
    if (!ctx || !src || !sink) return fail(CRX_INVALID_ARG, "NULL argument");
    if ((bytes % 16) != 0 || ((uintptr_t)src % 16) != 0)
        return fail(CRX_INVALID_ARG, "crx_probe_read needs a 16-byte aligned pointer and size");
    CRX_HIP(hipSetDevice(ctx->device));
    CRX_HIP(launch_read(src, bytes, ctx->numCUs, sink, as_stream(stream)));
    return CRX_OK;
}

crx_status crx_probe_atomic_min(crx_context* ctx, uint64_t* buffer, uint32_t logEntries, uint32_t atomicsPerLane,
                                int scattered, uint64_t* issued, void* stream) {
    if (!ctx || !buffer) return fail(CRX_INVALID_ARG, "NULL argument");
    if (logEntries < 6 || logEntries > 30) return fail(CRX_INVALID_ARG, "logEntries must be in [6, 30]");
    CRX_HIP(hipSetDevice(ctx->device));
    const uint32_t blocks = (uint32_t)ctx->numCUs * 8u;
    CRX_HIP(launch_atomic_probe(buffer, logEntries, atomicsPerLane, scattered, blocks, as_stream(stream)));
    if (issued) *issued = (uint64_t)blocks * 256ull * atomicsPerLane;
    return CRX_OK;
}

crx_status crx_probe_valu(crx_context* ctx, uint32_t iters, int mode, uint32_t* sink, uint64_t* waves,
                          void* stream) {
    if (!ctx || !sink) return fail(CRX_INVALID_ARG, "NULL argument");
    if (mode < 0 || mode > 5)
        return fail(CRX_INVALID_ARG, "mode must be 0 (fp32 fma), 1 (int32 add/xor), 2 (fma, 32 lanes), 3 (fma, 1 lane), "
                                     "4 (fma and int interleaved), 5 (fma and int waves alternating)");
    CRX_HIP(hipSetDevice(ctx->device));
    const uint32_t blocks = (uint32_t)ctx->numCUs * 8u;
    CRX_HIP(launch_valu_probe(iters, mode, blocks, sink, as_stream(stream)));
    if (waves) *waves = (uint64_t)blocks * 4ull;
    return CRX_OK;
}

namespace {

// Last level crx_texture_compute_mipmap writes: the reference computes levels
// 1 .. L-1 with L = floor(log2(max(w, h))) (textureTools.cu:47, :59); levels up
// to 7, the deepest TEXTURE samples, are always computed.
uint32_t mip_last_level(int32_t width, int32_t height) {
    const uint32_t m = (uint32_t)std::max(width, height);
    const uint32_t L = 31u - (uint32_t)__builtin_clz(m);
    return std::max(L, 8u) - 1u;
}

crx_status check_texture_size(int32_t width, int32_t height) {
    if (width <= 0 || height <= 0) return fail(CRX_INVALID_ARG, "texture width and height must be positive");
    if ((uint64_t)width * (uint64_t)height > 0x7fffffffull)
        return fail(CRX_INVALID_ARG, "texture has more than 2^31 - 1 texels");
    return CRX_OK;
}

}  // namespace

crx_status crx_texture_chain_texels(int32_t width, int32_t height, uint64_t* out) {
    if (!out) return fail(CRX_INVALID_ARG, "out is NULL");
    CRX_TRY(check_texture_size(width, height));
    uint64_t total = 0, w = (uint64_t)width, h = (uint64_t)height;
    for (uint32_t level = 0; level <= mip_last_level(width, height); level++) {
        total += w * h;
        w = (w + 1) / 2;
        h = (h + 1) / 2;
    }
    *out = total;
    return CRX_OK;
}

crx_status crx_texture_compute_mipmap(crx_context* ctx, uint32_t* data, int32_t width, int32_t height, void* stream) {
    if (!ctx || !data) return fail(CRX_INVALID_ARG, "NULL argument");
    CRX_TRY(check_texture_size(width, height));
    CRX_HIP(hipSetDevice(ctx->device));
    uint32_t sw = (uint32_t)width, sh = (uint32_t)height;
    uint64_t srcOffset = 0;
    for (uint32_t level = 1; level <= mip_last_level(width, height); level++) {
        const uint32_t tw = (sw + 1) / 2, th = (sh + 1) / 2;
        const uint64_t dstOffset = srcOffset + (uint64_t)sw * sh;
        // alpha: the reference starts each texel from a per-level debug colour
        // (0x00RRGGBB for levels 1-9, 0xff0000ff beyond) and overwrites RGB
        const uint32_t alpha = level >= 10 ? 0xffu : 0u;
        CRX_HIP(launch_mip_level(data + srcOffset, data + dstOffset, sw, sh, tw, th, alpha, ctx->numCUs,
                                 as_stream(stream)));
        srcOffset = dstOffset;
        sw = tw;
        sh = th;
    }
    return CRX_OK;
}

crx_status crx_clear_framebuffer(uint64_t* framebuffer, uint32_t numPixels, void* stream) {
    if (!framebuffer && numPixels) return fail(CRX_INVALID_ARG, "framebuffer is NULL");
    if (numPixels == 0) return CRX_OK;
    int dev = 0, cus = 0;
    CRX_HIP(hipGetDevice(&dev));
    CRX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    CRX_HIP(launch_clear(framebuffer, numPixels, nullptr, cus, as_stream(stream)));
    return CRX_OK;
}

crx_status crx_draw_triangles_visbuffer(crx_context* ctx, const crx_cmesh* meshes, uint32_t numMeshes,
                                        const crx_cmesh* instances, uint32_t numInstances,
                                        const crx_mat4* transforms, const crx_render_target* target,
                                        int instanced, void* stream, crx_draw_stats* stats) {
    if (!target) return fail(CRX_INVALID_ARG, "target is NULL");
    return draw_views(ctx, meshes, numMeshes, instances, numInstances, transforms, target, 1, instanced,
                      as_stream(stream), stats);
}

crx_status crx_draw_triangles_visbuffer_views(crx_context* ctx, const crx_cmesh* meshes, uint32_t numMeshes,
                                              const crx_cmesh* instances, uint32_t numInstances,
                                              const crx_mat4* transforms, const crx_render_target* targets,
                                              uint32_t numViews, int instanced, void* stream, crx_draw_stats* stats) {
    return draw_views(ctx, meshes, numMeshes, instances, numInstances, transforms, targets, numViews, instanced,
                      as_stream(stream), stats);
}

crx_status crx_resolve_visbuffer(crx_context* ctx, const crx_cmesh* instances, uint32_t numInstances,
                                 const uint64_t* prefix, const crx_render_target* target,
                                 crx_display_attribute attribute, void* stream) {
    if (!target) return fail(CRX_INVALID_ARG, "target is NULL");
    return resolve_views(ctx, instances, numInstances, prefix, target, 1, attribute, as_stream(stream));
}

crx_status crx_resolve_visbuffer_views(crx_context* ctx, const crx_cmesh* instances, uint32_t numInstances,
                                       const uint64_t* prefix, const crx_render_target* targets, uint32_t numViews,
                                       crx_display_attribute attribute, void* stream) {
    return resolve_views(ctx, instances, numInstances, prefix, targets, numViews, attribute, as_stream(stream));
}

crx_status crx_clear_views(crx_context* ctx, const crx_render_target* targets, uint32_t numViews, int colorbuffers,
                           void* stream) {
    if (!ctx) return fail(CRX_INVALID_ARG, "ctx is NULL");
    if (numViews > 0 && !targets) return fail(CRX_INVALID_ARG, "targets is NULL");
    std::vector<uint64_t*> bufs;
    std::vector<uint32_t> px;
    for (uint32_t v = 0; v < numViews; v++) {
        CRX_TRY(check_target(targets + v));
        const uint32_t n = (uint32_t)targets[v].width * (uint32_t)targets[v].height;
        bufs.push_back(targets[v].framebuffer);
        px.push_back(n);
        if (colorbuffers) {
            if (!targets[v].colorbuffer) return fail(CRX_INVALID_ARG, "target colorbuffer is NULL");
            bufs.push_back(targets[v].colorbuffer);
            px.push_back(n);
        }
    }
    CRX_HIP(hipSetDevice(ctx->device));
    return clear_buffers(ctx, bufs.data(), px.data(), (uint32_t)bufs.size(), as_stream(stream));
}

// ---- camera / matrix helpers ---------------------------------------------

void crx_camera_projection(float near_, float fovy, float aspect, crx_mat4* out) {
    if (!out) return;
    // Camera::createProjectionMatrix, VKRenderer.h:141-148
    const float f = 1.0f / std::tan(fovy / 2.0f);
    const float m[16] = {f / aspect, 0.0f, 0.0f, 0.0f, 0.0f, f, 0.0f, 0.0f,
                         0.0f, 0.0f, 0.0f, -1.0f, 0.0f, 0.0f, near_, 0.0f};
    std::memcpy(out->m, m, sizeof(m));
}

void crx_camera_look_at(const float eye[3], const float center[3], const float up[3], crx_mat4* out) {
    if (!out || !eye || !center || !up) return;
    // glm::lookAtRH (gtc/matrix_transform.inl)
    const f3 e{eye[0], eye[1], eye[2]};
    const f3 f = normalize3(sub3(f3{center[0], center[1], center[2]}, e));
    const f3 s = normalize3(cross3(f, f3{up[0], up[1], up[2]}));
    const f3 u = cross3(s, f);
    float* m = out->m;
    m[0] = s.x; m[4] = s.y; m[8] = s.z;
    m[1] = u.x; m[5] = u.y; m[9] = u.z;
    m[2] = -f.x; m[6] = -f.y; m[10] = -f.z;
    m[3] = 0.0f; m[7] = 0.0f; m[11] = 0.0f;
    m[12] = -dot3(s, e);
    m[13] = -dot3(u, e);
    m[14] = dot3(f, e);
    m[15] = 1.0f;
}

void crx_mat4_inverse(const crx_mat4* m, crx_mat4* out) {
    if (m && out) *out = mat_inverse(*m);
}

void crx_mat4_mul(const crx_mat4* a, const crx_mat4* b, crx_mat4* out) {
    if (a && b && out) *out = mat_mul(*a, *b);
}

// ---- scene API --------------------------------------------------------------

crx_status crx_mesh_create(crx_context* ctx, const float* positions, uint32_t numVertices,
                           const uint32_t* indices, uint32_t numTriangles, const uint32_t* colors,
                           crx_mesh** out) {
    if (!ctx || !out) return fail(CRX_INVALID_ARG, "NULL argument");
    *out = nullptr;
    if (numVertices > 0 && !positions) return fail(CRX_INVALID_ARG, "positions is NULL");
    if (numTriangles > 1431655765u) return fail(CRX_INVALID_ARG, "more than 1431655765 triangles");
    if (!indices && (uint64_t)numVertices < 3ull * numTriangles)
        return fail(CRX_INVALID_ARG, "non-indexed mesh needs 3*numTriangles vertices");
    if (indices) {
        for (uint64_t i = 0; i < 3ull * numTriangles; i++) {
            if (indices[i] >= numVertices) return fail(CRX_INVALID_ARG, "index " + std::to_string(i) + " out of range");
        }
    }
    CRX_HIP(hipSetDevice(ctx->device));
    crx_mesh* m = new crx_mesh();
    m->ctx = ctx;
    m->owned = true;
    m->numVertices = numVertices;
    m->numTriangles = numTriangles;
    m->seq = ++ctx->meshSeq;
    crx_box3 box{{INFINITY, INFINITY, INFINITY}, {-INFINITY, -INFINITY, -INFINITY}};
    for (uint32_t v = 0; v < numVertices; v++) {
        const float* p = positions + 3ull * v;
        box.min.x = std::fmin(box.min.x, p[0]); box.max.x = std::fmax(box.max.x, p[0]);
        box.min.y = std::fmin(box.min.y, p[1]); box.max.y = std::fmax(box.max.y, p[1]);
        box.min.z = std::fmin(box.min.z, p[2]); box.max.z = std::fmax(box.max.z, p[2]);
    }
    m->aabb = box;
    auto upload = [&](void** dst, const void* src, size_t bytes) -> crx_status {
        if (!src || bytes == 0) return CRX_OK;
        hipError_t e = hipMalloc(dst, bytes);
        if (e != hipSuccess) { *dst = nullptr; return fail(CRX_OUT_OF_MEMORY, std::string("hipMalloc: ") + hipGetErrorString(e)); }
        e = hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice);
        if (e != hipSuccess) return fail(CRX_HIP_ERROR, std::string("hipMemcpy: ") + hipGetErrorString(e));
        return CRX_OK;
    };
    crx_status s = upload((void**)&m->positions, positions, 12ull * numVertices);
    if (s == CRX_OK && indices) s = upload((void**)&m->indices, indices, 12ull * numTriangles);
    if (s == CRX_OK && colors) s = upload((void**)&m->colors, colors, 4ull * numVertices);
    if (s != CRX_OK) {
        crx_mesh_destroy(m);
        return s;
    }
    if (indices && m->indices) {
        m->rowTris = detect_row_triangles(indices, numTriangles);
        if (m->rowTris) {
            std::lock_guard<std::mutex> lk(g_rowHintsMu);
            g_rowHints[m->indices] = m->rowTris;
        }
    }
    *out = m;
    return CRX_OK;
}

crx_status crx_compress_mesh(const float* positions, uint32_t numVertices, const uint32_t* indices,
                             uint32_t numTriangles, uint16_t* qpositions, uint32_t* packedIndices,
                             uint64_t* packedWords, crx_cmesh* meta) {
    if (!meta || !packedWords) return fail(CRX_INVALID_ARG, "NULL argument");
    if (numVertices > 0 && (!positions || !qpositions)) return fail(CRX_INVALID_ARG, "positions is NULL");
    if (numTriangles > 0 && (!indices || !packedIndices)) return fail(CRX_INVALID_ARG, "indices is NULL");
    if (numTriangles > 1431655765u) return fail(CRX_INVALID_ARG, "more than 1431655765 triangles");
    const uint64_t numIndices = 3ull * numTriangles;
    // LargeGlbLoader.h:482-504
    uint32_t imin = 0xffffffffu, imax = 0;
    for (uint64_t i = 0; i < numIndices; i++) {
        if (indices[i] >= numVertices) return fail(CRX_INVALID_ARG, "index " + std::to_string(i) + " out of range");
        imin = std::min(imin, indices[i]);
        imax = std::max(imax, indices[i]);
    }
    if (numIndices == 0) imin = imax = 0;
    const uint64_t range = (uint64_t)imax - imin;
    const uint64_t bits = (uint64_t)std::ceil(std::log2f((float)(range + 1)));
    const uint64_t words = ((numIndices * bits + 7) / 8 + 3) / 4;
    std::memset(packedIndices ? packedIndices : nullptr, 0, packedIndices ? (size_t)(numIndices + 1) * 4 : 0);
    for (uint64_t i = 0; i < numIndices; i++) {
        // BitEdit::writeU32 (BitEdit.h:28-49)
        const uint64_t off = i * bits;
        if (bits == 0) break;
        const uint64_t w = off / 32;
        const uint32_t sh = (uint32_t)(off % 32);
        const uint32_t mask = bits == 32 ? 0xffffffffu : ((1u << bits) - 1u);
        const uint32_t value = (indices[i] - imin) & mask;
        packedIndices[w] = (packedIndices[w] & ~(mask << sh)) | (value << sh);
        if (sh + bits > 32) {
            const uint32_t written = 32 - sh;
            packedIndices[w + 1] = (packedIndices[w + 1] & ~(mask >> written)) | (value >> written);
        }
    }
    // LargeGlbLoader.h:640-680: accessor min / max, then 16-bit quantisation
    crx_box3 box{{INFINITY, INFINITY, INFINITY}, {-INFINITY, -INFINITY, -INFINITY}};
    for (uint32_t v = 0; v < numVertices; v++) {
        const float* q = positions + 3ull * v;
        box.min.x = std::fmin(box.min.x, q[0]); box.max.x = std::fmax(box.max.x, q[0]);
        box.min.y = std::fmin(box.min.y, q[1]); box.max.y = std::fmax(box.max.y, q[1]);
        box.min.z = std::fmin(box.min.z, q[2]); box.max.z = std::fmax(box.max.z, q[2]);
    }
    if (numVertices == 0) box = crx_box3{{0, 0, 0}, {0, 0, 0}};
    const float size[3] = {box.max.x - box.min.x, box.max.y - box.min.y, box.max.z - box.min.z};
    const float lo[3] = {box.min.x, box.min.y, box.min.z};
    for (uint32_t v = 0; v < numVertices; v++) {
        for (int a = 0; a < 3; a++) {
            const float d = positions[3ull * v + a] - lo[a];
            double x = size[a] > 0.0f ? 65536.0 * (double)d / (double)size[a] : 0.0;
            x = x < 0.0 ? 0.0 : (x > 65535.0 ? 65535.0 : x);
            qpositions[3ull * v + a] = (uint16_t)x;
        }
    }
    std::memset(meta, 0, sizeof(*meta));
    meta->numTriangles = numTriangles;
    meta->numVertices = numVertices;
    meta->aabb = box;
    // CuRast_render.h:358-378
    meta->compressionFactor = crx_vec3{size[0] / 65536.0f, size[1] / 65536.0f, size[2] / 65536.0f};
    meta->index_min = imin;
    meta->index_max = imax;
    meta->bitsPerIndex = (uint32_t)bits;
    meta->compressed = 1;
    meta->isLoaded = 1;
    *packedWords = words;
    return CRX_OK;
}

crx_status crx_mesh_create_compressed(crx_context* ctx, const float* positions, uint32_t numVertices,
                                      const uint32_t* indices, uint32_t numTriangles, const uint32_t* colors,
                                      crx_mesh** out) {
    if (!ctx || !out) return fail(CRX_INVALID_ARG, "NULL argument");
    *out = nullptr;
    if (numTriangles > 0 && !indices) return fail(CRX_INVALID_ARG, "a compressed mesh needs indices");
    std::vector<uint16_t> q(3ull * numVertices);
    std::vector<uint32_t> packed(3ull * numTriangles + 1);
    uint64_t words = 0;
    crx_cmesh meta;
    CRX_TRY(crx_compress_mesh(positions, numVertices, indices, numTriangles, q.data(), packed.data(), &words, &meta));
    CRX_HIP(hipSetDevice(ctx->device));
    crx_mesh* m = new crx_mesh();
    m->ctx = ctx;
    m->owned = true;
    m->numVertices = numVertices;
    m->numTriangles = numTriangles;
    m->seq = ++ctx->meshSeq;
    m->aabb = meta.aabb;
    m->compressed = true;
    m->bitsPerIndex = meta.bitsPerIndex;
    m->indexMin = meta.index_min;
    m->indexMax = meta.index_max;
    m->compressionFactor = meta.compressionFactor;
    auto upload = [&](void** dst, const void* src, size_t bytes) -> crx_status {
        if (!src || bytes == 0) return CRX_OK;
        hipError_t e = hipMalloc(dst, bytes);
        if (e != hipSuccess) { *dst = nullptr; return fail(CRX_OUT_OF_MEMORY, std::string("hipMalloc: ") + hipGetErrorString(e)); }
        e = hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice);
        if (e != hipSuccess) return fail(CRX_HIP_ERROR, std::string("hipMemcpy: ") + hipGetErrorString(e));
        return CRX_OK;
    };
    crx_status st = upload((void**)&m->positions, q.data(), 6ull * numVertices);
    if (st == CRX_OK) st = upload((void**)&m->indices, packed.data(), 4ull * std::max<uint64_t>(words, 1));
    if (st == CRX_OK && colors) st = upload((void**)&m->colors, colors, 4ull * numVertices);
    if (st != CRX_OK) {
        crx_mesh_destroy(m);
        return st;
    }
    // the row length of the uncompressed index order, kept for the packed stream
    if (m->indices) {
        m->rowTris = detect_row_triangles(indices, numTriangles);
        if (m->rowTris) {
            std::lock_guard<std::mutex> lk(g_rowHintsMu);
            g_rowHints[m->indices] = m->rowTris;
        }
    }
    *out = m;
    return CRX_OK;
}

crx_status crx_mesh_create_device(crx_context* ctx, const float* d_positions, uint32_t numVertices,
                                  const uint32_t* d_indices, uint32_t numTriangles,
                                  const uint32_t* d_colors, const crx_box3* aabb, crx_mesh** out) {
    if (!ctx || !out || !aabb) return fail(CRX_INVALID_ARG, "NULL argument");
    *out = nullptr;
    if (numTriangles > 1431655765u) return fail(CRX_INVALID_ARG, "more than 1431655765 triangles");
    if (numTriangles > 0 && !d_positions) return fail(CRX_INVALID_ARG, "positions is NULL");
    crx_mesh* m = new crx_mesh();
    m->ctx = ctx;
    m->owned = false;
    m->positions = const_cast<float*>(d_positions);
    m->indices = const_cast<uint32_t*>(d_indices);
    m->colors = const_cast<uint32_t*>(d_colors);
    m->numVertices = numVertices;
    m->numTriangles = numTriangles;
    m->aabb = *aabb;
    m->seq = ++ctx->meshSeq;
    *out = m;
    return CRX_OK;
}

crx_status crx_mesh_destroy(crx_mesh* mesh) {
    if (!mesh) return CRX_OK;
    if (mesh->rowTris && mesh->indices) {
        std::lock_guard<std::mutex> lk(g_rowHintsMu);
        g_rowHints.erase(mesh->indices);
    }
    if (mesh->owned) {
        if (mesh->positions) (void)hipFree(mesh->positions);
        if (mesh->indices) (void)hipFree(mesh->indices);
        if (mesh->colors) (void)hipFree(mesh->colors);
    }
    if (mesh->uvs) (void)hipFree(mesh->uvs);
    if (mesh->normals) (void)hipFree(mesh->normals);
    delete mesh;
    return CRX_OK;
}

crx_status crx_mesh_set_attributes(crx_mesh* mesh, const float* uvs, const float* normals) {
    if (!mesh) return fail(CRX_INVALID_ARG, "mesh is NULL");
    CRX_HIP(hipSetDevice(mesh->ctx->device));
    // upload into a new buffer first: the mesh keeps its old one on failure
    auto replace = [&](float** dst, const float* src, size_t bytes) -> crx_status {
        if (!src || bytes == 0) return CRX_OK;
        float* d = nullptr;
        hipError_t e = hipMalloc((void**)&d, bytes);
        if (e != hipSuccess) return fail(CRX_OUT_OF_MEMORY, std::string("hipMalloc: ") + hipGetErrorString(e));
        e = hipMemcpy(d, src, bytes, hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            (void)hipFree(d);
            return fail(CRX_HIP_ERROR, std::string("hipMemcpy: ") + hipGetErrorString(e));
        }
        // frames already enqueued may still read the old buffer
        if (*dst) {
            CRX_HIP(hipDeviceSynchronize());
            (void)hipFree(*dst);
        }
        *dst = d;
        return CRX_OK;
    };
    CRX_TRY(replace(&mesh->uvs, uvs, 8ull * mesh->numVertices));
    CRX_TRY(replace(&mesh->normals, normals, 12ull * mesh->numVertices));
    mesh->ctx->attrEpoch++;
    return CRX_OK;
}

crx_status crx_mesh_row_triangles(const crx_mesh* mesh, uint32_t* out) {
    if (!mesh || !out) return fail(CRX_INVALID_ARG, "NULL argument");
    *out = mesh->rowTris;
    return CRX_OK;
}

crx_status crx_mesh_info(const crx_mesh* mesh, uint32_t* numTriangles, uint32_t* numVertices, crx_box3* aabb) {
    if (!mesh) return fail(CRX_INVALID_ARG, "mesh is NULL");
    if (numTriangles) *numTriangles = mesh->numTriangles;
    if (numVertices) *numVertices = mesh->numVertices;
    if (aabb) *aabb = mesh->aabb;
    return CRX_OK;
}

crx_status crx_mesh_buffers(const crx_mesh* mesh, const float** d_positions, const uint32_t** d_indices,
                            const uint32_t** d_colors) {
    if (!mesh) return fail(CRX_INVALID_ARG, "mesh is NULL");
    if (d_positions) *d_positions = mesh->positions;
    if (d_indices) *d_indices = mesh->indices;
    if (d_colors) *d_colors = mesh->colors;
    return CRX_OK;
}

crx_status crx_scene_create(crx_context* ctx, crx_scene** out) {
    if (!ctx || !out) return fail(CRX_INVALID_ARG, "NULL argument");
    crx_scene* s = new crx_scene();
    s->ctx = ctx;
    *out = s;
    return CRX_OK;
}

crx_status crx_scene_destroy(crx_scene* scene) {
    delete scene;
    return CRX_OK;
}

crx_status crx_scene_add_node(crx_scene* scene, crx_mesh* mesh, const crx_mat4* world, uint32_t* node_id) {
    if (!scene || !mesh || !world) return fail(CRX_INVALID_ARG, "NULL argument");
    if (mesh->ctx != scene->ctx) return fail(CRX_INVALID_ARG, "mesh belongs to another context");
    scene->nodes.push_back(crx_scene::Node{mesh, *world, world_box(mesh->aabb, *world), crx_texture{}});
    scene->version++;
    if (node_id) *node_id = (uint32_t)(scene->nodes.size() - 1);
    return CRX_OK;
}

crx_status crx_scene_set_transform(crx_scene* scene, uint32_t node_id, const crx_mat4* world) {
    if (!scene || !world) return fail(CRX_INVALID_ARG, "NULL argument");
    if (node_id >= scene->nodes.size()) return fail(CRX_INVALID_ARG, "node_id out of range");
    scene->nodes[node_id].world = *world;
    scene->nodes[node_id].wbox = world_box(scene->nodes[node_id].mesh->aabb, *world);
    scene->version++;
    return CRX_OK;
}

crx_status crx_scene_set_texture(crx_scene* scene, uint32_t node_id, const crx_texture* texture) {
    if (!scene) return fail(CRX_INVALID_ARG, "scene is NULL");
    if (node_id >= scene->nodes.size()) return fail(CRX_INVALID_ARG, "node_id out of range");
    crx_texture t{};
    if (texture) t = *texture;
    if (t.data && !t.huffmanTables && (t.width <= 0 || t.height <= 0))
        return fail(CRX_INVALID_ARG, "texture width and height must be positive");
    scene->nodes[node_id].texture = t;
    scene->version++;
    return CRX_OK;
}

crx_status crx_scene_num_nodes(const crx_scene* scene, uint32_t* out) {
    if (!scene || !out) return fail(CRX_INVALID_ARG, "NULL argument");
    *out = (uint32_t)scene->nodes.size();
    return CRX_OK;
}

void crx_render_settings_default(crx_render_settings* out) {
    if (!out) return;
    out->instanced = 1;                                  // RASTERIZER_VISBUFFER_INSTANCED
    out->frustum_culling = 1;                            // enableFrustumCulling
    out->resolve = 1;
    out->attribute = CRX_DISPLAY_VERTEX_COLORS;
}

crx_status crx_render(crx_context* ctx, crx_scene* scene, const crx_view* view,
                      const crx_render_settings* settings, uint64_t* framebuffer, uint64_t* colorbuffer,
                      void* stream, crx_draw_stats* stats) {
    if (!view) return fail(CRX_INVALID_ARG, "view is NULL");
    return render_views(ctx, scene, view, 1, settings, &framebuffer, &colorbuffer, as_stream(stream), stats);
}

crx_status crx_render_views(crx_context* ctx, crx_scene* scene, const crx_view* views, uint32_t numViews,
                            const crx_render_settings* settings, uint64_t* const* framebuffers,
                            uint64_t* const* colorbuffers, void* stream, crx_draw_stats* stats) {
    return render_views(ctx, scene, views, numViews, settings, framebuffers, colorbuffers, as_stream(stream), stats);
}

crx_status crx_autotune(crx_context* ctx, crx_scene* scene, const crx_view* views, uint32_t numViews,
                        uint64_t* const* framebuffers, uint32_t flags, uint32_t reps, crx_autotune_result* out) {
    if (!ctx || !scene || !views || !framebuffers || !out) return fail(CRX_INVALID_ARG, "NULL argument");
    if (reps == 0 || reps > 1000) return fail(CRX_INVALID_ARG, "reps must be in [1, 1000]");
    if ((flags & ~(CRX_TUNE_SCHEDULE | CRX_TUNE_ROUTING | CRX_TUNE_QUEUES)) != 0)
        return fail(CRX_INVALID_ARG, "unknown autotune flags");
    std::memset(out, 0, sizeof(*out));
    crx_render_settings settings;
    crx_render_settings_default(&settings);
    settings.resolve = 0;
    const crx_tuning start = ctx->tuning;
    const bool wasTiming = ctx->timing;
    ctx->timing = true;
    uint32_t trials = 0;
    crx_draw_stats lastStats{};

    // median device time of stages 1-3 over `reps` draws with tuning t
    auto score = [&](const crx_tuning& t, float* ms) -> crx_status {
        CRX_TRY(check_tuning(t));
        ctx->tuning = t;
        CRX_TRY(render_views(ctx, scene, views, numViews, &settings, framebuffers, nullptr, nullptr, &lastStats));
        std::vector<float> v;
        for (uint32_t r = 0; r < reps; r++) {
            CRX_TRY(render_views(ctx, scene, views, numViews, &settings, framebuffers, nullptr, nullptr, &lastStats));
            crx_timings tm;
            CRX_TRY(crx_context_timings(ctx, &tm));
            v.push_back(tm.pipeline_ms);
        }
        std::sort(v.begin(), v.end());
        *ms = v[v.size() / 2];
        trials++;
        return CRX_OK;
    };

    // knob table: field and candidate values
    struct Knob {
        uint32_t crx_tuning::*field;
        std::vector<uint32_t> values;
        uint32_t cls;
    };
    const std::vector<Knob> knobs = {
        {&crx_tuning::stage1_block, {128, 256, 512}, CRX_TUNE_SCHEDULE},
        {&crx_tuning::stage1_tris_per_lane, {1, 2, 4}, CRX_TUNE_SCHEDULE},
        {&crx_tuning::stage1_waves_per_cu, {8, 16, 24, 32, 48, 64, 0}, CRX_TUNE_SCHEDULE},
        {&crx_tuning::stage1_order, {0, 1, 2}, CRX_TUNE_SCHEDULE},
        {&crx_tuning::stage1_cull_order, {0, 1, 2, 3}, CRX_TUNE_SCHEDULE},
        {&crx_tuning::stage1_fast_tiny, {0, 1}, CRX_TUNE_SCHEDULE},
        {&crx_tuning::stage1_layout, {0, 1, 2}, CRX_TUNE_SCHEDULE},
        {&crx_tuning::stage1_strip_rows, {2, 4, 8, 16, 32}, CRX_TUNE_SCHEDULE},
        {&crx_tuning::stage1_strip_batches, {1, 2, 4}, CRX_TUNE_SCHEDULE},
        {&crx_tuning::stage2_waves_per_cu, {4, 8, 16, 32}, CRX_TUNE_SCHEDULE},
        {&crx_tuning::stage3_waves_per_cu, {4, 8, 16, 32}, CRX_TUNE_SCHEDULE},
        {&crx_tuning::small_threshold, {32, 64, 128, 256, 512}, CRX_TUNE_ROUTING},
        {&crx_tuning::tile_size, {32, 64, 128}, CRX_TUNE_ROUTING},
    };

    crx_status st = CRX_OK;
    crx_tuning best = start;
    float bestMs = 0.0f;
    do {
        st = score(best, &bestMs);
        if (st != CRX_OK) break;
        out->start_ms = bestMs;
        for (int pass = 0; pass < 2 && st == CRX_OK; pass++) {
            bool improved = false;
            for (const Knob& k : knobs) {
                if (!(flags & k.cls)) continue;
                for (uint32_t val : k.values) {
                    if (best.*(k.field) == val) continue;
                    crx_tuning cand = best;
                    cand.*(k.field) = val;
                    float ms = 0.0f;
                    st = score(cand, &ms);
                    if (st != CRX_OK) break;
                    if (ms < bestMs * 0.98f) {   // keep a change only for a clear gain
                        best = cand;
                        bestMs = ms;
                        improved = true;
                    }
                }
                if (st != CRX_OK) break;
            }
            if (!improved) break;
        }
        if (st != CRX_OK) break;
        if (flags & CRX_TUNE_QUEUES) {
            // one more draw with the best tuning gives the queue counts; keep 2x headroom
            float ms = 0.0f;
            st = score(best, &ms);
            if (st != CRX_OK) break;
            best.max_nontrivial = std::max<uint32_t>(4096u, (uint32_t)std::min<uint64_t>(
                0xffffffffull, 2ull * lastStats.num_nontrivial));
            best.max_huge = std::max<uint32_t>(4096u, (uint32_t)std::min<uint64_t>(
                0xffffffffull, 2ull * lastStats.num_huge));
        }
    } while (false);
    ctx->timing = wasTiming;
    for (bool& r : ctx->recorded) r = false;
    if (st != CRX_OK) {
        ctx->tuning = start;
        return st;
    }
    ctx->tuning = best;
    out->tuning = best;
    out->best_ms = bestMs;
    out->trials = trials;
    return CRX_OK;
}

}  // extern "C"
