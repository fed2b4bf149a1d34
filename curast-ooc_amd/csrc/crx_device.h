// crx_device.h — device-side records and math of the gfx950 rasterizer.
//
// Shared by the kernels (crx_kernels.hip) and the host glue (crx_api.cpp):
// the host builds these compact records from the CMesh arrays of the C-ABI,
// uploads them once per draw and the kernels read nothing else.
#pragma once

#include <stdint.h>

#ifdef __HIPCC__
#define CRX_HD __host__ __device__ __forceinline__
#else
#define CRX_HD inline
#endif

namespace crx {

// Stage-1 work unit.  Instanced draws: one unit per unique mesh that loops
// over instCount consecutive instances.  Non-instanced draws: one unit per
// instance (instCount == 1).
struct DrawItem {
    const uint32_t* indices;    // IX_PU16: bit-packed (BitEdit layout); non-indexed: the positions
    uint32_t indexMask;         // ~0 = indexed, 0 = non-indexed (triangle t = vertices 3t..3t+2)
    const float* positions;     // xyz, 12-byte stride; IX_PU16: 3 x uint16, 6-byte stride
    uint64_t cumm;              // cummulativeTriangleCount of the first instance
    uint32_t numTriangles;
    uint32_t instOffset;        // first instance (index into transforms / instances)
    uint32_t instCount;
    uint32_t firstBatch;        // exclusive prefix sum of batches over items
    // IX_PU16 (CMesh::bitsPerIndex, index_min, compressionFactor, aabb.min)
    uint32_t bitsPerIndex;
    uint32_t indexMin;
    float cf[3];
    float amin[3];
    uint32_t numVertices;       // CMesh::numVertices (IX_PU16: bounds the 8-byte vertex reads)
    // Pair layout (k_stage1_pairs, whose batch is one wave's 128
    // triangles): the triangles as rows of rowTris triangles cut into
    // `strips` strips of unitCols batches of 128; a work unit is strip
    // s = unit / unitSegs over rows [g * unitRows, (g + 1) * unitRows),
    // g = unit % unitSegs, drawn by one wave row after row, each row's
    // batches left to right.  A contiguous walk is rowTris = 128, strips = 1,
    // unitCols = 1.  firstBatch then counts units.
    uint32_t rowTris;
    uint32_t strips;
    uint32_t unitRows;
    uint32_t unitSegs;
    uint32_t unitCols;
};

// Per-instance record for stage 2, stage 3 and resolve.
struct InstanceRec {
    const uint32_t* indices;
    const float* positions;
    const uint32_t* colors;
    const float* normals;
    const float* uvs;
    uint64_t cumm;
    uint32_t numTriangles;
    uint32_t transformIndex;    // CMesh::instances.offset
    int32_t id;
    uint32_t isLoaded;
    float world[16];
    // IX_PU16 (CMesh::compressed, bitsPerIndex, index_min, compressionFactor, aabb)
    uint32_t compressed;
    uint32_t bitsPerIndex;
    uint32_t indexMin;
    float cf[3];
    float amin[3];
    float amax[3];
    // TEXTURE (CMesh::texture): RGBA8 mip chain, level 0 first, each level
    // ((w + 1) / 2) x ((h + 1) / 2) of the one before; texJpeg = 1 for a
    // JPEG-coded texture (CMesh::texture.huffmanTables set)
    const uint32_t* texData;
    int32_t texW, texH;
    uint32_t texJpeg;
};

// world-to-view transform of one instance plus its mirror flag
// (det(worldView[0..2]) < 0, computed on the host in the same float order).
// skip != 0 drops the instance from this view (multi-view draws only: the
// host-side node cull of that view).
struct Transform {
    float m[16];
    uint32_t flip;
    uint32_t skip;
    uint32_t pad[2];
};

// Counter block, one per context, cleared by the clear kernel of each draw.
enum Counter : int {
    CTR_NONTRIVIAL = 0,   // stage-2 queue entries produced by stage 1
    CTR_HUGE = 1,         // stage-3 tile entries produced by stage 2
    CTR_CURSOR2 = 2,      // stage-2 work cursor
    CTR_CURSOR3 = 3,      // stage-3 work cursor
    // k_stage1_pairs: eight unit counters, one per XCD, each on its own
    // 128-byte line (word CTR_CLAIM + kClaimStride * xcd)
    CTR_CLAIM = 32,
    CTR_COUNT = 32 + 8 * 32
};
constexpr uint32_t kClaimStride = 32;

// One render target of a draw.  A plain draw has one; a multi-view draw has
// one per view, all sharing the draw items and instances.  Stage-2 and
// stage-3 queue entries carry the view index in bits 24..31 of their
// instance field.
struct ViewParams {
    uint64_t* fb;
    const Transform* transforms;   // this view's world-to-view transform per instance
    int32_t W, H;
    float Wf, Hf;            // float(W), float(H)
    float Wm1, Hm1;          // float(W - 1), float(H - 1)
    float f, aspect;         // proj[1][1], float(W)/float(H)
    float fa;                // f / aspect
    float faI, fI;           // 1/(f/aspect), 1/f (stage 3)
    float tinyMargin;        // bound on |approximate - IEEE| screen coordinates, with slack
    uint32_t pad[2];
};

// Instances per stage-1 item from which the single-view kernel reduces wave
// boxes and tests them per instance (the reduction is paid once per batch).
constexpr uint32_t kBoxInstances = 4;

constexpr uint32_t kViewShift = 24;                 // instance field: view << 24 | instance
constexpr uint32_t kInstanceMask = (1u << 24) - 1u;
constexpr uint32_t kMaxViews = 256;

struct RasterParams {
    const DrawItem* items;
    const InstanceRec* inst;
    const ViewParams* views;
    uint32_t* ctrNontrivial;      // stage-2 entries produced by stage 1
    uint32_t* ctrHuge;            // stage-3 entries produced by stage 2
    uint32_t* cursor2;            // stage-2 work cursor
    uint32_t* cursor3;            // stage-3 work cursor
    const uint32_t* batchCount;   // non-null: numBatches is read from here (plans built on the device)
    uint64_t* nontrivial;
    int4* huge;
    uint32_t numItems;
    uint32_t numBatches;
    uint32_t numViews;
    uint32_t ntCap;
    uint32_t hugeCap;
    uint32_t smallThreshold;
    int32_t tile;
    uint32_t batchShift;     // log2(triangles per stage-1 batch = workgroup size)
    uint32_t order;          // stage-1 batch order (crx_tuning::stage1_order)
    uint32_t cullOrder;      // 0 = backface first, 1 = frustum (y then x) first, 2 = frustum (x then y) first
    uint32_t fastTiny;       // filtered tiny-triangle cull (crx_tuning::stage1_fast_tiny)
    uint32_t numWaves2;      // waves in the stage-2 grid
    uint32_t numWaves3;      // waves in the stage-3 grid
    uint32_t pad0;
    uint32_t* claims;        // k_stage1_pairs: the per-XCD unit counters (CTR_CLAIM of the counter block)
};

struct ResolveParams {
    const InstanceRec* inst;
    const uint64_t* prefix;
    const uint64_t* fb;
    uint64_t* cb;
    uint32_t numInstances;
    int32_t W, H;
    int32_t attribute;
    float Wf, Hf;
    float p00I, p11I;          // 1/proj[0][0], 1/proj[1][1]
    float viewI[16];           // inverse(view), GLM order, computed once per frame
    float origin[3];           // viewI * (0,0,0,1)
    float view[16];            // the view matrix (STAGE: view-space vertices)
    float f, aspect;           // proj[1][1], float(W) / float(H) (STAGE)
    uint32_t firstBlock;       // multi-view resolve: first 64x4 tile block of this view
    uint32_t ownedOnly;        // 1: leave pixels whose triangle is not in the listed instances
                               // (a resolve in passes over groups of instances)
};

// The five stage counters of the kernel-level entry points (RasterArgs:
// numProcessedBatches, numProcessedBatches_nontrivial, hugeTrianglesCounter,
// nontrivialTrianglesCounter, numProcessedHugeTriangles).
struct CounterSet {
    uint32_t* p[5];
};

// One framebuffer of a multi-view clear.
struct ClearView {
    uint64_t* fb;
    uint32_t numPixels;
    uint32_t firstBlock;       // first block of this view (kClearChunk pixels per block)
};

constexpr uint32_t kClearChunk = 4096;

constexpr uint64_t kClearPixel = 0xFFFFFFF000000000ull;

// ---------------------------------------------------------------------------
// Math in GLM 0.9.9's evaluation order.  Compiled with -ffp-contract=off, so
// each * and + below is one rounded IEEE binary32 operation.
// ---------------------------------------------------------------------------

struct f2 { float x, y; };
struct f3 { float x, y, z; };

CRX_HD f3 sub3(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
CRX_HD f3 add3(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
CRX_HD f3 scale3(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
CRX_HD float dot3(f3 a, f3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
CRX_HD f3 cross3(f3 x, f3 y) {
    return f3{x.y * y.z - y.y * x.z, x.z * y.x - y.z * x.x, x.x * y.y - y.x * x.y};
}
CRX_HD f3 normalize3(f3 v) {
    float s = 1.0f / sqrtf(dot3(v, v));
    return scale3(v, s);
}

// xyz of M * (x, y, z, w): (c0*x + c1*y) + (c2*z + c3*w)
CRX_HD f3 xform(const float* m, f3 p, float w) {
    return f3{(m[0] * p.x + m[4] * p.y) + (m[8] * p.z + m[12] * w),
              (m[1] * p.x + m[5] * p.y) + (m[9] * p.z + m[13] * w),
              (m[2] * p.x + m[6] * p.y) + (m[10] * p.z + m[14] * w)};
}

// xyz of M * (x, y, z, 1) for a point: c3*1 is exact, so it is c3 itself.
CRX_HD f3 xform_point(const float* m, f3 p) {
    return f3{(m[0] * p.x + m[4] * p.y) + (m[8] * p.z + m[12]),
              (m[1] * p.x + m[5] * p.y) + (m[9] * p.z + m[13]),
              (m[2] * p.x + m[6] * p.y) + (m[10] * p.z + m[14])};
}

CRX_HD uint64_t pack_pixel(float depth, uint64_t id) {
    uint32_t bits = __builtin_bit_cast(uint32_t, depth) & 0x7FFFFFF8u;
    if (__builtin_isinf(depth)) bits = 0x7FFFFFF8u;
    return ((uint64_t)bits << 33) | id;
}

CRX_HD float unpack_depth(uint64_t packed) {
    uint32_t bits = (uint32_t)(packed >> 33) & 0x7FFFFFF8u;
    return bits == 0x7FFFFFF8u ? __builtin_inff() : __builtin_bit_cast(float, bits);
}

// float -> int32 with the reference device semantics: truncation, NaN -> 0,
// saturation (a plain C++ cast of an out-of-range value is undefined).
CRX_HD int32_t f2i(float x) {
    if (!(x == x)) return 0;
    if (x >= 2147483648.0f) return 2147483647;
    if (x <= -2147483648.0f) return (-2147483647 - 1);
    return (int32_t)x;
}

// float in [0,255] (or NaN) -> uint8 (NaN -> 0)
CRX_HD uint32_t f2u8(float x) {
    if (!(x > 0.0f)) return 0u;
    if (x >= 255.0f) return 255u;
    return (uint32_t)x;
}

CRX_HD float clampf(float v, float lo, float hi) {
    if (v < lo) return lo;
    if (v > hi) return hi;
    return v;
}

}  // namespace crx
