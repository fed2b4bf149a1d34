// crx_kernels.hip — gfx950 kernels of the visibility-buffer rasterizer.
//
//   k_clear    framebuffer clear + counter reset   (kernel_clearFramebuffer, resolve.cu:213-232)
//   k_stage1   one lane per small triangle         (stage1_drawSmallTriangles, triangles_visbuffer.cu:274-384)
//   k_stage2   one wave64 per medium triangle      (stage2_drawMediumTriangles, :388-563)
//   k_stage3   one wave64 per 64x64 tile           (stage3_drawHugeTriangles, :568-745)
//   k_resolve  one lane per pixel                  (kernel_resolve_visbuffer_to_colorbuffer2D, resolve.cu:640-1141)
//
// Every kernel reproduces the reference arithmetic operation by operation
// (IEEE binary32, GLM evaluation order, no contraction); what changes is the
// schedule.  Cull predicates are independent early-outs, so they are evaluated
// cheapest first; the frustum test is done without division by an exact
// rewrite of "RN(p/d) > 1" (see frustum_culled_axis).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "crx.h"
#include "crx_device.h"
#include "crx_internal.h"

namespace crx {

namespace {

constexpr int kWave = 64;

// Loads through the constant address space are scalar (s_load) when the index
// is uniform; the draw records are read-only for the whole kernel.  Loads
// through the global address space avoid flat instructions.  (The host pass
// of the compiler sees plain pointers.)
#if defined(__HIP_DEVICE_COMPILE__)
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(4))) T* kconst(const T* p) {
    return (const __attribute__((address_space(4))) T*)(p);
}
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(1))) T* kglobal(const T* p) {
    return (const __attribute__((address_space(1))) T*)(p);
}
#else
template <typename T>
__device__ __forceinline__ const T* kconst(const T* p) { return p; }
template <typename T>
__device__ __forceinline__ const T* kglobal(const T* p) { return p; }
#endif

// Short reciprocal and quotient sequences with the IEEE (round-to-nearest-
// even) results on the operand ranges where stage 1 uses them.
//
// rcp_refined(b): v_rcp_f32 (error below 1 ulp) refined by one Newton step
// with fused multiply-adds, e = 1 - b*r0 (exact), r1 = r0 + e*r0.  For every
// float b with kRcpLo <= |b| < kRcpHi it equals RN(1/b): crx_check_division
// compares it with the IEEE division for all 2 * 250 * 2^23 such floats.
// div_fast(a, b, y) with y = rcp_refined(b) = RN(1/b): q0 = RN(a*y) is within
// one ulp of a/b, rem = a - b*q0 is exact (one FMA) and q1 = RN(q0 + rem*y)
// is RN(a/b) (Markstein's theorem: P. Markstein, IBM J. Res. Dev. 34(1),
// 1990; Muller et al., Handbook of Floating-Point Arithmetic, on FMA-based
// division) when kDivLo <= |a|, |b| <= kDivHi, where nothing underflows or
// overflows; crx_check_division also compares it with the IEEE division on
// random operand pairs of that range.  For |a| below kDivLo (zero and
// subnormals included) both quotients are below 2^-26 in magnitude, so
// q * 0.5 + 0.5 -- the only use in stage 1 -- is 0.5 for both (also
// checked).
constexpr float kRcpLo = 0x1p-125f, kRcpHi = 0x1p125f;
constexpr float kDivLo = 0x1p-60f, kDivHi = 0x1p64f;

__device__ __forceinline__ float rcp_refined(float b) {
    const float r0 = __builtin_amdgcn_rcpf(b);
    const float e = __builtin_fmaf(-b, r0, 1.0f);
    return __builtin_fmaf(e, r0, r0);
}

__device__ __forceinline__ float div_fast(float a, float b, float y) {
    const float q0 = a * y;
    const float rem = __builtin_fmaf(-b, q0, a);
    return __builtin_fmaf(rem, y, q0);
}

// RN(1/b) for any b: the short sequence, or the IEEE division behind a
// wave-uniform branch for lanes out of range (so it is not evaluated
// speculatively for every lane).
__device__ __forceinline__ float rcp_rn(float b) {
    float r = rcp_refined(b);
    const float m = fabsf(b);
    const bool ok = m >= kRcpLo && m < kRcpHi;
    if (__ballot(!ok) != 0ull) {
        if (!ok) r = 1.0f / b;
    }
    return r;
}

// Orders LDS traffic between the lanes of one wave (no hardware barrier: the
// waves of a workgroup run independent loops here).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ void atomic_min_fb(uint64_t* fb, uint32_t pixel, uint64_t value) {
    atomicMin((unsigned long long*)(fb + pixel), (unsigned long long)value);
}

__device__ __forceinline__ f3 load_pos(const float* __restrict__ positions, uint32_t i) {
    const auto p = kglobal(positions) + 3ull * i;
    return f3{p[0], p[1], p[2]};
}

// Frustum test of one axis for a triangle whose depths d are all >= 2^-102
// (or +inf), without dividing.  With p = RN(fa*x) the reference culls when
// RN(p/d) > 1 holds for all three vertices, or RN(p/d) < -1 for all three
// (triangles_visbuffer.cu:113-118).  Round-to-nearest-even gives
// RN(q) > 1 <=> q > 1 + 2^-24, i.e. p - d > d*2^-24.  d*2^-24 is exact (no
// underflow for d >= 2^-102) and so is p - d wherever it matters: Sterbenz
// makes it exact for p in [d/2, 2d], and outside that range the rounded
// difference cannot cross the threshold.  The same argument gives
// RN(p/d) < -1 <=> p + d < -d*2^-24.  Infinite and NaN p compare exactly like
// the quotient they stand for.  The comparisons combine as lane masks.
__device__ __forceinline__ bool frustum_culled_axis(float pa, float pb, float pc, float da, float db, float dc) {
    const float ta = da * 0x1p-24f, tb = db * 0x1p-24f, tc = dc * 0x1p-24f;
    const bool above = (pa - da > ta) & (pb - db > tb) & (pc - dc > tc);
    const bool below = (pa + da < -ta) & (pb + db < -tb) & (pc + dc < -tc);
    return above | below;
}

// One view-space row of M * (x, y, z, 1): (r0*x + r1*y) + (r2*z + r3).
CRX_HD float xform_row(const float* m, int r, f3 p) {
    return (m[r] * p.x + m[4 + r] * p.y) + (m[8 + r] * p.z + m[12 + r]);
}

// Screen x (or y) of a vertex, -0.5 shifted (ndcToScreen, :103-108, :132-134),
// from an NDC value.
__device__ __forceinline__ float to_screen(float ndc, float size) {
    return (ndc * 0.5f + 0.5f) * size - 0.5f;
}

// Exact tiny-triangle cull of one axis decided from approximate coordinates.
// The reference culls the axis when its clipped bounding box [lo, hi] holds no
// sample: lo > floor(lo) and hi < floor(lo) + 1.  lo' and hi' are within `m`
// of lo and hi; if [lo'-m, lo'+m] lies strictly inside (k-1, k) with
// k = ceil(lo'-m), then lo is not an integer and ceil(lo) = k, and
// hi' + m < k gives hi < k: the reference culls the axis.
__device__ __forceinline__ bool surely_culled_axis(float lo, float hi, float m) {
    const float k = ceilf(lo - m);
    return (lo + m < k) && (hi + m < k);
}

// Cull order of an item: tuning value 3 picks backface first for meshes
// drawn over many instances (whose instances mostly sit inside the view, so
// the frustum rarely culls) and frustum-y first for the rest.
__device__ __forceinline__ uint32_t item_cull_order(const RasterParams& P, uint32_t instCount) {
    return P.cullOrder == 3u ? (instCount >= kBoxInstances ? 0u : 1u) : P.cullOrder;
}

// rasterize<INDEXBUFFER, UNCOMPRESSED>, triangles_visbuffer.cu:83-271, for one
// triangle-instance.  `id` is the packed triangle ID the reference forms at
// :253 and `queueEntry` the stage-2 entry it forms at :193-195.
//
// The reference's early-outs (all behind, frustum, tiny, backface) are pure
// predicates, so any order gives the same result; coordinates are computed
// lazily in the order the cheapest test needs them.  Two of the tests are
// decided without IEEE divisions when that is provably exact:
//  * frustum: frustum_culled_axis rewrites RN(p/d) > 1 as p - d > d*2^-24;
//  * tiny cull (P.fastTiny): screen coordinates from v_rcp_f32 (1 ulp) are
//    within P.tinyMargin of the IEEE ones for |ndc| <= 2 (the margin covers
//    the quotient error 2^-21*|ndc| scaled by W/2 plus the roundings of the
//    screen transform, with slack); only triangles that are culled for sure
//    leave early, all others take the exact path.
//
// stage1_core takes each vertex's depth d = -view.z and, with PRE, its view
// y row (VQ), from a caller that computed them once for vertices shared
// between triangles: the same operations, so the same bits.
struct VQ {
    float d, y;
};

__device__ __forceinline__ VQ vertex_q(const float* m, f3 p) {
    return VQ{-xform_row(m, 2, p), xform_row(m, 1, p)};
}

template <bool PRE>
__device__ __forceinline__ void stage1_core(const RasterParams& P, const ViewParams& V, const float* __restrict__ m,
                                            bool flip, f3 A, f3 B, f3 C, const VQ& qa, const VQ& qb, const VQ& qc,
                                            uint64_t id, uint64_t queueEntry, uint32_t cullOrder) {
    const float da = qa.d, db = qb.d, dc = qc.d;
    // all vertices behind the camera (:119)
    if (da <= 0.0f && db <= 0.0f && dc <= 0.0f) return;
    bool nontrivial = (da <= 0.0f || db <= 0.0f || dc <= 0.0f);
    const bool safeDepth = (da >= 0x1p-102f && db >= 0x1p-102f && dc >= 0x1p-102f);
    const float fa = V.fa, f = V.f;

    // frustum culling (:113-118), only when every depth is positive; the axis
    // the view culls most is tested before the other coordinate is computed
    const bool frustumFirst = cullOrder != 0;
    float vya = qa.y, vyb = qb.y, vyc = qc.y, vxa = 0.0f, vxb = 0.0f, vxc = 0.0f;
    bool haveX = false, haveY = PRE;
    if (frustumFirst && !nontrivial && safeDepth) {
        if (cullOrder == 1) {
            if (!PRE) { vya = xform_row(m, 1, A); vyb = xform_row(m, 1, B); vyc = xform_row(m, 1, C); }
            haveY = true;
            if (frustum_culled_axis(f * vya, f * vyb, f * vyc, da, db, dc)) return;
            vxa = xform_row(m, 0, A); vxb = xform_row(m, 0, B); vxc = xform_row(m, 0, C);
            haveX = true;
            if (frustum_culled_axis(fa * vxa, fa * vxb, fa * vxc, da, db, dc)) return;
        } else {
            vxa = xform_row(m, 0, A); vxb = xform_row(m, 0, B); vxc = xform_row(m, 0, C);
            haveX = true;
            if (frustum_culled_axis(fa * vxa, fa * vxb, fa * vxc, da, db, dc)) return;
            if (!PRE) { vya = xform_row(m, 1, A); vyb = xform_row(m, 1, B); vyc = xform_row(m, 1, C); }
            haveY = true;
            if (frustum_culled_axis(f * vya, f * vyb, f * vyc, da, db, dc)) return;
        }
    }
    if (!haveX) { vxa = xform_row(m, 0, A); vxb = xform_row(m, 0, B); vxc = xform_row(m, 0, C); }
    if (!haveY) { vya = xform_row(m, 1, A); vyb = xform_row(m, 1, B); vyc = xform_row(m, 1, C); }
    const f3 av{vxa, vya, -da}, bv{vxb, vyb, -db}, cv{vxc, vyc, -dc};
    const float pxa = fa * vxa, pxb = fa * vxb, pxc = fa * vxc;
    const float pya = f * vya, pyb = f * vyb, pyc = f * vyc;

    // backface culling in view space (:168-188)
    {
        const f3 ab = sub3(bv, av);
        const f3 ac = sub3(cv, av);
        float res = dot3(av, cross3(ab, ac));
        if (flip) res = -res;
        if (res >= 0.0f) return;
    }
    if (!frustumFirst && !nontrivial) {
        if (safeDepth) {
            if (frustum_culled_axis(pxa, pxb, pxc, da, db, dc) || frustum_culled_axis(pya, pyb, pyc, da, db, dc))
                return;
        }
    }
    if (!nontrivial && !safeDepth) {
        const float ax = pxa / da, bx = pxb / db, cx = pxc / dc;
        const float ay = pya / da, by = pyb / db, cy = pyc / dc;
        if ((ax > 1.0f && bx > 1.0f && cx > 1.0f) || (ax < -1.0f && bx < -1.0f && cx < -1.0f) ||
            (ay > 1.0f && by > 1.0f && cy > 1.0f) || (ay < -1.0f && by < -1.0f && cy < -1.0f))
            return;
    }

    // filtered tiny-triangle cull (see above); depths in [2^-100, 2^100]
    if (P.fastTiny && !nontrivial && da >= 0x1p-100f && db >= 0x1p-100f && dc >= 0x1p-100f &&
        da <= 0x1p100f && db <= 0x1p100f && dc <= 0x1p100f) {
        const float ra = __builtin_amdgcn_rcpf(da), rb = __builtin_amdgcn_rcpf(db), rc = __builtin_amdgcn_rcpf(dc);
        const float qxa = pxa * ra, qxb = pxb * rb, qxc = pxc * rc;
        const float qya = pya * ra, qyb = pyb * rb, qyc = pyc * rc;
        const float qmax = fmaxf(fmaxf(fmaxf(fabsf(qxa), fabsf(qxb)), fmaxf(fabsf(qxc), fabsf(qya))),
                                 fmaxf(fabsf(qyb), fabsf(qyc)));
        if (qmax <= 2.0f) {
            const float sxa = to_screen(qxa, V.Wf), sxb = to_screen(qxb, V.Wf), sxc = to_screen(qxc, V.Wf);
            const float sya = to_screen(qya, V.Hf), syb = to_screen(qyb, V.Hf), syc = to_screen(qyc, V.Hf);
            const float lox = fmaxf(fminf(sxa, fminf(sxb, sxc)), 0.0f);
            const float hix = fminf(fmaxf(sxa, fmaxf(sxb, sxc)), V.Wm1);
            const float loy = fmaxf(fminf(sya, fminf(syb, syc)), 0.0f);
            const float hiy = fminf(fmaxf(sya, fmaxf(syb, syc)), V.Hm1);
            if (surely_culled_axis(lox, hix, V.tinyMargin) || surely_culled_axis(loy, hiy, V.tinyMargin)) return;
        }
    }

    // viewToNDC (rasterization_helpers.cuh:95-101) and ndcToScreen - 0.5 (:103-108, :132-134)
    // The quotients here, the depth reciprocals and every per-pixel depth
    // reciprocal below take the short sequences (rcp_refined, div_fast) when
    // every active lane of the wave has its depths in [kDivLo, kDivHi] and
    // its numerators within kDivHi: then all of them are the IEEE values (a
    // per-pixel inv_depth is a convex combination of the depth reciprocals,
    // inside [2^-66, 2^62]).  Other waves take the IEEE division throughout.
    const float Wf = V.Wf, Hf = V.Hf;
    const float dmin = fminf(da, fminf(db, dc)), dmax = fmaxf(da, fmaxf(db, dc));
    const float pmax = fmaxf(fmaxf(fmaxf(fabsf(pxa), fabsf(pxb)), fmaxf(fabsf(pxc), fabsf(pya))),
                             fmaxf(fabsf(pyb), fabsf(pyc)));
    const bool fast = __ballot(!(dmin >= kDivLo && dmax <= kDivHi && pmax <= kDivHi)) == 0ull;
    float ax, bx, cx, ay, by, cy, iza, izb, izc;
    if (fast) {
        iza = rcp_refined(da);
        izb = rcp_refined(db);
        izc = rcp_refined(dc);
        ax = (div_fast(pxa, da, iza) * 0.5f + 0.5f) * Wf - 0.5f;
        bx = (div_fast(pxb, db, izb) * 0.5f + 0.5f) * Wf - 0.5f;
        cx = (div_fast(pxc, dc, izc) * 0.5f + 0.5f) * Wf - 0.5f;
        ay = (div_fast(pya, da, iza) * 0.5f + 0.5f) * Hf - 0.5f;
        by = (div_fast(pyb, db, izb) * 0.5f + 0.5f) * Hf - 0.5f;
        cy = (div_fast(pyc, dc, izc) * 0.5f + 0.5f) * Hf - 0.5f;
    } else {
        ax = ((pxa / da) * 0.5f + 0.5f) * Wf - 0.5f;
        bx = ((pxb / db) * 0.5f + 0.5f) * Wf - 0.5f;
        cx = ((pxc / dc) * 0.5f + 0.5f) * Wf - 0.5f;
        ay = ((pya / da) * 0.5f + 0.5f) * Hf - 0.5f;
        by = ((pyb / db) * 0.5f + 0.5f) * Hf - 0.5f;
        cy = ((pyc / dc) * 0.5f + 0.5f) * Hf - 0.5f;
        iza = 1.0f / da;
        izb = 1.0f / db;
        izc = 1.0f / dc;
    }

    float min_x = fminf(ax, fminf(bx, cx));
    float max_x = fmaxf(ax, fmaxf(bx, cx));
    float min_y = fminf(ay, fminf(by, cy));
    float max_y = fmaxf(ay, fmaxf(by, cy));
    min_x = fmaxf(min_x, 0.0f);
    max_x = fminf(max_x, V.Wm1);
    min_y = fmaxf(min_y, 0.0f);
    max_y = fminf(max_y, V.Hm1);

    // tiny-triangle cull (:153-158)
    if (!nontrivial) {
        const float sx = floorf(min_x);
        const float sy = floorf(min_y);
        if (min_x > sx + 0.0f && max_x < sx + 1.0f + 0.0f) return;
        if (min_y > sy + 0.0f && max_y < sy + 1.0f + 0.0f) return;
    }

    const int size_x = f2i(ceilf(max_x) - ceilf(min_x));
    const int size_y = f2i(ceilf(max_y) - ceilf(min_y));
    const uint32_t numFragments = (uint32_t)size_x * (uint32_t)size_y;
    if (numFragments > P.smallThreshold) nontrivial = true;   // unsigned compare, as the reference

    if (nontrivial) {
        const uint32_t slot = atomicAdd(P.ctrNontrivial, 1u);
        if (slot < P.ntCap) P.nontrivial[slot] = queueEntry;
        return;
    }

    // per-lane incremental rasterisation (:200-269)
    const float vabx = bx - ax, vaby = by - ay;
    const float vacx = cx - ax, vacy = cy - ay;
    const float factor = rcp_rn(vabx * vacy - vaby * vacx);
    const float ds_dx = vacy * factor;
    const float ds_dy = -vacx * factor;
    const float dt_dx = -vaby * factor;
    const float dt_dy = vabx * factor;
    const float start_x = ceilf(min_x);
    const float start_y = ceilf(min_y);
    const float s0x = start_x - ax + 0.0f;
    const float s0y = start_y - ay + 0.0f;
    float s_row = (s0x * vacy - s0y * vacx) * factor;
    float t_row = (vabx * s0y - vaby * s0x) * factor;
    uint32_t rowPix = (uint32_t)f2i(start_x) + (uint32_t)V.W * (uint32_t)f2i(start_y);
    // The reference also tests pixelID < W*H; every sample here is inside
    // the screen (the box is clamped to [0, W-1] x [0, H-1] before ceil, and
    // a NaN bound clamps to an edge), so that test is always true.
    for (int fy = 0; fy < size_y; ++fy) {
        float s = s_row, t = t_row;
        uint32_t pix = rowPix;
        for (int fx = 0; fx < size_x; ++fx) {
            const float v = 1.0f - (s + t);
            if (s > 0.0f && t > 0.0f && v > 0.0f) {
                const float inv_depth = (v * iza + s * izb) + t * izc;
                uint64_t packed;
                if (fast) {   // wave-uniform branch: a finite depth, no infinity to map
                    const uint32_t bits = __builtin_bit_cast(uint32_t, rcp_refined(inv_depth)) & 0x7FFFFFF8u;
                    packed = ((uint64_t)bits << 33) | id;
                } else {
                    packed = pack_pixel(1.0f / inv_depth, id);
                }
                atomic_min_fb(V.fb, pix, packed);
            }
            s += ds_dx;
            t += dt_dx;
            pix++;
        }
        s_row += ds_dy;
        t_row += dt_dy;
        rowPix += (uint32_t)V.W;
    }
}

__device__ __forceinline__ void stage1_triangle(const RasterParams& P, const ViewParams& V,
                                                const float* __restrict__ m,
                                                bool flip, f3 A, f3 B, f3 C,
                                                uint64_t id, uint64_t queueEntry, uint32_t cullOrder) {
    // depth first: ndc.z = depth = -view.z
    VQ qa{}, qb{}, qc{};
    qa.d = -xform_row(m, 2, A);
    qb.d = -xform_row(m, 2, B);
    qc.d = -xform_row(m, 2, C);
    stage1_core<false>(P, V, m, flip, A, B, C, qa, qb, qc, id, queueEntry, cullOrder);
}

}  // namespace

// Clear one or two buffers (a view's visibility and colour buffers) to
// 0xFFFFFFF0'00000000 (depth +inf, ID 0), one pixel per lane (a wave stores
// 512 contiguous bytes), and reset the draw counters (graph-safe: no memset
// node).
__global__ __launch_bounds__(256) void k_clear(uint64_t* __restrict__ fb, uint32_t numPixels, uint64_t* fb2,
                                               uint32_t numPixels2, uint32_t* __restrict__ ctr) {
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t first = blockIdx.x * blockDim.x + threadIdx.x;
    for (uint32_t i = first; i < numPixels; i += stride) fb[i] = kClearPixel;
    for (uint32_t i = first; i < numPixels2; i += stride) fb2[i] = kClearPixel;
    if (blockIdx.x == 0 && ctr)
        for (uint32_t i = threadIdx.x; i < CTR_COUNT; i += blockDim.x) ctr[i] = 0u;
}

__global__ __launch_bounds__(64) void k_reset_counters(uint32_t* __restrict__ ctr) {
    for (uint32_t i = threadIdx.x; i < CTR_COUNT; i += blockDim.x) ctr[i] = 0u;
}

// Counter reset of the kernel-level entry points: the five counters the
// reference's stage 1 zeroes before its grid barrier
// (triangles_visbuffer.cu:280-286), each at its own caller address.
__global__ __launch_bounds__(64) void k_reset_counter_set(CounterSet c) {
    if (threadIdx.x < 5u && c.p[threadIdx.x]) c.p[threadIdx.x][0] = 0u;
}

// Device-built plan of the kernel-level entry points, from caller-owned
// CMesh arrays (the same records plan_draw builds on the host).  One
// workgroup: each thread takes a contiguous run of items, sums their batch
// counts, a workgroup scan turns the sums into each run's first batch, and
// the thread writes its DrawItems; thread 0 writes the batch count.  Items
// whose instance range lies outside the instance list, or whose format is
// not the draw's, are given no triangles.
__global__ __launch_bounds__(1024) void k_plan_items(const crx_cmesh* __restrict__ list, uint32_t numItems,
                                                     uint32_t numInstances, int instanced, int pu16,
                                                     uint32_t batchShift, DrawItem* __restrict__ items,
                                                     uint32_t* __restrict__ batchCount) {
    __shared__ uint64_t sums[1024];
    const uint32_t t = threadIdx.x;
    const uint32_t per = (numItems + 1023u) / 1024u;
    const uint32_t lo = min(numItems, t * per), hi = min(numItems, lo + per);
    auto triangles = [&](uint32_t k) -> uint32_t {
        const crx_cmesh& m = list[k];
        const int64_t off = instanced ? (int64_t)m.instances.offset : (int64_t)k;
        const int64_t cnt = instanced ? (int64_t)m.instances.count : 1;
        const bool ok = off >= 0 && cnt >= 0 && off + cnt <= (int64_t)numInstances && (m.compressed != 0) == (pu16 != 0) &&
                        m.positions != nullptr && (!pu16 || (m.indices != nullptr && m.bitsPerIndex <= 32u)) &&
                        m.numTriangles <= 1431655765u;
        return ok ? m.numTriangles : 0u;
    };
    uint64_t sum = 0;
    for (uint32_t k = lo; k < hi; k++) sum += ((uint64_t)triangles(k) + (1ull << batchShift) - 1u) >> batchShift;
    sums[t] = sum;
    __syncthreads();
    for (uint32_t d = 1; d < 1024u; d <<= 1) {
        const uint64_t add = t >= d ? sums[t - d] : 0ull;
        __syncthreads();
        sums[t] += add;
        __syncthreads();
    }
    uint64_t first = sums[t] - sum;
    for (uint32_t k = lo; k < hi; k++) {
        const crx_cmesh& m = list[k];
        const uint32_t n = triangles(k);
        DrawItem it{};
        it.indices = m.indices ? m.indices : reinterpret_cast<const uint32_t*>(m.positions);
        it.indexMask = m.indices ? 0xFFFFFFFFu : 0u;
        it.positions = reinterpret_cast<const float*>(m.positions);
        it.cumm = m.cummulativeTriangleCount;
        it.numTriangles = n;
        it.instOffset = instanced ? (uint32_t)m.instances.offset : k;
        it.instCount = n == 0u ? 0u : (instanced ? (uint32_t)m.instances.count : 1u);
        it.firstBatch = (uint32_t)min(first, 0xFFFFFFFFull);
        it.bitsPerIndex = m.bitsPerIndex;
        it.indexMin = m.index_min;
        it.cf[0] = m.compressionFactor.x; it.cf[1] = m.compressionFactor.y; it.cf[2] = m.compressionFactor.z;
        it.amin[0] = m.aabb.min.x; it.amin[1] = m.aabb.min.y; it.amin[2] = m.aabb.min.z;
        it.numVertices = m.numVertices;
        // (the pair layout is not built here: these plans run k_stage1)
        it.rowTris = 128u;
        it.strips = 1u;
        it.unitRows = 1u;
        it.unitSegs = 1u;
        it.unitCols = 1u;
        items[k] = it;
        first += ((uint64_t)n + (1ull << batchShift) - 1u) >> batchShift;
    }
    if (t == 1023u) *batchCount = (uint32_t)min(sums[1023], 0xFFFFFFFFull);
}

// The view record of a kernel-level call, stored from the kernel argument
// (no host staging buffer, so the call needs no page-locked memory).
__global__ __launch_bounds__(64) void k_store_view(ViewParams v, ViewParams* __restrict__ dst) {
    if (threadIdx.x == 0) *dst = v;
}

// One lane per instance: its InstanceRec (make_instance on the host) and its
// Transform with the mirror flag det(worldView[0..2]) < 0 computed in the
// same float operations (triangles_visbuffer.cu:179-183).
__global__ __launch_bounds__(256) void k_plan_instances(const crx_cmesh* __restrict__ instances,
                                                        const crx_mat4* __restrict__ transforms,
                                                        uint32_t numInstances, InstanceRec* __restrict__ inst,
                                                        Transform* __restrict__ xf) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= numInstances) return;
    const crx_cmesh& m = instances[i];
    InstanceRec r{};
    r.indices = m.indices;
    r.positions = reinterpret_cast<const float*>(m.positions);
    r.colors = m.colors;
    r.normals = reinterpret_cast<const float*>(m.normals);
    r.uvs = reinterpret_cast<const float*>(m.uvs);
    r.cumm = m.cummulativeTriangleCount;
    r.numTriangles = m.numTriangles;
    const int32_t ti = m.instances.offset;
    r.transformIndex = (ti >= 0 && (uint32_t)ti < numInstances) ? (uint32_t)ti : i;
    r.id = m.id;
    r.isLoaded = m.isLoaded ? 1u : 0u;
    for (int e = 0; e < 16; e++) r.world[e] = m.world.m[e];
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
    inst[i] = r;
    Transform x{};
    for (int e = 0; e < 16; e++) x.m[e] = transforms[i].m[e];
    const f3 c0{x.m[0], x.m[1], x.m[2]}, c1{x.m[4], x.m[5], x.m[6]}, c2{x.m[8], x.m[9], x.m[10]};
    x.flip = dot3(cross3(c0, c1), c2) < 0.0f ? 1u : 0u;
    x.skip = 0u;
    xf[i] = x;
}

namespace {


// last draw item whose firstBatch <= batch
__device__ __forceinline__ uint32_t find_item(const RasterParams& P, uint32_t batch) {
    const auto items = kconst(P.items);
    uint32_t lo = 0, hi = P.numItems - 1u;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1u) >> 1;
        if (items[mid].firstBatch <= batch) lo = mid; else hi = mid - 1u;
    }
    return lo;
}

// Uniform fields of the draw item a batch belongs to, held in scalar
// registers (plain scalars, so the compiler never stages them in memory).
struct S1Item {
    const uint32_t* indices;    // non-indexed: the positions
    const float* positions;
    uint64_t cumm;
    uint32_t indexMask;         // ~0 = indexed, 0 = non-indexed
    uint32_t numTriangles, instOffset, instCount;
    uint32_t firstBatch, endBatch;
    uint32_t bits, indexMin;    // IX_PU16
    float cfx, cfy, cfz, mnx, mny, mnz;
    uint32_t numVertices;
};

// First batch past item `item`.
__device__ __forceinline__ uint32_t item_end(const RasterParams& P, uint32_t item, uint32_t numBatches) {
    return item + 1u < P.numItems ? kconst(P.items)[item + 1u].firstBatch : numBatches;
}

__device__ __forceinline__ S1Item s1_item(const RasterParams& P, uint32_t item, uint32_t numBatches) {
    const auto items = kconst(P.items);
    S1Item r;
    r.indices = items[item].indices;
    r.indexMask = items[item].indexMask;
    r.positions = items[item].positions;
    r.cumm = items[item].cumm;
    r.numTriangles = items[item].numTriangles;
    r.instOffset = items[item].instOffset;
    r.instCount = items[item].instCount;
    r.firstBatch = items[item].firstBatch;
    r.endBatch = item_end(P, item, numBatches);
    r.bits = items[item].bitsPerIndex;
    r.indexMin = items[item].indexMin;
    r.cfx = items[item].cf[0];
    r.cfy = items[item].cf[1];
    r.cfz = items[item].cf[2];
    r.mnx = items[item].amin[0];
    r.mny = items[item].amin[1];
    r.mnz = items[item].amin[2];
    r.numVertices = items[item].numVertices;
    return r;
}

// BitEdit::readU32 (BitEdit.h:8-26): `bits` bits (1..32) at bit `off`, the
// second word read only when the value straddles it.
__device__ __forceinline__ uint32_t bit_read(const uint32_t* data, uint64_t off, uint32_t bits) {
    if (bits == 0u || bits > 32u) return 0u;
    const auto d = kglobal(data);
    const uint64_t w = off >> 5;
    const uint32_t sh = (uint32_t)off & 31u;
    uint64_t combined = d[w];
    if (sh + bits > 32u) combined |= (uint64_t)d[w + 1] << 32;
    const uint32_t mask = bits == 32u ? 0xFFFFFFFFu : ((1u << bits) - 1u);
    return (uint32_t)(combined >> sh) & mask;
}

// Issue the index loads of triangle `tri` of the item (uncompressed formats).
// Triangles past the item's end load its last triangle instead (a valid
// address) and are masked by the caller.
__device__ __forceinline__ void s1_indices(const S1Item& it, uint32_t tri, uint32_t& i0, uint32_t& i1, uint32_t& i2) {
    const uint32_t t = min(tri, it.numTriangles - 1u);
    // A non-indexed item (vertices 3t..3t+2) points `indices` at its
    // positions and has a zero mask: the words read there are discarded.
    // The load is one dwordx3 on every path, so the vertex waits can
    // leave it in flight (the mask is data, not a branch).
    const auto ip = kglobal(it.indices) + 3ull * t;
    const uint32_t r0 = ip[0], r1 = ip[1], r2 = ip[2];
    const uint32_t mk = it.indexMask;
    i0 = (r0 & mk) | ((3u * t) & ~mk);
    i1 = (r1 & mk) | ((3u * t + 1u) & ~mk);
    i2 = (r2 & mk) | ((3u * t + 2u) & ~mk);
}

__device__ __forceinline__ f3 load_pos_g(const float* positions, uint32_t i) {
    const auto p = kglobal(positions) + 3ull * i;
    return f3{p[0], p[1], p[2]};
}

// IX_PU16 position: float(X) * compressionFactor + aabb.min per axis
// (triangles_visbuffer.cu:51-59; product then sum, no contraction).
__device__ __forceinline__ f3 load_pos_pu16(const float* positions, uint32_t i, const float* cf, const float* amin) {
    const auto q = kglobal(reinterpret_cast<const uint16_t*>(positions)) + 3ull * i;
    return f3{(float)q[0] * cf[0] + amin[0], (float)q[1] * cf[1] + amin[1], (float)q[2] * cf[2] + amin[2]};
}

// The batches of this workgroup: [begin, end) with stride `step`.
// Workgroup b is dispatched to XCD b % 8, and each XCD has its own L2.
// P.order 0 gives each XCD a contiguous share of the batch list (blocks
// b, b+8, ... split the share); 1 is a plain grid-stride loop; 2 gives each
// workgroup one contiguous range.
__device__ __forceinline__ void s1_range(uint32_t numBatches, uint32_t order, uint32_t& begin, uint32_t& end,
                                         uint32_t& step) {
    const uint32_t G = gridDim.x;
    if (order == 0) {
        const uint32_t perXcd = (numBatches + 7u) >> 3;
        begin = (blockIdx.x & 7u) * perXcd + (blockIdx.x >> 3);
        end = min(numBatches, (blockIdx.x & 7u) * perXcd + perXcd);
        step = G >> 3;                                  // G is a multiple of 8
    } else if (order == 1) {
        begin = blockIdx.x;
        end = numBatches;
        step = G;
    } else {
        const uint32_t per = (numBatches + G - 1u) / G;
        begin = blockIdx.x * per;
        end = min(numBatches, begin + per);
        step = 1;
    }
}

// IX_PU16 in stage 1, per wave group of 64 consecutive triangles
// [tw, tw + 64) (tw clamped into the item, as the lanes' triangles are): the
// group's indices occupy bits [3*bits*tw, 3*bits*(tw + 64)) of the stream,
// at most 6*bits + 1 words.  Every lane loads four consecutive words of that
// range with one coalesced dwordx4 (lane l from word 4l, clamped so that the
// four words stay inside the range and the stream), the wave stores them in
// LDS at their own positions (overlapping lanes store the same words), and
// each lane cuts its three indices out with BitEdit::readU32's arithmetic
// (BitEdit.h:8-26) plus index_min.  A vertex is one 8-byte read at the
// 4-byte boundary below its 6 bytes.  No load sits in a divergent branch, so
// the loads of the next batches stay in flight while a batch is drawn.
constexpr uint32_t kPu16Words = 256;     // LDS words per wave

// Raw buffer loads (buffer_load_dwordx2 / dwordx4 at any dword offset from
// a scalar base).  A dword at or past num_records bytes reads as zero without
// touching memory.  Word 3 of the descriptor is the gfx9 raw-buffer format.
#if defined(__HIP_DEVICE_COMPILE__)
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_rsrc(const void* base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ uint2 buffer_load_u2(__amdgpu_buffer_rsrc_t r, uint32_t offset) {
    return __builtin_bit_cast(uint2, __builtin_amdgcn_raw_buffer_load_b64(r, (int)offset, 0, 0));
}
__device__ __forceinline__ uint4 buffer_load_u4(__amdgpu_buffer_rsrc_t r, uint32_t offset) {
    return __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(r, (int)offset, 0, 0));
}
#endif

// threadIdx.x of the wave's lane 0, as a scalar.
__device__ __forceinline__ uint32_t wave_offset() {
    return __builtin_amdgcn_readfirstlane(threadIdx.x & ~(kWave - 1u));
}

// First word of a wave group's stream range, the bit offset of the group's
// first index inside it and the number of the range's words inside the
// stream (all uniform over the wave).
__device__ __forceinline__ void pu16_group(const S1Item& it, uint32_t groupBase, uint32_t& tw, uint64_t& firstWord,
                                           uint32_t& bit0, uint32_t& avail) {
    tw = min(groupBase, it.numTriangles - 1u);
    const uint64_t b = (uint64_t)it.bits * 3ull * tw;
    firstWord = b >> 5;
    bit0 = (uint32_t)b & 31u;
    const uint64_t numWords = ((uint64_t)it.numTriangles * 3ull * it.bits + 31u) >> 5;
    const uint32_t groupWords = (bit0 + 192u * it.bits + 31u) >> 5;
    avail = (uint32_t)min((uint64_t)groupWords, numWords - firstWord);
}

__device__ __forceinline__ void pu16_load_words(const S1Item& it, uint32_t groupBase, uint32_t& w0, uint32_t& w1,
                                                uint32_t& w2, uint32_t& w3) {
    w0 = w1 = w2 = w3 = 0u;
#if defined(__HIP_DEVICE_COMPILE__)
    if (it.bits == 0u || it.bits > 32u) return;
    uint32_t tw, bit0, avail;
    uint64_t firstWord;
    pu16_group(it, groupBase, tw, firstWord, bit0, avail);
    // lane l reads words 4l .. 4l + 3 of the group's range; those past the
    // range (or the stream) read as zero
    const uint4 v = buffer_load_u4(buffer_rsrc(it.indices + firstWord, 4u * avail), 16u * (threadIdx.x & (kWave - 1)));
    w0 = v.x;
    w1 = v.y;
    w2 = v.z;
    w3 = v.w;
#else
    (void)it;
    (void)groupBase;
#endif
}

__device__ __forceinline__ void pu16_decode(const S1Item& it, uint32_t groupBase, uint32_t* words, uint32_t w0,
                                            uint32_t w1, uint32_t w2, uint32_t w3, uint32_t& i0, uint32_t& i1,
                                            uint32_t& i2) {
    const uint32_t bits = it.bits;
    if (bits == 0u || bits > 32u) {                     // BitEdit::readU32 reads nothing
        i0 = i1 = i2 = it.indexMin;
        return;
    }
    uint32_t tw, bit0, avail;
    uint64_t firstWord;
    pu16_group(it, groupBase, tw, firstWord, bit0, avail);
    *reinterpret_cast<uint4*>(words + 4u * (threadIdx.x & (kWave - 1))) = make_uint4(w0, w1, w2, w3);
    wave_sync();
    const uint32_t t = min(groupBase + (threadIdx.x & (kWave - 1)), it.numTriangles - 1u);
    // bit offset of the lane's first index from the group's first word (< 32 * 194)
    uint32_t off = bit0 + 3u * bits * (t - tw);
    const uint32_t mask = bits >= 32u ? 0xFFFFFFFFu : ((1u << bits) - 1u);
    uint32_t v[3];
    #pragma unroll
    for (int j = 0; j < 3; j++) {
        // (hi:lo) >> sh of the two words holding the index; the second word is
        // inside the LDS words even when the index does not reach into it
        const uint32_t wd = off >> 5;
        const uint32_t lo = words[wd], hi = words[wd + 1];
        v[j] = __builtin_amdgcn_alignbit(hi, lo, off & 31u) & mask;
        off += bits;
    }
    wave_sync();
    i0 = v[0] + it.indexMin;
    i1 = v[1] + it.indexMin;
    i2 = v[2] + it.indexMin;
}

// IX_PU16 vertex i, issued: the two words of the 8-byte window at the 4-byte
// boundary below its 6 bytes (byte 6i, or 6i - 2 for odd i), carried as the
// bits of an f3 (window words, flag) until pu16_vertex_decode.  Every vertex
// but an even last one has its window inside the numVertices * 6 bytes of
// the buffer; that one reads a dummy in-range window here (flag 2) and its
// three uint16 in pu16_vertex_decode, in a branch only waves holding it
// enter.  Meshes of fewer than three vertices, or of more than 2^32 bytes
// (byte offsets of the buffer load are 32-bit), read uint16s directly.
__device__ __forceinline__ f3 pu16_vertex_issue(const S1Item& it, uint32_t i) {
    uint32_t lo = 0u, hi = 0u, flag = 0u;
#if defined(__HIP_DEVICE_COMPILE__)
    const auto base = reinterpret_cast<const char*>(it.positions);
    const uint64_t bytes = 6ull * it.numVertices;
    if (it.numVertices >= 3u && bytes < 0xFFFFFFFFull) {
        // an odd vertex's window ends at 6i + 6 <= 6V; an even one's at 6i + 8,
        // inside the buffer unless i is the last vertex.  (An index at or past
        // V, which no valid stream holds, reads some in-range window or zeros:
        // the buffer load never leaves the buffer.)
        const uint32_t odd = i & 1u;
        const bool fast = odd != 0u || i + 2u <= it.numVertices;
        const uint32_t at = 6u * i - (fast ? 2u * odd : 4u);
        const uint2 w = buffer_load_u2(buffer_rsrc(base, (uint32_t)bytes), at);
        lo = w.x;
        hi = w.y;
        flag = fast ? odd : 2u;
    } else {
        const auto h = kglobal(reinterpret_cast<const uint16_t*>(base)) + 3ull * i;
        lo = (uint32_t)h[0] | ((uint32_t)h[1] << 16);
        hi = h[2];
    }
#else
    (void)it;
    (void)i;
#endif
    return f3{__builtin_bit_cast(float, lo), __builtin_bit_cast(float, hi), __builtin_bit_cast(float, flag)};
}

// X * compressionFactor + aabb.min per axis (triangles_visbuffer.cu:51-59;
// product then sum, no contraction) from an issued window.
__device__ __forceinline__ f3 pu16_vertex_decode(const S1Item& it, const f3& raw) {
    uint32_t lo = __builtin_bit_cast(uint32_t, raw.x), hi = __builtin_bit_cast(uint32_t, raw.y);
    uint32_t flag = __builtin_bit_cast(uint32_t, raw.z);
    if (__ballot(flag == 2u) != 0ull && flag == 2u) {  // the even last vertex
        const auto h = kglobal(reinterpret_cast<const uint16_t*>(it.positions)) + 3ull * (it.numVertices - 1u);
        lo = (uint32_t)h[0] | ((uint32_t)h[1] << 16);
        hi = h[2];
        flag = 0u;
    }
    // the three uint16 sit at bits 16*odd + {0, 16, 32} of the window hi:lo
    const uint32_t sh = flag << 4;
    const uint32_t m = __builtin_amdgcn_alignbit(hi, lo, sh);
    const uint32_t x = m & 0xFFFFu;
    const uint32_t y = m >> 16;
    const uint32_t z = __builtin_amdgcn_ubfe(hi, sh, 16u);
    return f3{(float)x * it.cfx + it.mnx, (float)y * it.cfy + it.mny, (float)z * it.cfz + it.mnz};
}

// Index loads of the K triangles of `batch` for this lane (IX_PU16: the raw
// stream words of the lane's share of its wave groups, in i0..i3).
template <int BLOCK, int K, bool PU16>
__device__ __forceinline__ void s1_batch_indices(const RasterParams& P, const S1Item& it, uint32_t batch,
                                                 uint32_t (&i0)[K], uint32_t (&i1)[K], uint32_t (&i2)[K],
                                                 uint32_t (&i3)[K]) {
    const uint32_t base = (batch - it.firstBatch) << P.batchShift;
    #pragma unroll
    for (int k = 0; k < K; k++) {
        if (PU16) pu16_load_words(it, base + k * BLOCK + wave_offset(), i0[k], i1[k], i2[k], i3[k]);
        else s1_indices(it, base + k * BLOCK + threadIdx.x, i0[k], i1[k], i2[k]);
    }
}

// Vertex loads of `batch` (IX_PU16: after cutting the indices out of the
// stream words through the wave's LDS words; the vertices stay raw windows
// until pu16_vertex_decode).
template <int BLOCK, int K, bool PU16>
__device__ __forceinline__ void s1_batch_vertices(const RasterParams& P, const S1Item& it, uint32_t batch,
                                                  const uint32_t (&i0)[K], const uint32_t (&i1)[K],
                                                  const uint32_t (&i2)[K], const uint32_t (&i3)[K], uint32_t* words,
                                                  f3 (&A)[K], f3 (&B)[K], f3 (&C)[K]) {
    const uint32_t base = (batch - it.firstBatch) << P.batchShift;
    #pragma unroll
    for (int k = 0; k < K; k++) {
        if (PU16) {
            uint32_t a, b, c;
            pu16_decode(it, base + k * BLOCK + wave_offset(), words, i0[k], i1[k], i2[k], i3[k], a, b, c);
            A[k] = pu16_vertex_issue(it, a);
            B[k] = pu16_vertex_issue(it, b);
            C[k] = pu16_vertex_issue(it, c);
        } else {
            A[k] = load_pos_g(it.positions, i0[k]);
            B[k] = load_pos_g(it.positions, i1[k]);
            C[k] = load_pos_g(it.positions, i2[k]);
        }
    }
}

// The next batch of the sweep after `batch`, or `batch` itself (reloaded,
// a valid address) when there is none; *more tells which.
__device__ __forceinline__ uint32_t s1_next(uint32_t batch, uint32_t step, uint32_t end, bool* more) {
    const uint32_t next = batch + step;
    *more = next < end && next > batch;
    return *more ? next : batch;
}

// The batch loop of both stage-1 kernels.  A batch is BLOCK*K consecutive
// triangles of one draw item; lane t of the workgroup takes triangles
// t, t + BLOCK, ..., t + (K-1)*BLOCK (each of a wave's index loads reads 768
// contiguous bytes).  The loop is software-pipelined two deep: while batch i
// is drawn, the vertex loads of batch i+1 and the index loads of batch i+2
// are in flight, so a wave's memory latency overlaps its own arithmetic.
// All loads are issued on every iteration (past the last batch they reload
// it), so the waits for batch i never have to cover the younger loads.
// body(it, base, A, B, C) draws the batch whose first triangle is `base`
// within item `it`.
template <int BLOCK, int K, bool PU16, typename Body>
__device__ __forceinline__ void s1_sweep(const RasterParams& P, uint32_t* words, Body&& body) {
    const uint32_t numBatches = P.batchCount ? *kconst(P.batchCount) : P.numBatches;
    uint32_t begin, end, step;
    s1_range(numBatches, P.order, begin, end, step);
    if (begin >= end) return;
    // One pass per item: within a pass the item's fields are loop-invariant
    // scalars (holding the items of the batches in flight as well would spill
    // scalar registers), and the pipeline drains at the item's end.
    uint32_t batch = begin;
    for (;;) {
        const S1Item it = s1_item(P, find_item(P, batch), numBatches);
        const uint32_t iend = min(end, it.endBatch);
        uint32_t i0[K], i1[K], i2[K], i3[K];
        s1_batch_indices<BLOCK, K, PU16>(P, it, batch, i0, i1, i2, i3);
        bool more;
        uint32_t b1 = s1_next(batch, step, iend, &more);
        uint32_t j0[K], j1[K], j2[K], j3[K];
        s1_batch_indices<BLOCK, K, PU16>(P, it, b1, j0, j1, j2, j3);
        f3 A[K], B[K], C[K];
        s1_batch_vertices<BLOCK, K, PU16>(P, it, batch, i0, i1, i2, i3, words, A, B, C);
        for (;;) {
            f3 An[K], Bn[K], Cn[K];
            s1_batch_vertices<BLOCK, K, PU16>(P, it, b1, j0, j1, j2, j3, words, An, Bn, Cn);
            bool more2;
            const uint32_t b2 = more ? s1_next(b1, step, iend, &more2) : b1;
            if (!more) more2 = false;
            s1_batch_indices<BLOCK, K, PU16>(P, it, b2, j0, j1, j2, j3);
            if (PU16) {
                #pragma unroll
                for (int k = 0; k < K; k++) {
                    A[k] = pu16_vertex_decode(it, A[k]);
                    B[k] = pu16_vertex_decode(it, B[k]);
                    C[k] = pu16_vertex_decode(it, C[k]);
                }
            }
            body(it, (batch - it.firstBatch) << P.batchShift, A, B, C);
            if (!more) break;
            batch = b1;
            #pragma unroll
            for (int k = 0; k < K; k++) {
                A[k] = An[k];
                B[k] = Bn[k];
                C[k] = Cn[k];
            }
            b1 = b2;
            more = more2;
        }
        const uint32_t next = batch + step;
        if (next >= end || next <= batch) break;
        batch = next;
    }
}


// Order-preserving integer key of a float: signed-integer order of the keys
// is the numeric order of the non-NaN floats (-0 sorts just below +0).  A
// positive NaN sorts above +inf, a negative NaN below -inf.
__device__ __forceinline__ int32_t fkey(float f) {
    const int32_t b = __builtin_bit_cast(int32_t, f);
    return b ^ ((b >> 31) & 0x7FFFFFFF);
}

__device__ __forceinline__ float fkey_inv(int32_t k) {
    return __builtin_bit_cast(float, k ^ ((k >> 31) & 0x7FFFFFFF));
}

constexpr int32_t kKeyPosInf = 0x7F800000;                  // fkey(+inf)
constexpr int32_t kKeyNegInf = (int32_t)0xFF800000 ^ 0x7FFFFFFF;   // fkey(-inf)

// Integer minimum / maximum over the 64 lanes (all active), as a scalar.
// Each step is one v_min_i32 / v_max_i32 with a DPP operand: quad exchanges
// and row rotations reduce each row of 16 lanes, row broadcasts fold rows
// 0-1 into 1 and 2-3 into 3, then rows 0-1 into 2-3; lane 63 holds the
// result.  Lanes a DPP source does not feed read the identity, which lets
// the compiler fold each move into the min / max.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ int32_t dpp_or(int32_t identity, int32_t x) {
    return __builtin_amdgcn_update_dpp(identity, x, CTRL, ROWMASK, 0xF, false);
}

// Minimum / maximum over each group of G = 8 or 16 consecutive lanes; every
// lane of the group ends with the group's value.  Quad exchanges reduce each
// quad; the half-row mirror folds quads 0-1 and 2-3 of a row (G = 8), row
// rotations by 4 and 8 fold the whole row (G = 16).
template <int G>
__device__ __forceinline__ int32_t group_min_i(int32_t x) {
    static_assert(G == 8 || G == 16, "lane groups of 8 or 16");
    constexpr int32_t I = 0x7FFFFFFF;
    x = min(x, dpp_or<0xB1, 0xF>(I, x));    // quad_perm [1,0,3,2]
    x = min(x, dpp_or<0x4E, 0xF>(I, x));    // quad_perm [2,3,0,1]
    if constexpr (G == 8) {
        x = min(x, dpp_or<0x141, 0xF>(I, x));   // row_half_mirror
    } else {
        x = min(x, dpp_or<0x124, 0xF>(I, x));   // row_ror:4
        x = min(x, dpp_or<0x128, 0xF>(I, x));   // row_ror:8
    }
    return x;
}

template <int G>
__device__ __forceinline__ int32_t group_max_i(int32_t x) {
    static_assert(G == 8 || G == 16, "lane groups of 8 or 16");
    constexpr int32_t I = (int32_t)0x80000000;
    x = max(x, dpp_or<0xB1, 0xF>(I, x));
    x = max(x, dpp_or<0x4E, 0xF>(I, x));
    if constexpr (G == 8) {
        x = max(x, dpp_or<0x141, 0xF>(I, x));
    } else {
        x = max(x, dpp_or<0x124, 0xF>(I, x));
        x = max(x, dpp_or<0x128, 0xF>(I, x));
    }
    return x;
}

__device__ __forceinline__ int32_t wave_min_i(int32_t x) {
    constexpr int32_t I = 0x7FFFFFFF;
    x = group_min_i<16>(x);
    x = min(x, dpp_or<0x142, 0xA>(I, x));   // row_bcast:15 into rows 1, 3
    x = min(x, dpp_or<0x143, 0xC>(I, x));   // row_bcast:31 into rows 2, 3
    return __builtin_amdgcn_readlane(x, 63);
}

__device__ __forceinline__ int32_t wave_max_i(int32_t x) {
    constexpr int32_t I = (int32_t)0x80000000;
    x = group_max_i<16>(x);
    x = max(x, dpp_or<0x142, 0xA>(I, x));
    x = max(x, dpp_or<0x143, 0xC>(I, x));
    return __builtin_amdgcn_readlane(x, 63);
}

// Object-space box of a wave's 64 triangles.
struct WaveBox {
    f3 lo, hi;
    float M;   // max |coordinate| over the box
};

// Culling planes of one view for one instance, from its world-to-view
// transform m: g(p) = a.p + b, with g > 0 meaning "outside".
//   x+ : fa*x_view + z_view    (RN(fa*x/d) > 1,  d = -z_view)
//   x- : -fa*x_view + z_view   (RN(fa*x/d) < -1)
//   y+ / y- : the same with f and y_view
//   z  : z_view                (d < 0: behind the camera)
// slope/cnst bound the magnitudes that enter the rounding errors:
// S_x(p) = sum_k |m_xk||p_k| + |m_x3| <= A_x*M + B_x.
struct ViewCull {
    float px[4], nx[4], py[4], ny[4], z[4];
    float slopeX, cnstX, slopeY, cnstY, slopeZ, cnstZ;
    bool skip;   // the instance is dropped from this view
};

__device__ __forceinline__ ViewCull view_cull(const float* m, float fa, float f) {
    ViewCull c;
    c.skip = false;
    #pragma unroll
    for (int k = 0; k < 4; k++) {
        const float x = m[4 * k], y = m[4 * k + 1], z = m[4 * k + 2];
        c.px[k] = fa * x + z;
        c.nx[k] = -(fa * x) + z;
        c.py[k] = f * y + z;
        c.ny[k] = -(f * y) + z;
        c.z[k] = z;
    }
    const float A0 = (fabsf(m[0]) + fabsf(m[4])) + fabsf(m[8]);
    const float A1 = (fabsf(m[1]) + fabsf(m[5])) + fabsf(m[9]);
    const float A2 = (fabsf(m[2]) + fabsf(m[6])) + fabsf(m[10]);
    c.slopeX = fabsf(fa) * A0 + A2;
    c.cnstX = fabsf(fa) * fabsf(m[12]) + fabsf(m[14]);
    c.slopeY = fabsf(f) * A1 + A2;
    c.cnstY = fabsf(f) * fabsf(m[13]) + fabsf(m[14]);
    c.slopeZ = A2;
    c.cnstZ = fabsf(m[14]);
    return c;
}

// min over the box of a.p + b: per coordinate the smaller of the two corners.
__device__ __forceinline__ float box_min(const float* a, const WaveBox& B) {
    return ((fminf(a[0] * B.lo.x, a[0] * B.hi.x) + fminf(a[1] * B.lo.y, a[1] * B.hi.y)) +
            fminf(a[2] * B.lo.z, a[2] * B.hi.z)) + a[3];
}

// max over the box of a.p + b.
__device__ __forceinline__ float box_max(const float* a, const WaveBox& B) {
    return ((fmaxf(a[0] * B.lo.x, a[0] * B.hi.x) + fmaxf(a[1] * B.lo.y, a[1] * B.hi.y)) +
            fmaxf(a[2] * B.lo.z, a[2] * B.hi.z)) + a[3];
}

// Range of the computed view coordinate of row r of m over the box.  With
// S = (|m_r0| + |m_r1| + |m_r2|) * M + |m_r3|, the per-triangle code's
// xform_row and this bound are each within 8u*S of the exact affine value
// (u = 2^-24: three products and three sums); the widening is 2^-18*S, 8x more.
__device__ __forceinline__ void row_range(const float* m, int r, const WaveBox& B, float& lo, float& hi) {
    const float a[4] = {m[r], m[4 + r], m[8 + r], m[12 + r]};
    const float S = ((fabsf(a[0]) + fabsf(a[1])) + fabsf(a[2])) * B.M + fabsf(a[3]);
    const float e = S * 0x1p-18f;
    lo = box_min(a, B) - e;
    hi = box_max(a, B) + e;
}

// True when every computed screen coordinate of one axis lies strictly
// between the same two neighbouring sample centres inside the viewport.
// [qlo, qhi] contains every view coordinate / depth quotient of the box;
// the per-triangle code's RN(RN(g*v)/d) is within 2u of g*v/d, our products
// with v_rcp_f32 within a few ulps, and the widening is 2^-16 of the
// magnitude; the screen transform adds a few ulps of `size`, widened by
// 2^-16*size.
__device__ __forceinline__ bool axis_between_samples(float qlo, float qhi, float g, float size, float sizeM1) {
    const float a = g * qlo, b = g * qhi;
    const float w = 0x1p-16f * (fabsf(a) + fabsf(b)) + 0x1p-60f;
    const float nlo = fminf(a, b) - w, nhi = fmaxf(a, b) + w;
    const float ws = size * 0x1p-16f + 0x1p-20f;
    const float slo = ((nlo * 0.5f + 0.5f) * size - 0.5f) - ws;
    const float shi = ((nhi * 0.5f + 0.5f) * size - 0.5f) + ws;
    const float k = floorf(slo);
    return slo > k && shi < k + 1.0f && k >= 0.0f && k + 1.0f <= sizeM1;
}

// True only when stage1_triangle culls every triangle inside the box for the
// transform m and view V by its tiny-triangle test (triangles_visbuffer.cu:
// 153-158) or an earlier early-out: every depth is positive (so no triangle
// is nontrivial), and in x or in y every computed screen coordinate lies
// strictly between the same two sample centres k and k + 1 with 0 <= k and
// k + 1 <= size - 1, so the clamp changes nothing and each triangle's box
// holds no sample.
__device__ __forceinline__ bool wave_tiny_culled(const float* m, const ViewParams& V, const WaveBox& B) {
    float xlo, xhi, ylo, yhi, zlo, zhi;
    row_range(m, 2, B, zlo, zhi);
    const float dlo = -zhi, dhi = -zlo;                 // depth = -view z
    if (!(dlo > 0x1p-60f) || !(dhi < 0x1p60f)) return false;
    const float ilo = __builtin_amdgcn_rcpf(dhi), ihi = __builtin_amdgcn_rcpf(dlo);
    row_range(m, 0, B, xlo, xhi);
    if (axis_between_samples(fminf(xlo * ilo, xlo * ihi), fmaxf(xhi * ilo, xhi * ihi), V.fa, V.Wf, V.Wm1)) return true;
    row_range(m, 1, B, ylo, yhi);
    return axis_between_samples(fminf(ylo * ilo, ylo * ihi), fmaxf(yhi * ilo, yhi * ihi), V.f, V.Hf, V.Hm1);
}

// False only when every triangle inside the box is culled by the exact
// per-triangle tests of stage1_triangle for this view: the box lies behind the
// camera, or in front of it and beyond one side plane, by a margin.
//
// Why the margin suffices: with u = 2^-24, the float plane value at the chosen
// corner is within 8u*(|fa|*S_x + S_z) of the exact affine minimum over the
// box (coefficient roundings, products, three sums), and the per-triangle
// code's own x_view, z_view and fa*x_view are within 4u*S_x, 4u*S_z and one
// more rounding of the exact values; the cull there fires once
// fa*x_view/d > 1 + 2^-22.  All of this is below 22u*(|fa|*S_x + S_z);
// the margin is 2^-16 = 256u times that bound, plus 2^-90 for underflow.
// The side planes are only used when every depth is above that margin too,
// so each vertex has d >= 2^-102 and the division-free frustum test applies.
__device__ __forceinline__ bool view_may_see(const ViewCull& c, const WaveBox& B) {
    const float mz = (c.slopeZ * B.M + c.cnstZ) * 0x1p-16f + 0x1p-90f;
    if (box_min(c.z, B) > mz) return false;                 // behind the camera
    const float negz[4] = {-c.z[0], -c.z[1], -c.z[2], -c.z[3]};
    if (!(box_min(negz, B) > mz)) return true;              // some depth may be small or negative
    const float mx = (c.slopeX * B.M + c.cnstX) * 0x1p-16f + 0x1p-90f;
    const float my = (c.slopeY * B.M + c.cnstY) * 0x1p-16f + 0x1p-90f;
    if (box_min(c.px, B) > mx || box_min(c.nx, B) > mx) return false;
    if (box_min(c.py, B) > my || box_min(c.ny, B) > my) return false;
    return true;
}

__device__ __forceinline__ ViewCull lane_view_cull(const RasterParams& P, uint32_t view, uint32_t inst) {
    const ViewParams* V = P.views + view;
    const Transform& T = V->transforms[inst];
    float mm[16];
    #pragma unroll
    for (int e = 0; e < 16; e++) mm[e] = T.m[e];
    ViewCull c = view_cull(mm, V->fa, V->f);
    c.skip = T.skip != 0u;
    return c;
}

// Object-space box of the 64 triangles (a, b, c) of the wave, reduced on
// order-preserving integer keys.  Returns false when a coordinate is NaN or
// infinite: then there is no box and no box test may drop the wave.
__device__ __forceinline__ bool wave_box(const f3& A, const f3& B, const f3& C, WaveBox& box) {
    int32_t klo[3], khi[3];
    const int32_t ax = fkey(A.x), ay = fkey(A.y), az = fkey(A.z);
    const int32_t bx = fkey(B.x), by = fkey(B.y), bz = fkey(B.z);
    const int32_t cx = fkey(C.x), cy = fkey(C.y), cz = fkey(C.z);
    klo[0] = wave_min_i(min(ax, min(bx, cx)));
    klo[1] = wave_min_i(min(ay, min(by, cy)));
    klo[2] = wave_min_i(min(az, min(bz, cz)));
    khi[0] = wave_max_i(max(ax, max(bx, cx)));
    khi[1] = wave_max_i(max(ay, max(by, cy)));
    khi[2] = wave_max_i(max(az, max(bz, cz)));
    const bool unbounded = min(klo[0], min(klo[1], klo[2])) <= kKeyNegInf ||
                           max(khi[0], max(khi[1], khi[2])) >= kKeyPosInf;
    box.lo = f3{fkey_inv(klo[0]), fkey_inv(klo[1]), fkey_inv(klo[2])};
    box.hi = f3{fkey_inv(khi[0]), fkey_inv(khi[1]), fkey_inv(khi[2])};
    box.M = fmaxf(fmaxf(fmaxf(fabsf(box.lo.x), fabsf(box.hi.x)), fmaxf(fabsf(box.lo.y), fabsf(box.hi.y))),
                  fmaxf(fabsf(box.lo.z), fabsf(box.hi.z)));
    return !unbounded;
}

// Lanes per group of the per-group instance lists of k_stage1 (8 or 16).
constexpr int kBoxGroup = 16;

// Object-space boxes of the kWave / G lane groups of the wave (group g is
// lanes G*g .. G*g + G - 1, one triangle each), as wave-uniform values.
// bounded[g] is false when a coordinate of group g is NaN or infinite.
template <int G>
__device__ __forceinline__ void group_boxes(const f3& A, const f3& B, const f3& C, WaveBox (&box)[kWave / G],
                                            bool (&bounded)[kWave / G]) {
    const int32_t lo[3] = {group_min_i<G>(min(fkey(A.x), min(fkey(B.x), fkey(C.x)))),
                           group_min_i<G>(min(fkey(A.y), min(fkey(B.y), fkey(C.y)))),
                           group_min_i<G>(min(fkey(A.z), min(fkey(B.z), fkey(C.z))))};
    const int32_t hi[3] = {group_max_i<G>(max(fkey(A.x), max(fkey(B.x), fkey(C.x)))),
                           group_max_i<G>(max(fkey(A.y), max(fkey(B.y), fkey(C.y)))),
                           group_max_i<G>(max(fkey(A.z), max(fkey(B.z), fkey(C.z))))};
    #pragma unroll
    for (int g = 0; g < kWave / G; g++) {
        int32_t l[3], h[3];
        #pragma unroll
        for (int c = 0; c < 3; c++) {
            l[c] = __builtin_amdgcn_readlane(lo[c], G * g);
            h[c] = __builtin_amdgcn_readlane(hi[c], G * g);
        }
        bounded[g] = min(l[0], min(l[1], l[2])) > kKeyNegInf && max(h[0], max(h[1], h[2])) < kKeyPosInf;
        box[g].lo = f3{fkey_inv(l[0]), fkey_inv(l[1]), fkey_inv(l[2])};
        box[g].hi = f3{fkey_inv(h[0]), fkey_inv(h[1]), fkey_inv(h[2])};
        box[g].M = fmaxf(fmaxf(fmaxf(fabsf(box[g].lo.x), fabsf(box[g].hi.x)),
                               fmaxf(fabsf(box[g].lo.y), fabsf(box[g].hi.y))),
                         fmaxf(fabsf(box[g].lo.z), fabsf(box[g].hi.z)));
    }
}

}  // namespace

// Stage 1 (stage1_drawSmallTriangles, triangles_visbuffer.cu:274-384): every
// lane draws its K triangles into the single view, once per instance of the
// item.
// The variant that draws meshes over many instances with one triangle per
// lane (the per-draw schedule for such draws) is held to 72 VGPRs, seven
// waves per SIMD: its instance-list steps hide latency with waves, and the
// few spilled registers cost less (1.56 -> 1.44 ms of stage 1 for 3000
// instances of a 1M-triangle mesh); other variants keep the default, where a
// VGPR limit spills far more.
template <int BLOCK, int K, bool PU16, bool BOXES>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(BOXES && K == 1 ? 7 : 1))) void k_stage1(RasterParams P) {
    // the view's record in device memory (a frame graph rewrites it between
    // launches; the kernel arguments stay as captured)
    const ViewParams V0 = kconst(P.views)[0];
    const auto transforms = kconst(V0.transforms);
    uint32_t* words = nullptr;
    if constexpr (PU16) {
        __shared__ uint32_t pu16Words[BLOCK / kWave][kPu16Words];
        words = pu16Words[threadIdx.x / kWave];
    }
    uint32_t* queue = nullptr;
    if constexpr (BOXES) {
        __shared__ uint32_t instQueue[BLOCK / kWave][2 * kWave];
        queue = instQueue[threadIdx.x / kWave];
    }
    s1_sweep<BLOCK, K, PU16>(P, words, [&](const S1Item& it, uint32_t base, const f3 (&A)[K], const f3 (&B)[K],
                                           const f3 (&C)[K]) __attribute__((always_inline)) {
        if (BOXES && it.instCount >= kBoxInstances) {
            // A mesh drawn over many instances.  Per wave group: 64 instances
            // at a time, one per lane, are tested against the box of the
            // wave's 64 triangles (wave_tiny_culled), and the instances it
            // keeps are queued in LDS.  Each 64 queued instances are tested
            // against the boxes of the wave's lane groups of kBoxGroup
            // triangles, which leaves each group its own list of instances
            // to draw.  Every step then draws, in each group, the next
            // instance of that group's list, with the transform read from the
            // lane that tested it: the wave takes as many steps as its
            // longest list holds instances.
            constexpr int NG = kWave / kBoxGroup;
            const uint32_t lane = threadIdx.x & (kWave - 1);
            const uint32_t grp = lane / kBoxGroup;
            const uint32_t cullOrder = item_cull_order(P, it.instCount);
            const auto xf = kglobal(V0.transforms) + it.instOffset;
            #pragma unroll
            for (int k = 0; k < K; k++) {
                const uint32_t tri = base + k * BLOCK + threadIdx.x;
                if (__ballot(tri < it.numTriangles) == 0ull) continue;
                WaveBox wbox, gbox[NG];
                bool gbounded[NG];
                const bool wbounded = wave_box(A[k], B[k], C[k], wbox);
                group_boxes<kBoxGroup>(A[k], B[k], C[k], gbox, gbounded);
                uint32_t queued = 0;
                for (uint32_t jb = 0;; jb += kWave) {
                    const bool more = jb < it.instCount;
                    if (more) {
                        const uint32_t j = jb + lane;
                        const auto Tj = xf + min(j, it.instCount - 1u);
                        float mm[16];
                        #pragma unroll
                        for (int e = 0; e < 16; e++) mm[e] = Tj->m[e];
                        const bool keep = j < it.instCount && !(wbounded && wave_tiny_culled(mm, V0, wbox));
                        const uint64_t kept = __ballot(keep);
                        const uint32_t slot = __builtin_amdgcn_mbcnt_hi((uint32_t)(kept >> 32),
                                                  __builtin_amdgcn_mbcnt_lo((uint32_t)kept, 0u));
                        if (keep) queue[queued + slot] = j;   // queued < 64: slot < 128
                        queued += (uint32_t)__popcll(kept);
                    }
                    if (queued >= kWave || (!more && queued > 0u)) {
                        wave_sync();
                        const uint32_t n = min(queued, (uint32_t)kWave);
                        const uint32_t qj = queue[min(lane, n - 1u)];
                        const uint32_t rest = queue[kWave + lane];
                        wave_sync();
                        if (kWave + lane < queued) queue[lane] = rest;
                        queued -= n;
                        const auto Tq = xf + qj;
                        float mq[16];
                        #pragma unroll
                        for (int e = 0; e < 16; e++) mq[e] = Tq->m[e];
                        const uint32_t fq = Tq->flip;
                        uint64_t q[NG];
                        #pragma unroll
                        for (int g = 0; g < NG; g++)
                            q[g] = __ballot(lane < n && !(gbounded[g] && wave_tiny_culled(mq, V0, gbox[g])));
                        for (;;) {
                            uint64_t any = 0ull;
                            #pragma unroll
                            for (int g = 0; g < NG; g++) any |= q[g];
                            if (!any) break;
                            uint32_t src = 0;
                            bool act = false;
                            #pragma unroll
                            for (int g = 0; g < NG; g++) {
                                if (q[g]) {
                                    const uint32_t sg = (uint32_t)__builtin_ctzll(q[g]);
                                    q[g] &= q[g] - 1ull;
                                    if (grp == (uint32_t)g) {
                                        src = sg;
                                        act = true;
                                    }
                                }
                            }
                            float m[16];
                            #pragma unroll
                            for (int e = 0; e < 16; e++) m[e] = __shfl(mq[e], (int)src);
                            const uint32_t flip = (uint32_t)__shfl((int)fq, (int)src);
                            const uint32_t jj = (uint32_t)__shfl((int)qj, (int)src);
                            if (act && tri < it.numTriangles) {
                                const uint32_t inst = it.instOffset + jj;
                                const uint64_t id = it.cumm + (uint64_t)tri + (uint64_t)(jj * it.numTriangles);
                                stage1_triangle(P, V0, m, flip != 0u, A[k], B[k], C[k], id,
                                                ((uint64_t)inst << 32) | (uint64_t)tri, cullOrder);
                            }
                        }
                    }
                    if (!more) break;
                }
            }
            return;
        }
        for (uint32_t j = 0; j < it.instCount; j++) {
            const uint32_t inst = it.instOffset + j;
            const auto T = transforms + inst;
            float m[16];
            #pragma unroll
            for (int e = 0; e < 16; e++) m[e] = T->m[e];
            const bool flip = T->flip != 0u;
            #pragma unroll
            for (int k = 0; k < K; k++) {
                const uint32_t tri = base + k * BLOCK + threadIdx.x;
                if (tri < it.numTriangles) {
                    const uint64_t id = it.cumm + (uint64_t)tri + (uint64_t)(j * it.numTriangles);
                    stage1_triangle(P, V0, m, flip, A[k], B[k], C[k], id, ((uint64_t)inst << 32) | (uint64_t)tri,
                                    item_cull_order(P, it.instCount));
                }
            }
        }
    });
}

// ---------------------------------------------------------------------------
// Stage 1 over pairs of consecutive triangles (k_stage1_pairs).
//
// Lane t of a wave takes triangles base + 2t and base + 2t + 1.  Meshes
// stored in order put the two triangles of a quad next to each other, and
// those share two vertices: B = (A.a, A.c, n) for a grid cut along one
// diagonal, B = (A.a, n, A.b) for the other (a lat-long sphere).  When every
// lane of a wave holds one of these two patterns (the same one), the wave
// gathers 4 vertices instead of 6 and computes each vertex's depth and view
// y row once for both triangles (stage1_core<true>: the same operations, so
// the same bits); other waves draw both triangles separately.  The six
// indices are read with one 16- and one 8-byte buffer load.  Each wave
// claims units of unitRows consecutive rows of a 128-triangle strip (see
// DrawItem::rowTris) and draws them row after row, so for a grid the vertex
// row two neighbouring rows share is still in its XCD's L2 when the second
// row reads it; claims are dynamic, which keeps the view-dependent cost of
// the units balanced.  Only the order and the sharing of identical values
// differ from k_stage1.
// ---------------------------------------------------------------------------

constexpr uint32_t kPairNone = 0, kPairGrid = 1, kPairFan = 2;

// The layout of the item a unit walk is in (wave-uniform).
struct RunLayout {
    uint32_t rowTris, unitRows, unitSegs, unitCols, rows, firstUnit;
};

// A batch of the unit walk: unit (absolute), its strip, current row, the
// batch's column in the strip and the unit's row end (wave-uniform scalars).
struct RunBatch {
    uint32_t unit, strip, row, col, rowEnd;
};

__device__ __forceinline__ uint32_t uni(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }

// Triangles [base, end) of batch b (one wave's 128 triangles of a row; a
// column past the row's end, in the last strip, is empty).
__device__ __forceinline__ void run_bounds(const S1Item& it, const RunLayout& L, const RunBatch& b, uint32_t& base,
                                           uint32_t& end) {
    constexpr uint32_t waveTris = 2u * kWave;
    const uint64_t x = ((uint64_t)b.strip * L.unitCols + b.col) * waveTris;
    const uint64_t start = (uint64_t)b.row * L.rowTris + x;
    const uint64_t w = x < L.rowTris ? min((uint64_t)waveTris, (uint64_t)L.rowTris - x) : 0ull;
    base = uni((uint32_t)min(start, (uint64_t)it.numTriangles));
    end = uni((uint32_t)min(start + w, (uint64_t)it.numTriangles));
}

// First batch of unit u.
__device__ __forceinline__ RunBatch run_first(const RunLayout& L, uint32_t u) {
    const uint32_t rel = u - L.firstUnit;
    const uint32_t s = uni(rel / L.unitSegs);
    const uint32_t g = rel - s * L.unitSegs;
    RunBatch b;
    b.unit = uni(u);
    b.strip = s;
    b.col = 0u;
    b.row = uni(g * L.unitRows);
    b.rowEnd = uni(min(L.rows, g * L.unitRows + L.unitRows));
    return b;
}

// Claims this wave's next unit: units x, x + 8, x + 16, ... belong to the
// waves of XCD x (workgroup b runs on XCD b % 8), counted on that XCD's own
// counter, so the claims of different XCDs never meet on one address, and a
// unit is drawn row after row by one wave (the vertex row two rows share is
// read again from the XCD's L2).  The reference claims batches from one
// atomic cursor (triangles_visbuffer.cu:313).
__device__ __forceinline__ uint32_t run_claim(uint32_t* claims) {
    const uint32_t x = blockIdx.x & 7u;
    uint32_t k = 0u;
    if ((threadIdx.x & (kWave - 1)) == 0) k = atomicAdd(claims + kClaimStride * x, 1u);
    k = __builtin_amdgcn_readfirstlane(k);
    return x + 8u * k;
}

// The batch after `a` in this wave's walk of the item: the next column of
// the row (while it starts inside the row), else the unit's next row, else
// the first row of a newly claimed unit of the same item; more = 0 when the
// claimed unit lies past the item (left in `pending` for the next pass; the
// batch returned is then `a` again, a valid address to reload).
__device__ __forceinline__ RunBatch run_next(const RunLayout& L, uint32_t* claims, uint32_t iend, const RunBatch& a,
                                             uint32_t& more, uint32_t& pending) {
    more = 1u;
    if (a.col + 1u < L.unitCols && ((uint64_t)a.strip * L.unitCols + a.col + 1u) * (2u * kWave) < L.rowTris) {
        RunBatch b = a;
        b.col = uni(a.col + 1u);
        return b;
    }
    if (a.row + 1u < a.rowEnd) {
        RunBatch b = a;
        b.row = uni(a.row + 1u);
        b.col = 0u;
        return b;
    }
    const uint32_t u = run_claim(claims);
    if (u < iend) return run_first(L, u);
    pending = u;
    more = 0u;
    return a;
}

// Indices of the lane's two triangles (lanes past `end` read zeros).  A
// non-indexed item reads the same words of its positions (at least 36
// bytes per triangle) and replaces them with the vertex numbers 3t + v.
// IX_PU16: the raw stream words, cut into indices by pair_pu16_decode.  The
// wave's triangles [base, end) hold bits [3*bits*base, 3*bits*end) of the
// stream, at most 384 words from the word holding the first bit; lane l
// loads words 6l .. 6l + 5 of that range (zeros past the range or the
// stream), with the same two buffer loads.
constexpr uint32_t kPairPu16Words = 392;   // 384 words, the word after them and padding

__device__ __forceinline__ void pair_pu16_range(const S1Item& it, uint32_t base, uint32_t end, uint64_t& firstWord,
                                                uint32_t& bit0, uint32_t& avail) {
    const uint64_t b = (uint64_t)it.bits * 3ull * base;
    firstWord = b >> 5;
    bit0 = (uint32_t)b & 31u;
    const uint64_t numWords = ((uint64_t)it.numTriangles * 3ull * it.bits + 31u) >> 5;
    const uint64_t need = ((uint64_t)bit0 + 3ull * it.bits * (end - base) + 31u) >> 5;
    avail = (uint32_t)min(need, numWords - firstWord);
}

template <bool PU16>
__device__ __forceinline__ void pair_indices(const S1Item& it, uint32_t base, uint32_t end, uint32_t lane,
                                             uint32_t (&I)[6]) {
#if defined(__HIP_DEVICE_COMPILE__)
    __amdgpu_buffer_rsrc_t rs;
    if constexpr (PU16) {
        uint64_t firstWord;
        uint32_t bit0, avail;
        pair_pu16_range(it, base, end, firstWord, bit0, avail);
        if (it.bits == 0u || it.bits > 32u) avail = 0u;     // BitEdit::readU32 reads nothing
        rs = buffer_rsrc(it.indices + firstWord, 4u * avail);
    } else {
        rs = buffer_rsrc(it.indices + 3ull * base, 12u * (end - base));
    }
    const uint4 v = buffer_load_u4(rs, 24u * lane);
    const uint2 w = buffer_load_u2(rs, 24u * lane + 16u);
    I[0] = v.x; I[1] = v.y; I[2] = v.z; I[3] = v.w; I[4] = w.x; I[5] = w.y;
    if (!PU16 && it.indexMask == 0u) {          // wave-uniform; no loads in this branch
        const uint32_t t = min(base + 2u * lane, it.numTriangles - 1u), t1 = min(t + 1u, it.numTriangles - 1u);
        I[0] = 3u * t; I[1] = 3u * t + 1u; I[2] = 3u * t + 2u;
        I[3] = 3u * t1; I[4] = 3u * t1 + 1u; I[5] = 3u * t1 + 2u;
    }
#else
    (void)it; (void)base; (void)end; (void)lane;
    for (int j = 0; j < 6; j++) I[j] = 0u;
#endif
}

// IX_PU16: the six indices of the lane's two triangles from the stream words
// of pair_indices, through the wave's LDS words (BitEdit::readU32's
// arithmetic, BitEdit.h:8-26, plus index_min).  A lane past the batch's end
// cuts the batch's last triangle, so every read stays inside the words
// loaded.
__device__ __forceinline__ void pair_pu16_decode(const S1Item& it, uint32_t base, uint32_t end, uint32_t lane,
                                                 uint32_t* words, uint32_t (&I)[6]) {
    const uint32_t bits = it.bits;
    if (bits == 0u || bits > 32u) {                     // wave-uniform
        #pragma unroll
        for (int j = 0; j < 6; j++) I[j] = it.indexMin;
        return;
    }
    uint64_t firstWord;
    uint32_t bit0, avail;
    pair_pu16_range(it, base, end, firstWord, bit0, avail);
    uint32_t* row = words + 6u * lane;
    *reinterpret_cast<uint2*>(row) = make_uint2(I[0], I[1]);
    *reinterpret_cast<uint2*>(row + 2) = make_uint2(I[2], I[3]);
    *reinterpret_cast<uint2*>(row + 4) = make_uint2(I[4], I[5]);
    wave_sync();
    const uint32_t last = end > base ? end - base - 1u : 0u;
    const uint32_t mask = bits >= 32u ? 0xFFFFFFFFu : ((1u << bits) - 1u);
    uint32_t v[6];
    #pragma unroll
    for (int h = 0; h < 2; h++) {
        // bit offset of the triangle's first index from the range's first word
        uint32_t off = bit0 + 3u * bits * min(2u * lane + (uint32_t)h, last);
        #pragma unroll
        for (int j = 0; j < 3; j++) {
            const uint32_t wd = off >> 5;
            v[3 * h + j] = __builtin_amdgcn_alignbit(words[wd + 1u], words[wd], off & 31u) & mask;
            off += bits;
        }
    }
    wave_sync();
    #pragma unroll
    for (int j = 0; j < 6; j++) I[j] = v[j] + it.indexMin;
}

// The sharing pattern of the wave's pairs, and the gathers: A's three
// vertices, then B's new vertex (a pattern) or B's three (none).  A lane
// whose second triangle (or both) lies past the batch's end takes indices
// that fit both patterns, so a partly filled batch keeps the pattern of its
// drawn pairs (those triangles are not drawn).  IX_PU16: the indices are
// cut out of the stream words first, and the gathers are raw windows until
// pu16_vertex_decode.
template <bool PU16>
__device__ __forceinline__ uint32_t pair_vertices(const S1Item& it, uint32_t base, uint32_t end, uint32_t lane,
                                                  uint32_t* words, uint32_t (&I)[6], f3 (&S)[6]) {
    if constexpr (PU16) pair_pu16_decode(it, base, end, lane, words, I);
    const uint32_t rel = 2u * lane;
    if (base + rel >= end) {
        I[1] = I[2] = I[0];
    }
    if (base + rel + 1u >= end) {
        I[3] = I[0];
        I[4] = I[2];
        I[5] = I[1];
    }
    const bool grid = I[3] == I[0] && I[4] == I[2];
    const bool fan = I[3] == I[0] && I[5] == I[1];
    const uint32_t pat = __ballot(!grid) == 0ull ? kPairGrid : __ballot(!fan) == 0ull ? kPairFan : kPairNone;
    const uint32_t i3 = pat == kPairGrid ? I[5] : pat == kPairFan ? I[4] : I[3];
    if constexpr (PU16) {
        S[0] = pu16_vertex_issue(it, I[0]);
        S[1] = pu16_vertex_issue(it, I[1]);
        S[2] = pu16_vertex_issue(it, I[2]);
        S[3] = pu16_vertex_issue(it, i3);
        if (pat == kPairNone) {
            S[4] = pu16_vertex_issue(it, I[4]);
            S[5] = pu16_vertex_issue(it, I[5]);
        }
    } else {
        S[0] = load_pos_g(it.positions, I[0]);
        S[1] = load_pos_g(it.positions, I[1]);
        S[2] = load_pos_g(it.positions, I[2]);
        S[3] = load_pos_g(it.positions, i3);
        if (pat == kPairNone) {
            S[4] = load_pos_g(it.positions, I[4]);
            S[5] = load_pos_g(it.positions, I[5]);
        }
    }
    return pat;
}

// The batch walk of k_stage1_pairs, software-pipelined two deep like
// s1_sweep: while batch i is drawn, the gathers of batch i+1 and the index
// loads of batch i+2 are in flight.  Units (strip, group of rows) are dealt
// to the workgroups by P.order as s1_range deals batches; a workgroup walks
// its unit's rows in turn.  body(it, base, end, S, pattern) draws one batch.
template <bool PU16, typename Body>
__device__ __forceinline__ void s1_pairs_sweep(const RasterParams& P, uint32_t* words, Body&& body) {
    const uint32_t numUnits = P.batchCount ? *kconst(P.batchCount) : P.numBatches;
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const auto items = kconst(P.items);
    uint32_t unit = run_claim(P.claims);
    while (unit < numUnits) {
        const uint32_t item = find_item(P, unit);
        const S1Item it = s1_item(P, item, numUnits);
        RunLayout L;
        L.rowTris = items[item].rowTris;
        L.unitRows = items[item].unitRows;
        L.unitSegs = items[item].unitSegs;
        L.unitCols = items[item].unitCols;
        L.rows = (uint32_t)(((uint64_t)it.numTriangles + L.rowTris - 1u) / L.rowTris);
        L.firstUnit = it.firstBatch;
        const uint32_t iend = it.endBatch;
        uint32_t pending = numUnits;
        uint32_t base0, end0, base1, end1, base2, end2;
        const RunBatch b0 = run_first(L, unit);
        run_bounds(it, L, b0, base0, end0);
        uint32_t I[6], J[6];
        pair_indices<PU16>(it, base0, end0, lane, I);
        uint32_t more1;
        RunBatch b1 = run_next(L, P.claims, iend, b0, more1, pending);
        run_bounds(it, L, b1, base1, end1);
        pair_indices<PU16>(it, base1, end1, lane, J);
        f3 S[6];
        uint32_t pat = pair_vertices<PU16>(it, base0, end0, lane, words, I, S);
        for (;;) {
            f3 Sn[6];
            const uint32_t patN = pair_vertices<PU16>(it, base1, end1, lane, words, J, Sn);
            uint32_t more2 = 0u;
            const RunBatch b2 = more1 ? run_next(L, P.claims, iend, b1, more2, pending) : b1;
            run_bounds(it, L, b2, base2, end2);
            pair_indices<PU16>(it, base2, end2, lane, J);
            if constexpr (PU16) {
                #pragma unroll
                for (int j = 0; j < 4; j++) S[j] = pu16_vertex_decode(it, S[j]);
                if (pat == kPairNone) {
                    S[4] = pu16_vertex_decode(it, S[4]);
                    S[5] = pu16_vertex_decode(it, S[5]);
                }
            }
            body(it, base0, end0, S, pat);
            if (!more1) break;
            base0 = base1;
            end0 = end1;
            b1 = b2;
            base1 = base2;
            end1 = end2;
            more1 = more2;
            #pragma unroll
            for (int j = 0; j < 6; j++) S[j] = Sn[j];
            pat = patN;
        }
        // a unit claimed past this item (or past the end) starts the next pass
        unit = pending;
    }
}

// Stage 1 over pairs of consecutive triangles for draws of one view whose
// items are drawn over few instances (the others take k_stage1).
template <int BLOCK, bool PU16>
__global__ __launch_bounds__(BLOCK) void k_stage1_pairs(RasterParams P) {
    const ViewParams V0 = kconst(P.views)[0];
    const auto transforms = kconst(V0.transforms);
    uint32_t* words = nullptr;
    if constexpr (PU16) {
        __shared__ __align__(16) uint32_t pairWords[BLOCK / kWave][kPairPu16Words];
        words = pairWords[threadIdx.x / kWave];
    }
    s1_pairs_sweep<PU16>(P, words, [&](const S1Item& it, uint32_t base, uint32_t end, const f3 (&S)[6], uint32_t pat)
                          __attribute__((always_inline)) {
        const uint32_t ta = base + 2u * (threadIdx.x & (kWave - 1)), tb = ta + 1u;
        for (uint32_t j = 0; j < it.instCount; j++) {
            const uint32_t inst = it.instOffset + j;
            const auto T = transforms + inst;
            float m[16];
            #pragma unroll
            for (int e = 0; e < 16; e++) m[e] = T->m[e];
            const bool flip = T->flip != 0u;
            const uint32_t cullOrder = item_cull_order(P, it.instCount);
            const uint64_t off = it.cumm + (uint64_t)(j * it.numTriangles);
            const uint64_t hi = (uint64_t)inst << 32;
            if (pat != kPairNone) {
                // the four distinct vertices' quantities, once
                const VQ q0 = vertex_q(m, S[0]), q1 = vertex_q(m, S[1]), q2 = vertex_q(m, S[2]), q3 = vertex_q(m, S[3]);
                if (ta < end) stage1_core<true>(P, V0, m, flip, S[0], S[1], S[2], q0, q1, q2, off + ta, hi | ta, cullOrder);
                // B = (A.a, A.c, n) or (A.a, n, A.b): a wave-uniform choice
                if (pat == kPairGrid) {
                    if (tb < end) stage1_core<true>(P, V0, m, flip, S[0], S[2], S[3], q0, q2, q3, off + tb, hi | tb, cullOrder);
                } else {
                    if (tb < end) stage1_core<true>(P, V0, m, flip, S[0], S[3], S[1], q0, q3, q1, off + tb, hi | tb, cullOrder);
                }
            } else {
                if (ta < end) stage1_triangle(P, V0, m, flip, S[0], S[1], S[2], off + ta, hi | ta, cullOrder);
                if (tb < end) stage1_triangle(P, V0, m, flip, S[3], S[4], S[5], off + tb, hi | tb, cullOrder);
            }
        }
    });
}

// Stage 1 of a multi-view draw.  Each batch of triangles is fetched once and
// drawn into every view that may see it: a wave reduces the object-space box
// of its 64 triangles, each lane tests that box against one view
// (view_may_see: conservative, so it never drops a triangle the view would
// draw), and the wave runs the exact per-triangle code of the single-view
// stage 1 for the views left in the ballot.  Queue entries carry the view in
// the top 8 bits of their instance field.
template <int BLOCK, int K, bool PU16>
__global__ __launch_bounds__(BLOCK) void k_stage1_views(RasterParams P) {
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t numViews = P.numViews;
    const bool cacheCull = numViews <= (uint32_t)kWave;   // lane v keeps view v's planes
    ViewCull cull{};
    uint32_t cullInst = 0xffffffffu;

    uint32_t* words = nullptr;
    if constexpr (PU16) {
        __shared__ uint32_t pu16Words[BLOCK / kWave][kPu16Words];
        words = pu16Words[threadIdx.x / kWave];
    }
    s1_sweep<BLOCK, K, PU16>(P, words, [&](const S1Item& it, uint32_t base, const f3 (&A)[K], const f3 (&B)[K],
                                           const f3 (&C)[K]) __attribute__((always_inline)) {
        #pragma unroll
        for (int k = 0; k < K; k++) {
            const uint32_t tri = base + k * BLOCK + threadIdx.x;
            const bool valid = tri < it.numTriangles;
            if (__ballot(valid) == 0ull) continue;
            // object-space box of the wave's triangles (lanes past the item's
            // end hold its last triangle, which is inside the mesh anyway)
            WaveBox box;
            const bool unbounded = !wave_box(A[k], B[k], C[k], box);

            for (uint32_t j = 0; j < it.instCount; j++) {
                const uint32_t inst = it.instOffset + j;
                const uint64_t id = it.cumm + (uint64_t)tri + (uint64_t)(j * it.numTriangles);
                for (uint32_t vb = 0; vb < numViews; vb += kWave) {
                    const uint32_t v = vb + lane;
                    bool maybe = false;
                    if (v < numViews) {
                        if (!cacheCull || inst != cullInst) cull = lane_view_cull(P, v, inst);
                        maybe = !cull.skip && (unbounded || view_may_see(cull, box));
                    }
                    if (cacheCull) cullInst = inst;
                    uint64_t mask = __ballot(maybe);
                    while (mask) {
                        const uint32_t vi = vb + (uint32_t)__builtin_ctzll(mask);
                        mask &= mask - 1ull;
                        const ViewParams V = kconst(P.views)[vi];
                        const auto T = kconst(V.transforms) + inst;
                        float m[16];
                        #pragma unroll
                        for (int e = 0; e < 16; e++) m[e] = T->m[e];
                        if (valid) {
                            const uint64_t entry = ((uint64_t)((vi << kViewShift) | inst) << 32) | (uint64_t)tri;
                            stage1_triangle(P, V, m, T->flip != 0u, A[k], B[k], C[k], id, entry,
                                            item_cull_order(P, it.instCount));
                        }
                    }
                }
            }
        }
    });
}

namespace {

// SutherlandHodgman::windowClipping, rasterization_helpers.cuh:126-210,
// extending the running box with the clipped polygon's vertices.
__device__ __forceinline__ bool sh_inside(f2 p, int e) {
    // edges: 0 x >= -1, 1 -x >= -1, 2 y >= -1, 3 -y >= -1  (sign*val >= sign*value)
    const float val = (e < 2) ? p.x : p.y;
    const float sign = (e & 1) ? -1.0f : 1.0f;
    const float value = (e & 1) ? 1.0f : -1.0f;
    return sign * val >= sign * value;
}

__device__ __forceinline__ f2 sh_intersect(f2 p1, f2 p2, int e) {
    const float value = (e & 1) ? 1.0f : -1.0f;
    if (e < 2) {
        const float t = (value - p1.x) / (p2.x - p1.x);
        return f2{value, p1.y + t * (p2.y - p1.y)};
    }
    const float t = (value - p1.y) / (p2.y - p1.y);
    return f2{p1.x + t * (p2.x - p1.x), value};
}

__device__ void sh_clip_extend(f2 a, f2 b, f2 c, f2& bmin, f2& bmax) {
    // A triangle inside all four edges passes every edge unchanged (each
    // vertex is kept, no intersection is made): extend by a, b, c in order,
    // as the loop below would, without its per-lane polygon buffers.
    if (sh_inside(a, 0) && sh_inside(a, 1) && sh_inside(a, 2) && sh_inside(a, 3) &&
        sh_inside(b, 0) && sh_inside(b, 1) && sh_inside(b, 2) && sh_inside(b, 3) &&
        sh_inside(c, 0) && sh_inside(c, 1) && sh_inside(c, 2) && sh_inside(c, 3)) {
        for (const f2 p : {a, b, c}) {
            bmin.x = fminf(bmin.x, p.x);
            bmin.y = fminf(bmin.y, p.y);
            bmax.x = fmaxf(bmax.x, p.x);
            bmax.y = fmaxf(bmax.y, p.y);
        }
        return;
    }
    f2 bufA[8], bufB[8];
    f2* in = bufA;
    f2* out = bufB;
    in[0] = a; in[1] = b; in[2] = c;
    int inCount = 3;
    for (int e = 0; e < 4; e++) {
        int outCount = 0;
        if (inCount == 0) break;
        f2 S = in[inCount - 1];
        for (int i = 0; i < inCount; i++) {
            const f2 E = in[i];
            const bool Ein = sh_inside(E, e);
            const bool Sin = sh_inside(S, e);
            if (Ein) {
                if (!Sin) out[outCount++] = sh_intersect(S, E, e);
                out[outCount++] = E;
            } else if (Sin) {
                out[outCount++] = sh_intersect(S, E, e);
            }
            S = E;
        }
        f2* tmp = in; in = out; out = tmp;
        inCount = outCount;
    }
    for (int i = 0; i < inCount; i++) {
        bmin.x = fminf(bmin.x, in[i].x);
        bmin.y = fminf(bmin.y, in[i].y);
        bmax.x = fmaxf(bmax.x, in[i].x);
        bmax.y = fmaxf(bmax.y, in[i].y);
    }
}

__device__ __forceinline__ f2 view_to_ndc_xy(f3 v, float f, float aspect) {
    const float depth = -v.z;
    return f2{(f / aspect) * v.x / depth, f * v.y / depth};
}

// computeScreenSpaceBoundingBox, rasterization_helpers.cuh:350-460 (mask != 7).
__device__ void stage2_bbox(f3 av, f3 bv, f3 cv, float da, float db, float dc,
                            f2 an, f2 bn, f2 cn, float f, float aspect, float W, float H,
                            float& min_x, float& max_x, float& min_y, float& max_y) {
    f2 bmin{1e10f, 1e10f};
    f2 bmax{-1e10f, -1e10f};
    const int mask = (da <= 0.0f ? 1 : 0) | (db <= 0.0f ? 2 : 0) | (dc <= 0.0f ? 4 : 0);
    if (mask == 0) {
        sh_clip_extend(an, bn, cn, bmin, bmax);
    } else {
        f3 v0, v1, v2;
        if (mask == 1 || mask == 6) { v0 = av; v1 = bv; v2 = cv; }
        else if (mask == 2 || mask == 5) { v0 = bv; v1 = cv; v2 = av; }
        else { v0 = cv; v1 = av; v2 = bv; }
        const float near_ = 0.01f;
        const float t1 = (-near_ - v0.z) / (v1.z - v0.z);
        const f3 I1 = add3(v0, scale3(sub3(v1, v0), t1));
        const float t2 = (-near_ - v0.z) / (v2.z - v0.z);
        const f3 I2 = add3(v0, scale3(sub3(v2, v0), t2));
        if (mask == 1 || mask == 2 || mask == 4) {
            const f2 p0 = view_to_ndc_xy(v1, f, aspect);
            const f2 p1 = view_to_ndc_xy(v2, f, aspect);
            const f2 pI1 = view_to_ndc_xy(I1, f, aspect);
            const f2 pI2 = view_to_ndc_xy(I2, f, aspect);
            sh_clip_extend(p0, p1, pI1, bmin, bmax);
            sh_clip_extend(pI1, p1, pI2, bmin, bmax);
        } else {
            sh_clip_extend(view_to_ndc_xy(v0, f, aspect), view_to_ndc_xy(I1, f, aspect),
                           view_to_ndc_xy(I2, f, aspect), bmin, bmax);
        }
    }
    bmin.x = clampf(bmin.x, -1.0f, 1.0f);
    bmax.x = clampf(bmax.x, -1.0f, 1.0f);
    bmin.y = clampf(bmin.y, -1.0f, 1.0f);
    bmax.y = clampf(bmax.y, -1.0f, 1.0f);
    const float halfW = W * 0.5f, halfH = H * 0.5f;
    min_x = bmin.x * halfW + halfW;
    min_y = bmin.y * halfH + halfH;
    max_x = bmax.x * halfW + halfW;
    max_y = bmax.y * halfH + halfH;
}

// Per-entry setup of stage 2 (:418-471), computed by one lane for its entry.
struct S2Setup {
    float ax, ay, vabx, vaby, vacx, vacy, inv_det;
    float iza, izb, izc;
    float start_x, start_y;
    int32_t size_x;
    uint32_t numFragments;
    uint32_t id_lo, id_hi;
    uint64_t* fb;            // the entry's view
    int32_t W, H;
};

// Instance field of a queue entry: the instance index for a plain draw,
// view << 24 | instance for a multi-view draw.
__device__ __forceinline__ uint32_t entry_instance(const RasterParams& P, uint32_t field) {
    return P.numViews == 1u ? field : (field & kInstanceMask);
}

// Index of the view of a queue entry (0 for a plain draw).
__device__ __forceinline__ uint32_t entry_view(const RasterParams& P, uint32_t field) {
    return P.numViews == 1u ? 0u : (field >> kViewShift);
}

// Vertex indices of triangle `tri` of an instance (plain, non-indexed or
// IX_PU16 bit-packed).
__device__ __forceinline__ void triangle_indices(const InstanceRec& mesh, uint32_t tri, uint32_t& i0, uint32_t& i1,
                                                 uint32_t& i2) {
    if (mesh.compressed) {
        const uint32_t bits = mesh.bitsPerIndex;
        const uint64_t off = (uint64_t)bits * (3ull * tri);
        i0 = bit_read(mesh.indices, off, bits) + mesh.indexMin;
        i1 = bit_read(mesh.indices, off + bits, bits) + mesh.indexMin;
        i2 = bit_read(mesh.indices, off + 2u * bits, bits) + mesh.indexMin;
    } else if (mesh.indices) {
        const auto ip = kglobal(mesh.indices) + 3ull * tri;
        i0 = ip[0];
        i1 = ip[1];
        i2 = ip[2];
    } else {
        i0 = 3u * tri; i1 = i0 + 1u; i2 = i0 + 2u;
    }
}

// Object-space vertex as the raster stages decode it (getVertex).
__device__ __forceinline__ f3 raster_vertex(const InstanceRec& mesh, uint32_t i) {
    if (mesh.compressed) return load_pos_pu16(mesh.positions, i, mesh.cf, mesh.amin);
    return load_pos(mesh.positions, i);
}

// Object-space vertex as the resolve decodes it (getVertex_resolved,
// resolve.cu:69-89): (X / 65536) * (aabb.max - aabb.min) + aabb.min.
__device__ __forceinline__ f3 resolve_vertex(const InstanceRec& mesh, uint32_t i) {
    if (!mesh.compressed) return load_pos(mesh.positions, i);
    const auto q = kglobal(reinterpret_cast<const uint16_t*>(mesh.positions)) + 3ull * i;
    return f3{((float)q[0] / 65536.0f) * (mesh.amax[0] - mesh.amin[0]) + mesh.amin[0],
              ((float)q[1] / 65536.0f) * (mesh.amax[1] - mesh.amin[1]) + mesh.amin[1],
              ((float)q[2] / 65536.0f) * (mesh.amax[2] - mesh.amin[2]) + mesh.amin[2]};
}

}  // namespace

// Stage 2.  A persistent grid of waves; each wave claims up to 64 queue
// entries at once, every lane sets up one entry (bounding box with
// Sutherland-Hodgman clipping), huge or near-plane triangles are split into
// tiles appended to the stage-3 queue with one atomic per wave, and the
// remaining entries are rasterised one after the other by all 64 lanes.
__global__ __launch_bounds__(256) void k_stage2(RasterParams P) {
    __shared__ S2Setup sh_setup[4][kWave];
    __shared__ uint32_t sh_tiles[4][kWave][4];   // tile_x, tile_y, tiles_x, numTiles
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t wid = threadIdx.x / kWave;
    S2Setup* setup = sh_setup[wid];
    const uint32_t count = min(*P.ctrNontrivial, P.ntCap);
    const int TILE = P.tile;
    // Entries per claim: 64 (one per lane) once the queue fills the grid's
    // waves; a shorter queue is spread over more waves in smaller claims, as
    // each wave rasterises its claim's triangles one after the other.  The
    // first claim is static (wave w takes entries [w*chunk, (w+1)*chunk)),
    // later claims come from the shared cursor, so waves beyond the queue
    // exit without an atomic.
    const uint32_t chunk =
        (uint32_t)min((uint64_t)kWave, max((uint64_t)1, ((uint64_t)count + P.numWaves2 - 1u) / P.numWaves2));
    const uint32_t waveId = blockIdx.x * 4u + wid;
    bool first = true;

    while (true) {
        wave_sync();
        uint32_t base = 0;
        if (first) {
            base = waveId * chunk;
            first = false;
        } else {
            if (lane == 0) base = P.numWaves2 * chunk + atomicAdd(P.cursor2, chunk);
            base = __shfl(base, 0);
        }
        if (base >= count) break;
        const uint32_t idx = base + lane;
        bool raster = false;
        uint32_t numTiles = 0;
        int tile_x = 0, tile_y = 0, tiles_x = 0;
        uint32_t meshIndex = 0, triangleIndex = 0;
        if (lane < chunk && idx < count) {
            const uint64_t packed = P.nontrivial[idx];
            meshIndex = (uint32_t)(packed >> 32);    // view << 24 | instance, kept for the tile entries
            triangleIndex = (uint32_t)packed;
            const InstanceRec& mesh = P.inst[entry_instance(P, meshIndex)];
            const ViewParams& V = P.views[entry_view(P, meshIndex)];
            const float* m = V.transforms[mesh.transformIndex].m;
            uint32_t i0, i1, i2;
            triangle_indices(mesh, triangleIndex, i0, i1, i2);
            const f3 av = xform_point(m, raster_vertex(mesh, i0));
            const f3 bv = xform_point(m, raster_vertex(mesh, i1));
            const f3 cv = xform_point(m, raster_vertex(mesh, i2));
            const float f = V.f, aspect = V.aspect;
            const float da = -av.z, db = -bv.z, dc = -cv.z;
            const f2 an = view_to_ndc_xy(av, f, aspect);
            const f2 bn = view_to_ndc_xy(bv, f, aspect);
            const f2 cn = view_to_ndc_xy(cv, f, aspect);
            float min_x, max_x, min_y, max_y;
            stage2_bbox(av, bv, cv, da, db, dc, an, bn, cn, f, aspect, V.Wf, V.Hf, min_x, max_x, min_y, max_y);
            min_x -= 0.5f; min_y -= 0.5f;
            max_x += 0.5f; max_y += 0.5f;
            const float asx = (an.x * 0.5f + 0.5f) * V.Wf, asy = (an.y * 0.5f + 0.5f) * V.Hf;
            const float bsx = (bn.x * 0.5f + 0.5f) * V.Wf, bsy = (bn.y * 0.5f + 0.5f) * V.Hf;
            const float csx = (cn.x * 0.5f + 0.5f) * V.Wf, csy = (cn.y * 0.5f + 0.5f) * V.Hf;
            const float vabx = bsx - asx, vaby = bsy - asy;
            const float vacx = csx - asx, vacy = csy - asy;
            const float inv_det = 1.0f / (vabx * vacy - vaby * vacx);
            const int size_x = f2i(ceilf(max_x)) - f2i(floorf(min_x));
            const int size_y = f2i(ceilf(max_y)) - f2i(floorf(min_y));
            const uint32_t numFragments = (uint32_t)size_x * (uint32_t)size_y;
            if (!(min_x >= max_x || min_y >= max_y)) {
                if ((int32_t)numFragments >= TILE * TILE || da <= 0.0f || db <= 0.0f || dc <= 0.0f) {
                    tile_x = f2i(min_x) / TILE;
                    tile_y = f2i(min_y) / TILE;
                    tiles_x = (f2i(max_x) + TILE - 1) / TILE - tile_x;
                    const int tiles_y = (f2i(max_y) + TILE - 1) / TILE - tile_y;
                    numTiles = (uint32_t)tiles_x * (uint32_t)tiles_y;
                } else {
                    raster = true;
                    S2Setup s;
                    s.ax = asx; s.ay = asy;
                    s.vabx = vabx; s.vaby = vaby; s.vacx = vacx; s.vacy = vacy;
                    s.inv_det = inv_det;
                    s.iza = 1.0f / da; s.izb = 1.0f / db; s.izc = 1.0f / dc;
                    s.start_x = floorf(min_x);
                    s.start_y = floorf(min_y);
                    s.size_x = size_x;
                    s.numFragments = numFragments;
                    // packed_id = (uint32)(uint32(cumm) + triangleIndex), :414-519
                    const uint32_t id32 = (uint32_t)mesh.cumm + triangleIndex;
                    s.id_lo = id32;
                    s.id_hi = 0u;
                    s.fb = V.fb;
                    s.W = V.W;
                    s.H = V.H;
                    setup[lane] = s;
                }
            }
        }

        // LDS-staged append of tile entries: one atomic per wave for all lanes' tiles.
        uint32_t prefix = numTiles;
        #pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
            const uint32_t n = __shfl_up(prefix, o);
            if ((int)lane >= o) prefix += n;
        }
        const uint32_t total = __shfl(prefix, kWave - 1);
        if (total > 0) {
            uint32_t hbase = 0;
            if (lane == kWave - 1) hbase = atomicAdd(P.ctrHuge, total);
            hbase = __shfl(hbase, kWave - 1);
            sh_tiles[wid][lane][0] = (uint32_t)tile_x;
            sh_tiles[wid][lane][1] = (uint32_t)tile_y;
            sh_tiles[wid][lane][2] = (uint32_t)tiles_x;
            sh_tiles[wid][lane][3] = prefix - numTiles;
            wave_sync();
            uint64_t pending = __ballot(numTiles > 0);
            while (pending) {
                const int src = __builtin_ctzll(pending);
                pending &= pending - 1;
                const uint32_t n = __shfl(numTiles, src);
                const uint32_t mi = __shfl(meshIndex, src);
                const uint32_t ti = __shfl(triangleIndex, src);
                const uint32_t tx0 = sh_tiles[wid][src][0];
                const uint32_t ty0 = sh_tiles[wid][src][1];
                const uint32_t ntx = sh_tiles[wid][src][2];
                const uint32_t off = hbase + sh_tiles[wid][src][3];
                for (uint32_t i = lane; i < n; i += kWave) {
                    const uint32_t slot = off + i;
                    if (slot < P.hugeCap) {
                        P.huge[slot] = make_int4((int)mi, (int)ti, (int)(tx0 + i % ntx), (int)(ty0 + i / ntx));
                    }
                }
            }
        }

        // Rasterise the medium triangles of this claim, all lanes per triangle (:516-561).
        wave_sync();
        uint64_t todo = __ballot(raster);
        while (todo) {
            const int src = __builtin_ctzll(todo);
            todo &= todo - 1;
            const S2Setup s = setup[src];
            const uint64_t id = (uint64_t)s.id_lo;
            for (uint32_t off = lane; off < s.numFragments; off += kWave) {
                const int fragX = (int)off % s.size_x;
                const int fragY = (int)off / s.size_x;
                const float px = s.start_x + (float)fragX;
                const float py = s.start_y + (float)fragY;
                const float sx = px - s.ax + 0.5f;
                const float sy = py - s.ay + 0.5f;
                const float ss = (sx * s.vacy - sy * s.vacx) * s.inv_det;
                const float tt = (s.vabx * sy - s.vaby * sx) * s.inv_det;
                const float vv = 1.0f - (ss + tt);
                if (ss >= 0.0f && tt >= 0.0f && vv >= 0.0f) {
                    const int ix = f2i(px), iy = f2i(py);
                    // samples outside the screen are discarded (the reference wraps them)
                    if (ix >= 0 && ix < s.W && iy >= 0 && iy < s.H) {
                        const float inv_depth = (vv * s.iza + ss * s.izb) + tt * s.izc;
                        const float depth = 1.0f / inv_depth;
                        atomic_min_fb(s.fb, (uint32_t)(ix + s.W * iy), pack_pixel(depth, id));
                    }
                }
            }
        }
    }
}

namespace {

// intersectTriangle_mt, double-sided branch (rasterization_helpers.cuh:213-270),
// with orig = 0.  e1, e2, qvec and dot(e2, qvec) do not depend on the ray.
struct MTTri {
    f3 e1, e2, tvec, qvec;
};

__device__ __forceinline__ float intersect_mt(const MTTri& T, f3 dir) {
    const f3 pvec = cross3(dir, T.e2);
    const float det = dot3(T.e1, pvec);
    if (det > -1e-7f && det < 1e-7f) return __builtin_inff();
    const float inv_det = 1.0f / det;
    const float u = dot3(T.tvec, pvec) * inv_det;
    if (u < 0.0f || u > 1.0f) return __builtin_inff();
    const float v = dot3(dir, T.qvec) * inv_det;
    if (v < 0.0f || u + v > 1.0f) return __builtin_inff();
    const float t = dot3(T.e2, T.qvec) * inv_det;
    return t > 0.0f ? t : __builtin_inff();
}

}  // namespace

// Stage 3.  A persistent grid of waves; each wave takes one tile entry at a
// time and ray-casts its TILE x TILE pixels (RAYTRACE branch, :681-741).
// With TILE <= 64 the 64 lanes cover 64/TILE rows per step, otherwise each
// lane covers TILE/64 columns.
__global__ __launch_bounds__(256) void k_stage3(RasterParams P) {
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t count = min(*P.ctrHuge, P.hugeCap);
    const int TILE = P.tile;
    const int lanesPerRow = TILE < kWave ? TILE : kWave;
    const int rowsPerStep = kWave / lanesPerRow;
    const int colsPerLane = TILE / lanesPerRow;
    // static first claim (entry = wave id), then the shared cursor
    const uint32_t waveId = blockIdx.x * 4u + (threadIdx.x / kWave);
    bool first = true;
    while (true) {
        uint32_t idx = 0;
        if (first) {
            idx = waveId;
            first = false;
        } else {
            if (lane == 0) idx = P.numWaves3 + atomicAdd(P.cursor3, 1u);
            idx = __shfl(idx, 0);
        }
        if (idx >= count) break;
        const int4 e = P.huge[idx];
        // the entry is the same for all lanes: its view record is read by scalar loads
        const uint32_t field = (uint32_t)__builtin_amdgcn_readfirstlane(e.x);
        const InstanceRec& mesh = P.inst[entry_instance(P, field)];
        const auto& V = kconst(P.views)[entry_view(P, field)];
        const float* m = V.transforms[mesh.transformIndex].m;
        const uint32_t tri = (uint32_t)e.y;
        uint32_t i0, i1, i2;
        triangle_indices(mesh, tri, i0, i1, i2);
        const f3 av = xform_point(m, raster_vertex(mesh, i0));
        const f3 bv = xform_point(m, raster_vertex(mesh, i1));
        const f3 cv = xform_point(m, raster_vertex(mesh, i2));
        MTTri T;
        T.e1 = sub3(bv, av);
        T.e2 = sub3(cv, av);
        T.tvec = f3{-av.x, -av.y, -av.z};
        T.qvec = cross3(T.tvec, T.e1);
        const uint64_t id = mesh.cumm + (uint64_t)(int64_t)e.y;
        const float min_x = clampf((float)(e.z * TILE), 0.0f, V.Wf);
        const float min_y = clampf((float)(e.w * TILE), 0.0f, V.Hf);
        const float fx0 = floorf(min_x), fy0 = floorf(min_y);
        for (int c = 0; c < colsPerLane; c++) {
            const int fragX = (int)(lane % (uint32_t)lanesPerRow) + c * lanesPerRow;
            const int px = f2i(fx0 + (float)fragX);
            if (px >= V.W) continue;   // the reference wraps these into the next row
            const float u = 2.0f * ((float)px + 0.5f) / V.Wf - 1.0f;
            const float rx = V.faI * u;
            for (int y0 = 0; y0 < TILE; y0 += rowsPerStep) {
                const int fragY = y0 + (int)(lane / (uint32_t)lanesPerRow);
                const int py = f2i(fy0 + (float)fragY);
                if (py >= V.H) break;
                const float v = 2.0f * ((float)py + 0.5f) / V.Hf - 1.0f;
                const f3 dir = normalize3(f3{rx, V.fI * v, -1.0f});
                const float t = intersect_mt(T, dir);
                if (t == __builtin_inff()) continue;
                // depth = dot(t*dir, (0,0,-1)) (:734)
                const float depth = ((t * dir.x) * 0.0f + (t * dir.y) * 0.0f) + (t * dir.z) * -1.0f;
                atomic_min_fb(V.fb, (uint32_t)(px + V.W * py), pack_pixel(depth, id));
            }
        }
    }
}

namespace {

// intersectPlane, rasterization_helpers.cuh:318-335
__device__ __forceinline__ float intersect_plane(f3 origin, f3 dir, f3 normal, float constant) {
    const float denominator = dot3(normal, dir);
    if (denominator == 0.0f) return __builtin_inff();
    const float t = -(dot3(origin, normal) + constant) / denominator;
    return (t >= 0.0f) ? t : __builtin_inff();
}

// Colour of a covered pixel whose triangle id `total` lies in `mesh`
// (resolve.cu:736-1035).  `mesh` is a register copy of the instance record.
// computeRayDir (resolve.cu:662-676) through the screen point (px, py).
__device__ __forceinline__ f3 resolve_ray(const ResolveParams& R, const f3& origin, float px, float py) {
    const float u = 2.0f * px / R.Wf - 1.0f;
    const float v = 2.0f * py / R.Hf - 1.0f;
    const f3 rview = normalize3(f3{R.p00I * u, R.p11I * v, -1.0f});
    return normalize3(sub3(xform_point(R.viewI, rview), origin));
}

// sampleColor_linear (resolve.cu:148-195): the four texels around the sample
// (fmodf wraps them into the level, float -> int truncates), weighted by the
// fractions of the shifted texel coordinates.
__device__ __forceinline__ uint32_t sample_linear(const uint32_t* data, int width, int height, float ux, float uy) {
    const float fw = (float)width, fh = (float)height;
    const float ftx = (ux - floorf(ux)) * fw;
    const float fty = (uy - floorf(uy)) * fh;
    const auto tex = kglobal(data);
    const int last = width * height - 1;
    auto texel = [&](float fx, float fy) -> uint32_t {
        const int tx = f2i(fmodf(fx, fw));
        const int ty = f2i(fmodf(fy, fh));
        int id = tx + ty * width;
        id = id < 0 ? 0 : (id > last ? last : id);
        return tex[id];
    };
    const uint32_t t00 = texel(ftx - 0.5f, fty - 0.5f);
    const uint32_t t01 = texel(ftx - 0.5f, fty + 0.5f);
    const uint32_t t10 = texel(ftx + 0.5f, fty - 0.5f);
    const uint32_t t11 = texel(ftx + 0.5f, fty + 0.5f);
    const float wx = fmodf(ftx + 0.5f, 1.0f);
    const float wy = fmodf(fty + 0.5f, 1.0f);
    const float w00 = (1.0f - wx) * (1.0f - wy), w10 = wx * (1.0f - wy), w01 = (1.0f - wx) * wy, w11 = wx * wy;
    uint32_t rgb = 0u;
    #pragma unroll
    for (int k = 0; k < 3; k++) {
        const float c = ((w00 * (float)((t00 >> (8 * k)) & 0xffu) + w10 * (float)((t10 >> (8 * k)) & 0xffu)) +
                         w01 * (float)((t01 >> (8 * k)) & 0xffu)) +
                        w11 * (float)((t11 >> (8 * k)) & 0xffu);
        rgb |= f2u8(c) << (8 * k);
    }
    return 0xff000000u | rgb;
}

// Mip level of a texture sample: trunc(0.5 * log2(m)) clamped to [0, 7]
// (resolve.cu:849-856), taken exactly from m's exponent: for m >= 1 the
// value lies in [e/2, (e+1)/2) with e the exponent of m, so the level is
// e >> 1; m < 1, zero and NaN give level 0.
__device__ __forceinline__ int mip_level(float m) {
    if (!(m >= 1.0f)) return 0;
    const int e = (int)((__builtin_bit_cast(uint32_t, m) >> 23) & 0xffu) - 127;
    return min(e >> 1, 7);
}

// TEX: the variant launched for the TEXTURE and STAGE attributes; the other
// attributes run without their code, whose registers would lower every
// resolve's occupancy.
template <bool TEX>
__device__ __forceinline__ uint32_t shade_pixel(const ResolveParams& R, const InstanceRec& mesh, uint64_t total,
                                                int x, int y) {
    uint32_t color = 0;
    uint32_t tri = (uint32_t)(total - mesh.cumm);
    tri = min(tri, mesh.numTriangles - 1u);
    // an empty mesh has no triangle to fetch: shade it like an unloaded one
    if (!mesh.isLoaded || mesh.numTriangles == 0) {
        color = 0xffff00ffu;
    } else {
        uint32_t i0, i1, i2;
        triangle_indices(mesh, tri, i0, i1, i2);   // resolve.cu:744-760
        // ray through the pixel centre (computeRayDir, resolve.cu:662-680)
        const f3 origin{R.origin[0], R.origin[1], R.origin[2]};
        const f3 rayDir = resolve_ray(R, origin, (float)x + 0.5f, (float)y + 0.5f);

        const f3 aw = xform_point(mesh.world, resolve_vertex(mesh, i0));
        const f3 bw = xform_point(mesh.world, resolve_vertex(mesh, i1));
        const f3 cw = xform_point(mesh.world, resolve_vertex(mesh, i2));
        const f3 v0 = sub3(bw, aw);
        const f3 v1 = sub3(cw, aw);
        // resolve.cu:790-801
        const f3 normal = normalize3(cross3(v0, v1));
        const float d = dot3(aw, normal);
        const float t = intersect_plane(origin, rayDir, normal, -d);
        // resolve.cu:805-834
        const float d00 = dot3(v0, v0);
        const float d01 = dot3(v0, v1);
        const float d11 = dot3(v1, v1);
        const float denom = d00 * d11 - d01 * d01;
        const float denomI = 1.0f / denom;
        // barycentric (s, t) of the ray's hit on the triangle's plane
        auto stv = [&](float th, const f3& rd, float& s_, float& t_) {
            const f3 p = add3(origin, f3{th * rd.x, th * rd.y, th * rd.z});
            const f3 v2 = sub3(p, aw);
            const float d20 = dot3(v2, v0);
            const float d21 = dot3(v2, v1);
            s_ = (d11 * d20 - d01 * d21) * denomI;
            t_ = (d00 * d21 - d01 * d20) * denomI;
        };
        float stx, sty;
        stv(t, rayDir, stx, sty);
        const float w = 1.0f - stx - sty;
        color = 0xffff00ffu;
        const int attr = R.attribute;
        if (attr == 0) {                       // NONE
            color = 0xffffffffu;
        } else if (attr == 5) {                // TRIANGLE_ID
            color = tri * 12345678u;
        } else if (attr == 6) {                // MESH_ID
            color = ((uint32_t)mesh.id + 1u) * 12345678u;
        } else if (TEX && attr == 7) {                 // STAGE (resolve.cu:974-1031)
            // the stage the triangle's size and near-plane test would route it
            // to: view space, viewToNDC, ndcToScreen (no half-pixel shift
            // here, rasterization_helpers.cuh:95-108), the box clipped to the
            // screen, its fragment count (an int product compared unsigned)
            const f3 av = xform_point(R.view, aw), bv = xform_point(R.view, bw), cv = xform_point(R.view, cw);
            const float fa = R.f / R.aspect;
            const float da = -av.z, db = -bv.z, dc = -cv.z;
            const float sxa = ((fa * av.x) / da * 0.5f + 0.5f) * R.Wf, sya = ((R.f * av.y) / da * 0.5f + 0.5f) * R.Hf;
            const float sxb = ((fa * bv.x) / db * 0.5f + 0.5f) * R.Wf, syb = ((R.f * bv.y) / db * 0.5f + 0.5f) * R.Hf;
            const float sxc = ((fa * cv.x) / dc * 0.5f + 0.5f) * R.Wf, syc = ((R.f * cv.y) / dc * 0.5f + 0.5f) * R.Hf;
            float min_x = fminf(sxa, fminf(sxb, sxc)), max_x = fmaxf(sxa, fmaxf(sxb, sxc));
            float min_y = fminf(sya, fminf(syb, syc)), max_y = fmaxf(sya, fmaxf(syb, syc));
            min_x = fmaxf(min_x, 0.0f);
            max_x = fminf(max_x, (float)(R.W - 1));
            min_y = fmaxf(min_y, 0.0f);
            max_y = fminf(max_y, (float)(R.H - 1));
            const int32_t size_x = f2i(ceilf(max_x) - floorf(min_x));
            const int32_t size_y = f2i(ceilf(max_y) - floorf(min_y));
            const uint32_t frags = (uint32_t)size_x * (uint32_t)size_y;
            uint32_t stage = frags <= 128u ? 0u : frags <= 4096u ? 1u : 2u;    // THRESHOLD_SMALL, THRESHOLD_LARGE
            if (da <= 0.0f || db <= 0.0f || dc <= 0.0f) stage = 2u;
            // SCHEME_SPECTRAL[9], [7], [3] (utils.cuh:21-33)
            color = stage == 0u ? 0xffbd8832u : stage == 1u ? 0xffa4ddabu : 0xff61aefdu;
        } else if (TEX && attr == 1 && mesh.uvs) {    // TEXTURE (resolve.cu:783-876)
            // the neighbouring pixels' rays, hits and texture coordinates give
            // the sample's footprint, and with it the mip level
            const f3 r10 = resolve_ray(R, origin, (float)x + 1.5f, (float)y + 0.5f);
            const f3 r01 = resolve_ray(R, origin, (float)x + 0.5f, (float)y + 1.5f);
            const f3 r11 = resolve_ray(R, origin, (float)x + 1.5f, (float)y + 1.5f);
            const float t10 = intersect_plane(origin, r10, normal, -d);
            const float t01 = intersect_plane(origin, r01, normal, -d);
            const float t11 = intersect_plane(origin, r11, normal, -d);
            float s10, u10, s01, u01, s11, u11;
            stv(t10, r10, s10, u10);
            stv(t01, r01, s01, u01);
            stv(t11, r11, s11, u11);
            const float* ua = mesh.uvs + 2ull * i0;
            const float* ub = mesh.uvs + 2ull * i1;
            const float* uc = mesh.uvs + 2ull * i2;
            const float a0 = ua[0], a1 = ua[1], b0 = ub[0], b1 = ub[1], c0 = uc[0], c1 = uc[1];
            // uv_a * (1 - s - t) + uv_b * s + uv_c * t per component
            auto interp = [&](float s_, float t_, float& px, float& py) {
                const float wq = 1.0f - s_ - t_;
                px = (a0 * wq + b0 * s_) + c0 * t_;
                py = (a1 * wq + b1 * s_) + c1 * t_;
            };
            float uvx, uvy, x10, y10, x01, y01, x11, y11;
            interp(stx, sty, uvx, uvy);
            interp(s10, u10, x10, y10);
            interp(s01, u01, x01, y01);
            interp(s11, u11, x11, y11);
            const float maxx = fmaxf(fmaxf(uvx, x10), fmaxf(x01, x11));
            const float maxy = fmaxf(fmaxf(uvy, y10), fmaxf(y01, y11));
            const float minx = fminf(fminf(uvx, x10), fminf(x01, x11));
            const float miny = fminf(fminf(uvy, y10), fminf(y01, y11));
            uvx = uvx - floorf(uvx);
            uvy = uvy - floorf(uvy);
            const float tw = (float)mesh.texW, th = (float)mesh.texH;
            const float dxx = (maxx - minx) * tw, dxy = (maxx - minx) * th;
            const float dyx = (maxy - miny) * tw, dyy = (maxy - miny) * th;
            const int level = mip_level(fmaxf(dxx * dxx + dxy * dxy, dyx * dyx + dyy * dyy));
            if (!mesh.texJpeg && mesh.texData == nullptr) {
                color = 0u;   // no texture: every level samples 0
            } else if (!mesh.texJpeg) {
                // level k of the chain: k levels of (w + 1) / 2 x (h + 1) / 2 in
                const uint32_t* data = mesh.texData;
                int w = mesh.texW, h = mesh.texH;
                for (int i = 1; i <= level; i++) {
                    data = data + (ptrdiff_t)w * h;
                    w = (w + 1) / 2;
                    h = (h + 1) / 2;
                }
                color = sample_linear(data, w, h, uvx, uvy);
            }
            // a JPEG-coded texture: the reference queues its blocks for the
            // decoder and this pass leaves the pixel magenta (resolve.cu:858-876)
        } else if (attr == 2 && mesh.uvs) {    // UV
            const float* ua = mesh.uvs + 2ull * i0;
            const float* ub = mesh.uvs + 2ull * i1;
            const float* uc = mesh.uvs + 2ull * i2;
            float ux = (ua[0] * w + ub[0] * stx) + uc[0] * sty;
            float uy = (ua[1] * w + ub[1] * stx) + uc[1] * sty;
            ux = ux - floorf(ux);
            uy = uy - floorf(uy);
            const uint32_t r = f2u8(fmodf(ux - floorf(ux), 1.0f) * 256.0f);
            const uint32_t g = f2u8(fmodf(uy - floorf(uy), 1.0f) * 256.0f);
            color = (color & 0xff000000u) | r | (g << 8);
        } else if (attr == 3) {                // NORMAL
            f3 N;
            if (mesh.normals) {
                const f3 n0 = xform(mesh.world, load_pos(mesh.normals, i0), 0.0f);
                const f3 n1 = xform(mesh.world, load_pos(mesh.normals, i1), 0.0f);
                const f3 n2 = xform(mesh.world, load_pos(mesh.normals, i2), 0.0f);
                N = normalize3(add3(add3(scale3(n0, w), scale3(n1, stx)), scale3(n2, sty)));
            } else {
                N = normalize3(cross3(v0, v1));
            }
            const uint32_t r = f2u8(clampf(N.x * 255.0f, 0.0f, 255.0f));
            const uint32_t g = f2u8(clampf(N.y * 255.0f, 0.0f, 255.0f));
            const uint32_t b = f2u8(clampf(N.z * 255.0f, 0.0f, 255.0f));
            color = (color & 0xff000000u) | r | (g << 8) | (b << 16);
        } else if (attr == 4 && mesh.colors) { // VERTEX_COLORS (resolve.cu:952-973)
            const auto colors = kglobal(mesh.colors);
        const uint32_t Ca = colors[i0], Cb = colors[i1], Cc = colors[i2];
            uint32_t rgb = 0;
            #pragma unroll
            for (int k = 0; k < 3; k++) {
                const float ca = (float)((Ca >> (8 * k)) & 0xffu);
                const float cb = (float)((Cb >> (8 * k)) & 0xffu);
                const float cc = (float)((Cc >> (8 * k)) & 0xffu);
                const float ch = (ca * w + cb * stx) + cc * sty;
                rgb |= f2u8(clampf(ch, 0.0f, 255.0f)) << (8 * k);
            }
            color = (Ca & 0xff000000u) | rgb;
        } else {
            color = 0xff0000ffu;
        }
        color |= 0xff000000u;
    }
    return color;
}

// Resolve of one pixel (x, y) < (W, H).
template <bool TEX>
__device__ __forceinline__ void resolve_pixel(const ResolveParams& R, int x, int y) {
    const uint32_t pixelID = (uint32_t)x + (uint32_t)R.W * (uint32_t)y;

    const uint64_t pixel = kglobal(R.fb)[pixelID];
    const uint64_t pixel_cb = kglobal((const uint64_t*)R.cb)[pixelID];
    const float depth = unpack_depth(pixel);
    const uint64_t total = pixel & 0xFFFFFFFFFull;
    const float depth_cb = __builtin_bit_cast(float, (uint32_t)(pixel_cb >> 32));
    if (depth_cb < depth) return;

    uint32_t color = 0;
    if (!__builtin_isinf(depth) && R.numInstances > 0) {
        // binary search of the instance holding `total` (resolve.cu:697-721)
        const uint32_t n = R.numInstances;
        const auto prefix = kglobal(R.prefix);
        int left = 0, right = (int)(n - 1u), found = 0;
        while (left <= right) {
            const int mid = left + (right - left) / 2;
            const int64_t before = (int64_t)prefix[mid];
            const int64_t after = ((uint32_t)mid == n - 1u) ? 0xffffffffffll : (int64_t)prefix[mid + 1];
            if (total >= (uint64_t)before && total < (uint64_t)after) { found = mid; break; }
            else if ((uint64_t)after <= total) left = mid + 1;
            else right = mid - 1;
            if (mid == left) break;
        }
        // a wave whose covered pixels share one instance reads its record
        // with scalar loads; otherwise every lane loads its own
        const uint32_t f0 = (uint32_t)__builtin_amdgcn_readfirstlane(found);
        // (ownedOnly: a pixel whose triangle lies outside the listed
        // instances belongs to another pass and is left as it is)
        if (__ballot((uint32_t)found != f0) == 0ull) {
            const InstanceRec mesh = kconst(R.inst)[f0];
            if (R.ownedOnly && total - mesh.cumm >= (uint64_t)mesh.numTriangles) return;
            color = shade_pixel<TEX>(R, mesh, total, x, y);
        } else {
            const InstanceRec mesh = R.inst[found];
            if (R.ownedOnly && total - mesh.cumm >= (uint64_t)mesh.numTriangles) return;
            color = shade_pixel<TEX>(R, mesh, total, x, y);
        }
    } else if (R.ownedOnly && !__builtin_isinf(depth)) {
        return;
    }
    if (depth != __builtin_inff()) {
        R.cb[pixelID] = ((uint64_t)__builtin_bit_cast(uint32_t, depth) << 32) | (uint64_t)color;
    } else {
        R.cb[pixelID] = ((uint64_t)0x7f800000u << 32);
    }
}

// last view whose first block is <= b (firstBlock read through `get`)
template <typename Get>
__device__ __forceinline__ uint32_t find_view(uint32_t numViews, uint32_t b, Get get) {
    uint32_t lo = 0, hi = numViews - 1u;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1u) >> 1;
        if (get(mid) <= b) lo = mid; else hi = mid - 1u;
    }
    return lo;
}

}  // namespace

// Resolve: one lane per pixel, 64x4 pixels per workgroup (a wave reads 512
// contiguous bytes of the visibility buffer).
template <bool TEX>
__global__ __launch_bounds__(256) void k_resolve(ResolveParams R) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= R.W || y >= R.H) return;
    resolve_pixel<TEX>(R, x, y);
}

// Resolve of several views in one launch: block b resolves 64x4 tile
// b - firstBlock[v] of view v, whose record is read by scalar loads.
template <bool TEX>
__global__ __launch_bounds__(256) void k_resolve_views(const ResolveParams* __restrict__ views, uint32_t numViews) {
    const auto rv = kconst(views);
    const uint32_t v = find_view(numViews, blockIdx.x, [&](uint32_t i) { return rv[i].firstBlock; });
    const ResolveParams R = rv[v];
    const uint32_t local = blockIdx.x - R.firstBlock;
    const uint32_t tilesX = ((uint32_t)R.W + 63u) / 64u;
    const int x = (int)(local % tilesX) * 64 + (int)(threadIdx.x & 63);
    const int y = (int)(local / tilesX) * 4 + (int)(threadIdx.x >> 6);
    if (x >= R.W || y >= R.H) return;
    resolve_pixel<TEX>(R, x, y);
}

// Clear of several framebuffers in one launch: block b clears the 4096-pixel
// chunk b - firstBlock[v] of view v (each wave stores 512 contiguous bytes per
// step), and block 0 resets the draw counters.
__global__ __launch_bounds__(256) void k_clear_views(const ClearView* __restrict__ views, uint32_t numViews,
                                                     uint32_t* __restrict__ ctr) {
    const auto cv = kconst(views);
    const uint32_t v = find_view(numViews, blockIdx.x, [&](uint32_t i) { return cv[i].firstBlock; });
    const ClearView c = cv[v];
    const uint32_t base = (blockIdx.x - c.firstBlock) * kClearChunk;
    const uint32_t end = min(c.numPixels, base + kClearChunk);
    for (uint32_t i = base + threadIdx.x; i < end; i += 256u) c.fb[i] = kClearPixel;
    if (blockIdx.x == 0 && ctr)
        for (uint32_t i = threadIdx.x; i < CTR_COUNT; i += blockDim.x) ctr[i] = 0u;
}

// Plain streaming copy, 16 bytes per lane, four nontemporal loads in flight
// per lane and nontemporal stores (calibrates the HBM traffic counters and
// gives a measured copy bandwidth; at 32 workgroups per CU it matches the
// runtime's device-to-device copy).
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void k_copy(u32x4* __restrict__ dst, const u32x4* __restrict__ src, size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + 3 * stride < n; i += 4 * stride) {
        u32x4 v[4];
        #pragma unroll
        for (int u = 0; u < 4; u++) v[u] = __builtin_nontemporal_load(src + i + u * stride);
        #pragma unroll
        for (int u = 0; u < 4; u++) __builtin_nontemporal_store(v[u], dst + i + u * stride);
    }
    for (; i < n; i += stride) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

// Check of rcp_refined / div_fast against the IEEE division: lanes sweep all
// floats b with kRcpLo <= |b| < kRcpHi (biased exponents 2..251, both signs)
// and count b with rcp_refined(b) != 1/b (out[0]); then `pairs` hashed
// operand pairs (a, b) with exponents in [-60, 63] and any mantissa and sign
// count div_fast(a, b, rcp_refined(b)) != a/b (out[1]), and as many pairs
// with |a| below kDivLo (zero and subnormals included) count
// q * 0.5 + 0.5 differing between the two quotients (out[2]).
__global__ __launch_bounds__(256) void k_check_division(uint64_t pairs, unsigned long long* __restrict__ out) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    constexpr uint64_t kPerSign = 250ull << 23;          // exponents 2..251
    uint32_t bad = 0u;
    for (uint64_t k = tid; k < 2ull * kPerSign; k += stride) {
        const uint64_t j = k % kPerSign;
        const uint32_t bits = (uint32_t)((k / kPerSign) << 31) | (uint32_t)((j + (2ull << 23)));
        const float b = __builtin_bit_cast(float, bits);
        const float r = rcp_refined(b);
        const float q = 1.0f / b;
        bad += __builtin_bit_cast(uint32_t, r) != __builtin_bit_cast(uint32_t, q) ? 1u : 0u;
    }
    if (bad) atomicAdd(out, (unsigned long long)bad);
    bad = 0u;
    uint32_t bad2 = 0u;
    for (uint64_t k = tid; k < pairs; k += stride) {
        uint64_t h = (k + 1ull) * 0x9E3779B97F4A7C15ull;
        h ^= h >> 31; h *= 0xBF58476D1CE4E5B9ull; h ^= h >> 29; h *= 0x94D049BB133111EBull; h ^= h >> 32;
        const uint32_t ha = (uint32_t)h, hb = (uint32_t)(h >> 32);
        // exponent in [-60, 63] from 7 bits, mantissa and sign from the rest
        const uint32_t ea = (uint32_t)(127 - 60) + ((ha >> 24) & 127u) % 124u;
        const uint32_t eb = (uint32_t)(127 - 60) + ((hb >> 24) & 127u) % 124u;
        const float a = __builtin_bit_cast(float, (ha & 0x807FFFFFu) | (ea << 23));
        const float b = __builtin_bit_cast(float, (hb & 0x807FFFFFu) | (eb << 23));
        const float q = div_fast(a, b, rcp_refined(b));
        const float w = a / b;
        bad += __builtin_bit_cast(uint32_t, q) != __builtin_bit_cast(uint32_t, w) ? 1u : 0u;
        // a tiny numerator: biased exponent 0..66 (|a| < 2^-60)
        const uint32_t et = ((ha >> 24) & 127u) % 67u;
        const float at = __builtin_bit_cast(float, (ha & 0x807FFFFFu) | (et << 23));
        const float qt = div_fast(at, b, rcp_refined(b)) * 0.5f + 0.5f;
        const float wt = (at / b) * 0.5f + 0.5f;
        bad2 += __builtin_bit_cast(uint32_t, qt) != __builtin_bit_cast(uint32_t, wt) ? 1u : 0u;
    }
    if (bad) atomicAdd(out + 1, (unsigned long long)bad);
    if (bad2) atomicAdd(out + 2, (unsigned long long)bad2);
}

// Plain streaming read, 16 bytes per lane and four nontemporal loads in
// flight per lane: the measured read bandwidth of HBM.  The lanes' XOR of
// the data is stored only when it equals a fixed value, so the loads stay.
__global__ __launch_bounds__(256) void k_read(const u32x4* __restrict__ src, size_t n, uint32_t* __restrict__ sink) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t acc = 0u;
    for (; i + 3 * stride < n; i += 4 * stride) {
        u32x4 v[4];
        #pragma unroll
        for (int u = 0; u < 4; u++) v[u] = __builtin_nontemporal_load(src + i + u * stride);
        #pragma unroll
        for (int u = 0; u < 4; u++) acc ^= (v[u].x ^ v[u].y) ^ (v[u].z ^ v[u].w);
    }
    for (; i < n; i += stride) {
        const u32x4 v = __builtin_nontemporal_load(src + i);
        acc ^= (v.x ^ v.y) ^ (v.z ^ v.w);
    }
    if (acc == 0x9E3779B9u) *sink = acc;
}

// 64-bit atomic minimum probe: every lane performs perLane atomicMin on a
// buffer of 2^logEntries entries, at xorshift-hashed positions (scattered)
// or at consecutive positions per wave (coalesced: a wave's 64 lanes hit
// 512 contiguous bytes, as a stage-2/3 row of pixels does).  The reference
// point for the stages' atomic throughput.
__global__ __launch_bounds__(256) void k_atomic_probe(uint64_t* __restrict__ buf, uint32_t logEntries,
                                                      uint32_t perLane, int scattered) {
    const uint32_t gid = blockIdx.x * 256u + threadIdx.x;
    const uint32_t mask = (1u << logEntries) - 1u;
    uint32_t h = gid * 0x9E3779B9u + 0x7F4A7C15u;
    for (uint32_t j = 0; j < perLane; j++) {
        h ^= h << 13;
        h ^= h >> 17;
        h ^= h << 5;
        const uint32_t wave = __builtin_amdgcn_readfirstlane(h);
        const uint32_t pos = scattered ? (h & mask) : ((wave & ~63u) + (threadIdx.x & 63u)) & mask;
        atomicMin((unsigned long long*)(buf + pos), ((unsigned long long)h << 24) | j);
    }
}

// VALU issue-rate probe: every lane runs `iters` rounds of 8 independent
// dependency chains, so enough waves per SIMD hide each chain's latency and
// the issue rate is what is measured.  mode 0: v_fma_f32 chains (8 VALU
// instructions per round); mode 1: 32-bit integer add and xor chains (16
// per round as compiled: the address and bit arithmetic that dominates
// stage 1); modes 2 and 3: the fma chains on 32 lanes or one lane of each
// wave; mode 4: 4 fma and 4 integer chains interleaved in every wave; mode
// 5: fma chains on even waves and integer chains on odd ones (the two
// instruction mixes under which a SIMD may issue two VALU instructions in
// one quad-cycle).  The SQ_INSTS_VALU counter of a launch confirms the
// count.  The
// result is stored only when it matches a value no input produces, which
// keeps the chains live.
__global__ __launch_bounds__(256) void k_valu_probe(uint32_t iters, int mode, uint32_t s1, uint32_t s2,
                                                    uint32_t* __restrict__ sink) {
    const uint32_t gid = blockIdx.x * 256u + threadIdx.x;
    // modes 2 and 3 run the fma chains with half of each wave's lanes, or
    // one lane, active: whether a VALU instruction's cost depends on EXEC
    if ((mode == 2 && (threadIdx.x & 63u) >= 32u) || (mode == 3 && (threadIdx.x & 63u) != 0u)) return;
    // mode 5: even waves run the fma chains, odd waves the integer chains
    if (mode == 5) mode = (threadIdx.x & 64u) ? 1 : 0;
    if (mode == 4) {
        // 4 fma chains and 4 integer add/xor chains interleaved in each wave
        float f[4];
        uint32_t u[4];
        const float m = __uint_as_float(0x3F7FFFF0u | (s1 & 15u)), c = __uint_as_float(0x3A800000u | (s2 & 255u));
        #pragma unroll
        for (int k = 0; k < 4; k++) {
            f[k] = (float)(gid + (uint32_t)k) * 1e-6f;
            u[k] = gid * 8u + (uint32_t)k;
        }
        for (uint32_t j = 0; j < iters; j++) {
            #pragma unroll
            for (int k = 0; k < 4; k++) {
                f[k] = __builtin_fmaf(f[k], m, c);
                u[k] = (u[k] + s1) ^ s2;
            }
        }
        uint32_t r = 0u;
        #pragma unroll
        for (int k = 0; k < 4; k++) r ^= u[k] ^ __float_as_uint(f[k]);
        if (r == 0x9E3779B9u && s1 == 0u) *sink = gid;
        return;
    }
    if (mode != 1) {
        float a[8];
        const float m = __uint_as_float(0x3F7FFFF0u | (s1 & 15u)), c = __uint_as_float(0x3A800000u | (s2 & 255u));
        #pragma unroll
        for (int k = 0; k < 8; k++) a[k] = (float)(gid + (uint32_t)k) * 1e-6f;
        for (uint32_t j = 0; j < iters; j++) {
            #pragma unroll
            for (int k = 0; k < 8; k++) a[k] = __builtin_fmaf(a[k], m, c);
        }
        float r = 0.0f;
        #pragma unroll
        for (int k = 0; k < 8; k++) r += a[k];
        if (__float_as_uint(r) == 0x7F800001u) *sink = gid;
    } else {
        uint32_t a[8];
        #pragma unroll
        for (int k = 0; k < 8; k++) a[k] = gid * 8u + (uint32_t)k;
        for (uint32_t j = 0; j < iters; j++) {
            #pragma unroll
            for (int k = 0; k < 8; k++) a[k] = (a[k] + s1) ^ s2;
        }
        uint32_t r = 0u;
        #pragma unroll
        for (int k = 0; k < 8; k++) r ^= a[k];
        if (r == 0x9E3779B9u && s1 == 0u) *sink = gid;
    }
}

// One level of a texture's mip chain (kernel_computeMipMap,
// textureTools.cu:41-133): texel (x, y) of the target level is the
// per-channel sum of the source texels (2x..2x+1, 2y..2y+1) that lie inside
// the source level (a missing one adds 0), divided by 4 and truncated; its
// alpha byte is `alpha`.  One lane per target texel, grid-stride; neighbouring
// lanes read neighbouring texel pairs of two source rows.
__global__ __launch_bounds__(256) void k_mip_level(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst,
                                                   uint32_t sw, uint32_t sh, uint32_t tw, uint32_t n,
                                                   uint32_t alpha) {
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) {
        const uint32_t y = i / tw;
        const uint32_t x = i - y * tw;
        const uint32_t sx = 2u * x, sy = 2u * y;
        uint32_t r = 0u, g = 0u, b = 0u;
        auto add = [&](uint32_t px, uint32_t py) {
            if (px >= sw || py >= sh) return;
            const uint32_t t = src[(uint64_t)py * sw + px];
            r += t & 0xffu;
            g += (t >> 8) & 0xffu;
            b += (t >> 16) & 0xffu;
        };
        add(sx, sy);
        add(sx, sy + 1u);
        add(sx + 1u, sy);
        add(sx + 1u, sy + 1u);
        dst[i] = (alpha << 24) | ((b >> 2) << 16) | ((g >> 2) << 8) | (r >> 2);
    }
}

// ---------------------------------------------------------------------------
// Launch wrappers used by the host glue.
// ---------------------------------------------------------------------------

hipError_t launch_mip_level(const uint32_t* src, uint32_t* dst, uint32_t sw, uint32_t sh, uint32_t tw, uint32_t th,
                            uint32_t alpha, int numCUs, hipStream_t stream) {
    const uint32_t n = tw * th;
    uint32_t blocks = (n + 255u) / 256u;
    const uint32_t cap = (uint32_t)numCUs * 16u;
    if (blocks > cap) blocks = cap;
    if (blocks == 0) blocks = 1;
    hipLaunchKernelGGL(k_mip_level, dim3(blocks), dim3(256), 0, stream, src, dst, sw, sh, tw, n, alpha);
    return hipGetLastError();
}

hipError_t launch_clear(uint64_t* fb, uint32_t numPixels, uint32_t* ctr, int numCUs, hipStream_t stream,
                        uint64_t* fb2, uint32_t numPixels2) {
    uint32_t blocks = (std::max(numPixels, numPixels2) + 255u) / 256u;
    const uint32_t cap = (uint32_t)numCUs * 16u;
    if (blocks > cap) blocks = cap;
    if (blocks == 0) blocks = 1;
    hipLaunchKernelGGL(k_clear, dim3(blocks), dim3(256), 0, stream, fb, numPixels, fb2, fb2 ? numPixels2 : 0u, ctr);
    return hipGetLastError();
}

hipError_t launch_reset_counters(uint32_t* ctr, hipStream_t stream) {
    hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(64), 0, stream, ctr);
    return hipGetLastError();
}

hipError_t launch_reset_counter_set(const CounterSet& c, hipStream_t stream) {
    hipLaunchKernelGGL(k_reset_counter_set, dim3(1), dim3(64), 0, stream, c);
    return hipGetLastError();
}

hipError_t launch_store_view(const ViewParams& v, ViewParams* dst, hipStream_t stream) {
    hipLaunchKernelGGL(k_store_view, dim3(1), dim3(64), 0, stream, v, dst);
    return hipGetLastError();
}

hipError_t launch_plan(const crx_cmesh* list, uint32_t numItems, const crx_cmesh* instances,
                       const crx_mat4* transforms, uint32_t numInstances, int instanced, int pu16,
                       uint32_t batchShift, DrawItem* items, InstanceRec* inst, Transform* xf, uint32_t* batchCount,
                       hipStream_t stream) {
    hipLaunchKernelGGL(k_plan_items, dim3(1), dim3(1024), 0, stream, list, numItems, numInstances, instanced, pu16,
                       batchShift, items, batchCount);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_plan_instances, dim3((numInstances + 255u) / 256u), dim3(256), 0, stream, instances,
                       transforms, numInstances, inst, xf);
    return hipGetLastError();
}

namespace {

template <int BLOCK, int K, bool PU16>
hipError_t launch_s1(const RasterParams& P, uint32_t grid, bool views, bool boxes, hipStream_t stream) {
    if (views) hipLaunchKernelGGL((k_stage1_views<BLOCK, K, PU16>), dim3(grid), dim3(BLOCK), 0, stream, P);
    else if (boxes) hipLaunchKernelGGL((k_stage1<BLOCK, K, PU16, true>), dim3(grid), dim3(BLOCK), 0, stream, P);
    else hipLaunchKernelGGL((k_stage1<BLOCK, K, PU16, false>), dim3(grid), dim3(BLOCK), 0, stream, P);
    return hipGetLastError();
}

template <int BLOCK, int K>
hipError_t launch_s1_fmt(const RasterParams& P, uint32_t grid, bool pu16, bool views, bool boxes, hipStream_t stream) {
    return pu16 ? launch_s1<BLOCK, K, true>(P, grid, views, boxes, stream)
                : launch_s1<BLOCK, K, false>(P, grid, views, boxes, stream);
}

template <int BLOCK>
hipError_t launch_s1_k(const RasterParams& P, uint32_t k, uint32_t grid, bool pu16, bool views, bool boxes,
                       hipStream_t stream) {
    switch (k) {
        case 1: return launch_s1_fmt<BLOCK, 1>(P, grid, pu16, views, boxes, stream);
        case 2: return launch_s1_fmt<BLOCK, 2>(P, grid, pu16, views, boxes, stream);
        case 4: return launch_s1_fmt<BLOCK, 4>(P, grid, pu16, views, boxes, stream);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_s1_any(const RasterParams& P, uint32_t block, uint32_t k, uint32_t grid, bool pu16, bool views,
                         bool boxes, hipStream_t stream) {
    switch (block) {
        case 128: return launch_s1_k<128>(P, k, grid, pu16, views, boxes, stream);
        case 256: return launch_s1_k<256>(P, k, grid, pu16, views, boxes, stream);
        case 512: return launch_s1_k<512>(P, k, grid, pu16, views, boxes, stream);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

hipError_t launch_stage1(const RasterParams& P, uint32_t block, uint32_t trisPerLane, uint32_t grid, bool pu16,
                         bool instanceBoxes, hipStream_t stream) {
    return launch_s1_any(P, block, trisPerLane, grid, pu16, false, instanceBoxes, stream);
}

hipError_t launch_stage1_pairs(const RasterParams& P, uint32_t block, uint32_t grid, bool pu16, hipStream_t stream) {
    if (block != 256) return hipErrorInvalidValue;
    if (pu16) hipLaunchKernelGGL((k_stage1_pairs<256, true>), dim3(grid), dim3(256), 0, stream, P);
    else hipLaunchKernelGGL((k_stage1_pairs<256, false>), dim3(grid), dim3(256), 0, stream, P);
    return hipGetLastError();
}

hipError_t launch_stage1_views(const RasterParams& P, uint32_t block, uint32_t trisPerLane, uint32_t grid, bool pu16,
                               hipStream_t stream) {
    return launch_s1_any(P, block, trisPerLane, grid, pu16, true, false, stream);
}

hipError_t launch_resolve_views(const ResolveParams* views, uint32_t numViews, uint32_t numBlocks, bool full,
                                hipStream_t stream) {
    if (numViews == 0 || numBlocks == 0) return hipSuccess;
    if (full) hipLaunchKernelGGL(k_resolve_views<true>, dim3(numBlocks), dim3(256), 0, stream, views, numViews);
    else hipLaunchKernelGGL(k_resolve_views<false>, dim3(numBlocks), dim3(256), 0, stream, views, numViews);
    return hipGetLastError();
}

hipError_t launch_clear_views(const ClearView* views, uint32_t numViews, uint32_t numBlocks, uint32_t* ctr,
                              hipStream_t stream) {
    if (numViews == 0 || numBlocks == 0) return hipSuccess;
    hipLaunchKernelGGL(k_clear_views, dim3(numBlocks), dim3(256), 0, stream, views, numViews, ctr);
    return hipGetLastError();
}

hipError_t launch_stage2(const RasterParams& P, uint32_t grid, hipStream_t stream) {
    hipLaunchKernelGGL(k_stage2, dim3(grid), dim3(256), 0, stream, P);
    return hipGetLastError();
}

hipError_t launch_stage3(const RasterParams& P, uint32_t grid, hipStream_t stream) {
    hipLaunchKernelGGL(k_stage3, dim3(grid), dim3(256), 0, stream, P);
    return hipGetLastError();
}

hipError_t launch_atomic_probe(uint64_t* buf, uint32_t logEntries, uint32_t perLane, int scattered, uint32_t blocks,
                               hipStream_t stream) {
    hipLaunchKernelGGL(k_atomic_probe, dim3(blocks), dim3(256), 0, stream, buf, logEntries, perLane, scattered);
    return hipGetLastError();
}

hipError_t launch_valu_probe(uint32_t iters, int mode, uint32_t blocks, uint32_t* sink, hipStream_t stream) {
    hipLaunchKernelGGL(k_valu_probe, dim3(blocks), dim3(256), 0, stream, iters, mode, 0x2545F491u, 0x6C8E9CF5u, sink);
    return hipGetLastError();
}

hipError_t launch_copy(void* dst, const void* src, size_t bytes, int numCUs, hipStream_t stream) {
    const size_t n = bytes / 16;
    uint32_t blocks = (uint32_t)std::min<size_t>((n + 255) / 256, (size_t)numCUs * 32);
    if (blocks == 0) blocks = 1;
    hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, stream, (u32x4*)dst, (const u32x4*)src, n);
    return hipGetLastError();
}

hipError_t launch_read(const void* src, size_t bytes, int numCUs, uint32_t* sink, hipStream_t stream) {
    const size_t n = bytes / 16;
    uint32_t blocks = (uint32_t)std::min<size_t>((n + 255) / 256, (size_t)numCUs * 32);
    if (blocks == 0) blocks = 1;
    hipLaunchKernelGGL(k_read, dim3(blocks), dim3(256), 0, stream, (const u32x4*)src, n, sink);
    return hipGetLastError();
}

hipError_t launch_check_division(uint64_t pairs, int numCUs, unsigned long long* out, hipStream_t stream) {
    hipLaunchKernelGGL(k_check_division, dim3((uint32_t)numCUs * 32u), dim3(256), 0, stream, pairs, out);
    return hipGetLastError();
}

hipError_t launch_resolve(const ResolveParams& R, hipStream_t stream) {
    const dim3 grid((uint32_t)(R.W + 63) / 64u, (uint32_t)(R.H + 3) / 4u);
    if (full_resolve((crx_display_attribute)R.attribute)) hipLaunchKernelGGL(k_resolve<true>, grid, dim3(256), 0, stream, R);
    else hipLaunchKernelGGL(k_resolve<false>, grid, dim3(256), 0, stream, R);
    return hipGetLastError();
}

}  // namespace crx
