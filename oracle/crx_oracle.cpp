// crx_oracle.cpp — CPU oracle (test infrastructure only; see crx_oracle.h).
//
// A line-by-line restatement of CuRast's triangle visibility-buffer kernels in
// scalar C++17.  Each function cites the reference code it follows (paths are
// relative to the CuRast-OOC source tree).  Build with -ffp-contract=off and
// without -ffast-math: every float operation below is one IEEE binary32
// operation, evaluated in the order GLM 0.9.9 evaluates it.
//
// Parity status of each part (how it is pinned) is listed in tests/; nothing
// here is shared with the HIP product code.

#include "crx_oracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr float kInf = std::numeric_limits<float>::infinity();

struct V2 { float x, y; };
struct V3 { float x, y, z; };
struct V4 { float x, y, z, w; };
typedef crx_mat4 M4;

inline uint32_t fbits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline float bitsf(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

// Float -> int32 as the reference's device code converts (round toward zero,
// NaN -> 0, saturating), i.e. PTX cvt.rzi.s32.f32.
inline int32_t f2i(float x) {
    if (x != x) return 0;
    if (x >= 2147483648.0f) return 2147483647;
    if (x <= -2147483648.0f) return (-2147483647 - 1);
    return (int32_t)x;
}
// Float -> uint8 of a value already clamped to [0,255] (NaN -> 0).
inline uint32_t f2u8(float x) {
    if (!(x > 0.0f)) return 0;
    if (x >= 255.0f) return 255;
    return (uint32_t)x;
}

inline V3 xyz(V4 v) { return V3{v.x, v.y, v.z}; }
inline V3 sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 add(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 mul(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
inline V3 mul(float s, V3 a) { return V3{s * a.x, s * a.y, s * a.z}; }
inline V2 sub(V2 a, V2 b) { return V2{a.x - b.x, a.y - b.y}; }

// glm func_geometric.inl:58-62: tmp = x*y; (tmp.x + tmp.y) + tmp.z
inline float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
// glm func_geometric.inl:74-84
inline V3 cross(V3 x, V3 y) {
    return V3{x.y * y.z - y.y * x.z, x.z * y.x - y.z * x.x, x.x * y.y - y.x * x.y};
}
// glm func_exponential.inl:128-133 (1/sqrt) and func_geometric.inl:86-95
inline V3 normalize(V3 v) { float s = 1.0f / std::sqrt(dot(v, v)); return mul(v, s); }
// rasterization_helpers.cuh:541-543
inline float cross2(V2 a, V2 b) { return a.x * b.y - a.y * b.x; }

// glm type_mat4x4.inl:507-550: (c0*x + c1*y) + (c2*z + c3*w)
inline V4 mulmv(const M4& m, V4 v) {
    const float* a = m.m;
    return V4{(a[0] * v.x + a[4] * v.y) + (a[8] * v.z + a[12] * v.w),
              (a[1] * v.x + a[5] * v.y) + (a[9] * v.z + a[13] * v.w),
              (a[2] * v.x + a[6] * v.y) + (a[10] * v.z + a[14] * v.w),
              (a[3] * v.x + a[7] * v.y) + (a[11] * v.z + a[15] * v.w)};
}

// glm type_mat4x4.inl:601-616: Result[j] = ((A0*b0 + A1*b1) + A2*b2) + A3*b3
M4 mulmm(const M4& A, const M4& B) {
    M4 R;
    for (int j = 0; j < 4; j++) {
        const float* b = B.m + 4 * j;
        for (int r = 0; r < 4; r++) {
            R.m[4 * j + r] = ((A.m[r] * b[0] + A.m[4 + r] * b[1]) + A.m[8 + r] * b[2]) + A.m[12 + r] * b[3];
        }
    }
    return R;
}

// glm func_matrix.inl:298-395 (compute_inverse<tmat4x4>)
M4 inverse(const M4& M) {
    auto m = [&](int c, int r) { return M.m[4 * c + r]; };
    float Coef00 = m(2,2) * m(3,3) - m(3,2) * m(2,3);
    float Coef02 = m(1,2) * m(3,3) - m(3,2) * m(1,3);
    float Coef03 = m(1,2) * m(2,3) - m(2,2) * m(1,3);
    float Coef04 = m(2,1) * m(3,3) - m(3,1) * m(2,3);
    float Coef06 = m(1,1) * m(3,3) - m(3,1) * m(1,3);
    float Coef07 = m(1,1) * m(2,3) - m(2,1) * m(1,3);
    float Coef08 = m(2,1) * m(3,2) - m(3,1) * m(2,2);
    float Coef10 = m(1,1) * m(3,2) - m(3,1) * m(1,2);
    float Coef11 = m(1,1) * m(2,2) - m(2,1) * m(1,2);
    float Coef12 = m(2,0) * m(3,3) - m(3,0) * m(2,3);
    float Coef14 = m(1,0) * m(3,3) - m(3,0) * m(1,3);
    float Coef15 = m(1,0) * m(2,3) - m(2,0) * m(1,3);
    float Coef16 = m(2,0) * m(3,2) - m(3,0) * m(2,2);
    float Coef18 = m(1,0) * m(3,2) - m(3,0) * m(1,2);
    float Coef19 = m(1,0) * m(2,2) - m(2,0) * m(1,2);
    float Coef20 = m(2,0) * m(3,1) - m(3,0) * m(2,1);
    float Coef22 = m(1,0) * m(3,1) - m(3,0) * m(1,1);
    float Coef23 = m(1,0) * m(2,1) - m(2,0) * m(1,1);
    float Fac0[4] = {Coef00, Coef00, Coef02, Coef03};
    float Fac1[4] = {Coef04, Coef04, Coef06, Coef07};
    float Fac2[4] = {Coef08, Coef08, Coef10, Coef11};
    float Fac3[4] = {Coef12, Coef12, Coef14, Coef15};
    float Fac4[4] = {Coef16, Coef16, Coef18, Coef19};
    float Fac5[4] = {Coef20, Coef20, Coef22, Coef23};
    float Vec0[4] = {m(1,0), m(0,0), m(0,0), m(0,0)};
    float Vec1[4] = {m(1,1), m(0,1), m(0,1), m(0,1)};
    float Vec2[4] = {m(1,2), m(0,2), m(0,2), m(0,2)};
    float Vec3[4] = {m(1,3), m(0,3), m(0,3), m(0,3)};
    float Inv0[4], Inv1[4], Inv2[4], Inv3[4];
    for (int i = 0; i < 4; i++) {
        // tvec4 expressions A*B - C*D + E*F evaluate as (A*B - C*D) + E*F
        Inv0[i] = (Vec1[i] * Fac0[i] - Vec2[i] * Fac1[i]) + Vec3[i] * Fac2[i];
        Inv1[i] = (Vec0[i] * Fac0[i] - Vec2[i] * Fac3[i]) + Vec3[i] * Fac4[i];
        Inv2[i] = (Vec0[i] * Fac1[i] - Vec1[i] * Fac3[i]) + Vec3[i] * Fac5[i];
        Inv3[i] = (Vec0[i] * Fac2[i] - Vec1[i] * Fac4[i]) + Vec2[i] * Fac5[i];
    }
    const float SignA[4] = {+1, -1, +1, -1};
    const float SignB[4] = {-1, +1, -1, +1};
    M4 I;
    for (int i = 0; i < 4; i++) {
        I.m[0 + i] = Inv0[i] * SignA[i];
        I.m[4 + i] = Inv1[i] * SignB[i];
        I.m[8 + i] = Inv2[i] * SignA[i];
        I.m[12 + i] = Inv3[i] * SignB[i];
    }
    float Row0[4] = {I.m[0], I.m[4], I.m[8], I.m[12]};
    float Dot0[4] = {m(0,0) * Row0[0], m(0,1) * Row0[1], m(0,2) * Row0[2], m(0,3) * Row0[3]};
    float Dot1 = (Dot0[0] + Dot0[1]) + (Dot0[2] + Dot0[3]);
    float OneOverDeterminant = 1.0f / Dot1;
    for (int i = 0; i < 16; i++) I.m[i] = I.m[i] * OneOverDeterminant;
    return I;
}

// rasterization_helpers.cuh:35-52
uint64_t pack_pixel(float depth, uint64_t trianglePrefix) {
    uint64_t udepth = fbits(depth);
    udepth = udepth & 0x7FFFFFF8ull;
    if (std::isinf(depth)) udepth = 0x7FFFFFF8ull;
    udepth = udepth << 33;
    return udepth | trianglePrefix;
}

// rasterization_helpers.cuh:54-66
void unpack_pixel(uint64_t packed, float* depth, uint64_t* trianglePrefix) {
    uint32_t udepth = (uint32_t)(packed >> 33);
    udepth = udepth & 0x7FFFFFF8u;
    if (udepth == 0x7FFFFFF8u) *depth = kInf; else *depth = bitsf(udepth);
    *trianglePrefix = packed & 0xFFFFFFFFFull;
}

// rasterization_helpers.cuh:95-101
inline V3 viewToNDC(V3 v, float f, float aspect) {
    float depth = -v.z;
    float x_ndc = (f / aspect) * v.x / depth;
    float y_ndc = f * v.y / depth;
    return V3{x_ndc, y_ndc, depth};
}

// rasterization_helpers.cuh:103-108
inline V2 ndcToScreen(V3 ndc, float width, float height) {
    return V2{(ndc.x * 0.5f + 0.5f) * width, (ndc.y * 0.5f + 0.5f) * height};
}

// utils.cuh:205-214
inline float clampf(float v, float lo, float hi) {
    if (v < lo) return lo;
    if (v > hi) return hi;
    return v;
}

inline void atomic_min_u64(uint64_t* p, uint64_t v) {
    uint64_t old = __atomic_load_n(p, __ATOMIC_RELAXED);
    while (v < old) {
        if (__atomic_compare_exchange_n(p, &old, v, true, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) break;
    }
}

// BitEdit.h:8-26
uint32_t bitedit_read(const uint32_t* data, uint64_t bitOffset, uint64_t bitCount) {
    if (bitCount == 0 || bitCount > 32) return 0;
    uint64_t wordIdx = bitOffset / 32;
    uint32_t bitShift = (uint32_t)(bitOffset % 32);
    uint64_t combined = data[wordIdx];
    if (bitShift + bitCount > 32) combined |= (uint64_t(data[wordIdx + 1]) << 32);
    uint32_t mask = (bitCount == 32) ? 0xFFFFFFFFu : ((1u << bitCount) - 1);
    return uint32_t((combined >> bitShift) & mask);
}

// BitEdit.h:28-49
void bitedit_write(uint32_t* data, uint64_t bitOffset, uint64_t bitCount, uint32_t value) {
    if (bitCount == 0 || bitCount > 32) return;
    uint64_t wordIdx = bitOffset / 32;
    uint32_t bitShift = (uint32_t)(bitOffset % 32);
    uint32_t mask = (bitCount == 32) ? 0xFFFFFFFFu : ((1u << bitCount) - 1);
    value &= mask;
    uint32_t firstMask = ~(mask << bitShift);
    data[wordIdx] = (data[wordIdx] & firstMask) | (value << bitShift);
    if (bitShift + bitCount > 32) {
        uint32_t bitsWritten = 32 - bitShift;
        uint32_t secondMask = ~(mask >> bitsWritten);
        data[wordIdx + 1] = (data[wordIdx + 1] & secondMask) | (value >> bitsWritten);
    }
}

// getVertex<INDEXBUFFER, UNCOMPRESSED>, triangles_visbuffer.cu:37-47, and
// getVertex<INDEXBUFFER, IX_PU16>, :48-62: bit-packed index + index_min,
// 3 x uint16 position scaled by compressionFactor plus aabb.min (a product
// then a sum: the contract has no contraction).
inline V4 getVertex(const crx_cmesh& mesh, uint32_t vertexIndex) {
    if (mesh.compressed) {
        const uint32_t resolvedIndex =
            bitedit_read(mesh.indices, (uint64_t)mesh.bitsPerIndex * vertexIndex, mesh.bitsPerIndex) + mesh.index_min;
        const uint16_t* q = reinterpret_cast<const uint16_t*>(mesh.positions) + 3ull * resolvedIndex;
        return V4{(float)q[0] * mesh.compressionFactor.x + mesh.aabb.min.x,
                  (float)q[1] * mesh.compressionFactor.y + mesh.aabb.min.y,
                  (float)q[2] * mesh.compressionFactor.z + mesh.aabb.min.z, 1.0f};
    }
    uint32_t resolvedIndex = mesh.indices ? mesh.indices[vertexIndex] : vertexIndex;
    const crx_vec3 p = mesh.positions[resolvedIndex];
    return V4{p.x, p.y, p.z, 1.0f};
}

// getVertex_resolved, resolve.cu:69-89: the resolve decodes compressed
// positions with its own formula, (X / 65536) * (aabb.max - aabb.min) + aabb.min.
inline crx_vec3 getVertex_resolved(const crx_cmesh& mesh, uint32_t resolvedIndex) {
    if (!mesh.compressed) return mesh.positions[resolvedIndex];
    const uint16_t* q = reinterpret_cast<const uint16_t*>(mesh.positions) + 3ull * resolvedIndex;
    const float sx = mesh.aabb.max.x - mesh.aabb.min.x;
    const float sy = mesh.aabb.max.y - mesh.aabb.min.y;
    const float sz = mesh.aabb.max.z - mesh.aabb.min.z;
    return crx_vec3{((float)q[0] / 65536.0f) * sx + mesh.aabb.min.x, ((float)q[1] / 65536.0f) * sy + mesh.aabb.min.y,
                    ((float)q[2] / 65536.0f) * sz + mesh.aabb.min.z};
}

struct Target {
    uint64_t* fb;
    int W, H;
    float f, aspect;
};

Target make_target(const crx_render_target* t) {
    Target g;
    g.fb = t->framebuffer;
    g.W = t->width;
    g.H = t->height;
    g.f = t->proj.m[5];                       // proj[1][1]
    g.aspect = float(t->width) / float(t->height);
    return g;
}

struct Counters { uint64_t px = 0; };

// rasterize<INDEXBUFFER, UNCOMPRESSED>, triangles_visbuffer.cu:83-271
void stage1_rasterize(const Target& tg, const crx_tuning& tu, const crx_cmesh& sh_mesh,
                      const M4& worldView, int triangleIndex, int instanceIndex,
                      V4 a_object, V4 b_object, V4 c_object,
                      std::vector<uint64_t>& queue, Counters& cnt) {
    if ((uint32_t)triangleIndex >= sh_mesh.numTriangles) return;
    const float f = tg.f, aspect = tg.aspect;

    V3 a_view = xyz(mulmv(worldView, a_object));
    V3 b_view = xyz(mulmv(worldView, b_object));
    V3 c_view = xyz(mulmv(worldView, c_object));
    V3 a_ndc = viewToNDC(a_view, f, aspect);
    V3 b_ndc = viewToNDC(b_view, f, aspect);
    V3 c_ndc = viewToNDC(c_view, f, aspect);

    bool isNontrivial = (a_ndc.z <= 0.0f || b_ndc.z <= 0.0f || c_ndc.z <= 0.0f);
    if (!isNontrivial) {
        if (a_ndc.x > +1.0f && b_ndc.x > +1.0f && c_ndc.x > +1.0f) return;
        if (a_ndc.x < -1.0f && b_ndc.x < -1.0f && c_ndc.x < -1.0f) return;
        if (a_ndc.y > +1.0f && b_ndc.y > +1.0f && c_ndc.y > +1.0f) return;
        if (a_ndc.y < -1.0f && b_ndc.y < -1.0f && c_ndc.y < -1.0f) return;
    }
    if (a_ndc.z <= 0.0f && b_ndc.z <= 0.0f && c_ndc.z <= 0.0f) return;

    const float Wf = float(tg.W), Hf = float(tg.H);
    V2 a_screen = ndcToScreen(a_ndc, Wf, Hf);
    V2 b_screen = ndcToScreen(b_ndc, Wf, Hf);
    V2 c_screen = ndcToScreen(c_ndc, Wf, Hf);
    a_screen.x -= 0.5f; a_screen.y -= 0.5f;
    b_screen.x -= 0.5f; b_screen.y -= 0.5f;
    c_screen.x -= 0.5f; c_screen.y -= 0.5f;
    constexpr float SAMPLE_OFFSET = 0.0f;

    float min_x = std::fmin(a_screen.x, std::fmin(b_screen.x, c_screen.x));
    float max_x = std::fmax(a_screen.x, std::fmax(b_screen.x, c_screen.x));
    float min_y = std::fmin(a_screen.y, std::fmin(b_screen.y, c_screen.y));
    float max_y = std::fmax(a_screen.y, std::fmax(b_screen.y, c_screen.y));

    min_x = std::fmax(min_x, 0.0f);
    max_x = std::fmin(max_x, float(tg.W - 1));
    min_y = std::fmax(min_y, 0.0f);
    max_y = std::fmin(max_y, float(tg.H - 1));

    if (!isNontrivial) {
        float sample_x = std::floor(min_x);
        float sample_y = std::floor(min_y);
        if (min_x > sample_x + SAMPLE_OFFSET && max_x < sample_x + 1.0f + SAMPLE_OFFSET) return;
        if (min_y > sample_y + SAMPLE_OFFSET && max_y < sample_y + 1.0f + SAMPLE_OFFSET) return;
    }

    int size_x = f2i(std::ceil(max_x) - std::ceil(min_x));
    int size_y = f2i(std::ceil(max_y) - std::ceil(min_y));
    int numFragments = (int)((uint32_t)size_x * (uint32_t)size_y);
    // `numFragments > THRESHOLD_SMALL` compares int with uint32_t, i.e. unsigned.
    if ((uint32_t)numFragments > tu.small_threshold) isNontrivial = true;

    {   // backface culling in view space, :168-188
        V3 ab = sub(b_view, a_view);
        V3 ac = sub(c_view, a_view);
        V3 N = cross(ab, ac);
        float res = dot(a_view, N);
        V3 c0{worldView.m[0], worldView.m[1], worldView.m[2]};
        V3 c1{worldView.m[4], worldView.m[5], worldView.m[6]};
        V3 c2{worldView.m[8], worldView.m[9], worldView.m[10]};
        float det = dot(cross(c0, c1), c2);
        bool flipTriangles = det < 0.0f;
        if (flipTriangles) res = -res;
        if (res >= 0.0f) return;
    }

    if (isNontrivial) {
        uint64_t packed = (uint64_t((uint32_t)(sh_mesh.instances.offset + instanceIndex)) << 32) |
                          uint64_t((uint32_t)triangleIndex);
        queue.push_back(packed);
        return;
    }

    V2 v_ab = sub(b_screen, a_screen);
    V2 v_ac = sub(c_screen, a_screen);
    float factor = 1.0f / cross2(v_ab, v_ac);
    float inv_z_a = 1.0f / a_ndc.z;
    float inv_z_b = 1.0f / b_ndc.z;
    float inv_z_c = 1.0f / c_ndc.z;
    float ds_dx = v_ac.y * factor;
    float ds_dy = -v_ac.x * factor;
    float dt_dx = -v_ab.y * factor;
    float dt_dy = v_ab.x * factor;
    float start_x = std::ceil(min_x);
    float start_y = std::ceil(min_y);
    float sample0_x = start_x - a_screen.x + SAMPLE_OFFSET;
    float sample0_y = start_y - a_screen.y + SAMPLE_OFFSET;
    float s_row_start = (sample0_x * v_ac.y - sample0_y * v_ac.x) * factor;
    float t_row_start = (v_ab.x * sample0_y - v_ab.y * sample0_x) * factor;
    int row_pixel_id = (int)((uint32_t)f2i(start_x) + (uint32_t)tg.W * (uint32_t)f2i(start_y));
    uint32_t numPixels = (uint32_t)tg.W * (uint32_t)tg.H;
    // 64-bit prefix: cumm + triangleIndex + (uint32)(instanceIndex * numTriangles)
    uint64_t id = sh_mesh.cummulativeTriangleCount + (uint64_t)(int64_t)triangleIndex +
                  (uint64_t)((uint32_t)instanceIndex * sh_mesh.numTriangles);

    for (int fragY = 0; fragY < size_y; ++fragY) {
        float s = s_row_start;
        float t = t_row_start;
        int pixelID = row_pixel_id;
        for (int fragX = 0; fragX < size_x; fragX++) {
            float v = 1.0f - (s + t);
            if (s > 0.0f && t > 0.0f && v > 0.0f) {
                if ((uint32_t)pixelID < numPixels) {
                    float inv_depth = v * inv_z_a + s * inv_z_b + t * inv_z_c;
                    float depth = 1.0f / inv_depth;
                    atomic_min_u64(&tg.fb[(uint32_t)pixelID], pack_pixel(depth, id));
                    cnt.px++;
                }
            }
            s += ds_dx;
            t += dt_dx;
            pixelID++;
        }
        s_row_start += ds_dy;
        t_row_start += dt_dy;
        row_pixel_id += tg.W;
    }
}

int num_threads(int n) {
#ifdef _OPENMP
    return n > 0 ? n : omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

// SutherlandHodgman, rasterization_helpers.cuh:126-210
struct ClipEdge { int axis; float sign; float value; };
inline bool sh_inside(V2 p, ClipEdge e) {
    float val = (e.axis == 0) ? p.x : p.y;
    return (e.sign * val >= e.sign * e.value);
}
inline V2 sh_intersect(V2 p1, V2 p2, ClipEdge e) {
    if (e.axis == 0) {
        float t = (e.value - p1.x) / (p2.x - p1.x);
        return V2{e.value, p1.y + t * (p2.y - p1.y)};
    } else {
        float t = (e.value - p1.y) / (p2.y - p1.y);
        return V2{p1.x + t * (p2.x - p1.x), e.value};
    }
}
void sh_windowClipping(V2 a, V2 b, V2 c, V2* boxMin, V2* boxMax) {
    V2 bufferA[8], bufferB[8];
    V2* in = bufferA;
    V2* out = bufferB;
    in[0] = a; in[1] = b; in[2] = c;
    int inCount = 3;
    const ClipEdge edges[4] = {{0, 1.0f, -1.0f}, {0, -1.0f, 1.0f}, {1, 1.0f, -1.0f}, {1, -1.0f, 1.0f}};
    for (int e = 0; e < 4; e++) {
        int outCount = 0;
        if (inCount == 0) break;
        V2 S = in[inCount - 1];
        for (int i = 0; i < inCount; i++) {
            V2 E = in[i];
            bool E_inside = sh_inside(E, edges[e]);
            bool S_inside = sh_inside(S, edges[e]);
            if (E_inside) {
                if (!S_inside) out[outCount++] = sh_intersect(S, E, edges[e]);
                out[outCount++] = E;
            } else if (S_inside) {
                out[outCount++] = sh_intersect(S, E, edges[e]);
            }
            S = E;
        }
        V2* tmp = in; in = out; out = tmp;
        inCount = outCount;
    }
    for (int i = 0; i < inCount; i++) {
        V2 p = in[i];
        boxMin->x = std::fmin(boxMin->x, p.x);
        boxMin->y = std::fmin(boxMin->y, p.y);
        boxMax->x = std::fmax(boxMax->x, p.x);
        boxMax->y = std::fmax(boxMax->y, p.y);
    }
}

inline V2 xy(V3 v) { return V2{v.x, v.y}; }

// computeScreenSpaceBoundingBox, rasterization_helpers.cuh:350-460
bool stage2_bbox(V3 a_ndc, V3 b_ndc, V3 c_ndc, V3 a_view, V3 b_view, V3 c_view,
                 float f, float aspect, float width, float height,
                 float* min_x, float* max_x, float* min_y, float* max_y) {
    V2 boxMin{1e10f, 1e10f};
    V2 boxMax{-1e10f, -1e10f};
    int mask = (a_ndc.z <= 0.0f ? 1 : 0) | (b_ndc.z <= 0.0f ? 2 : 0) | (c_ndc.z <= 0.0f ? 4 : 0);
    if (mask == 7) return false;
    constexpr float near_ = 0.01f;
    auto intersectNear = [&](V3 p1, V3 p2) {
        float t = (-near_ - p1.z) / (p2.z - p1.z);
        return add(p1, mul(t, sub(p2, p1)));
    };
    if (mask == 0) {
        sh_windowClipping(xy(a_ndc), xy(b_ndc), xy(c_ndc), &boxMin, &boxMax);
    } else {
        V3 v_view[3];
        if (mask == 1 || mask == 6) { v_view[0] = a_view; v_view[1] = b_view; v_view[2] = c_view; }
        else if (mask == 2 || mask == 5) { v_view[0] = b_view; v_view[1] = c_view; v_view[2] = a_view; }
        else { v_view[0] = c_view; v_view[1] = a_view; v_view[2] = b_view; }
        if (mask == 1 || mask == 2 || mask == 4) {
            V3 I1 = intersectNear(v_view[0], v_view[1]);
            V3 I2 = intersectNear(v_view[0], v_view[2]);
            V2 p0 = xy(viewToNDC(v_view[1], f, aspect));
            V2 p1 = xy(viewToNDC(v_view[2], f, aspect));
            V2 pI1 = xy(viewToNDC(I1, f, aspect));
            V2 pI2 = xy(viewToNDC(I2, f, aspect));
            sh_windowClipping(p0, p1, pI1, &boxMin, &boxMax);
            sh_windowClipping(pI1, p1, pI2, &boxMin, &boxMax);
        } else {
            V3 I1 = intersectNear(v_view[0], v_view[1]);
            V3 I2 = intersectNear(v_view[0], v_view[2]);
            sh_windowClipping(xy(viewToNDC(v_view[0], f, aspect)), xy(viewToNDC(I1, f, aspect)),
                              xy(viewToNDC(I2, f, aspect)), &boxMin, &boxMax);
        }
    }
    boxMin.x = clampf(boxMin.x, -1.0f, 1.0f);
    boxMax.x = clampf(boxMax.x, -1.0f, 1.0f);
    boxMin.y = clampf(boxMin.y, -1.0f, 1.0f);
    boxMax.y = clampf(boxMax.y, -1.0f, 1.0f);
    float halfW = width * 0.5f;
    float halfH = height * 0.5f;
    *min_x = boxMin.x * halfW + halfW;
    *min_y = boxMin.y * halfH + halfH;
    *max_x = boxMax.x * halfW + halfW;
    *max_y = boxMax.y * halfH + halfH;
    return true;
}

// stage2_drawMediumTriangles, triangles_visbuffer.cu:388-563 (one queue entry)
void stage2_entry(const Target& tg, const crx_tuning& tu, const crx_cmesh* instances,
                  const M4* transforms, uint64_t packed,
                  std::vector<crx_huge_triangle>& huge, Counters& cnt) {
    uint32_t meshIndex = (uint32_t)(packed >> 32);
    uint32_t triangleIndex = (uint32_t)(packed & 0xffffffffull);
    const crx_cmesh& mesh = instances[meshIndex];
    uint32_t cummulativeOffset = (uint32_t)mesh.cummulativeTriangleCount;
    const M4& worldView = transforms[mesh.instances.offset];
    const float f = tg.f, aspect = tg.aspect;

    V3 a = xyz(getVertex(mesh, 3 * triangleIndex + 0));
    V3 b = xyz(getVertex(mesh, 3 * triangleIndex + 1));
    V3 c = xyz(getVertex(mesh, 3 * triangleIndex + 2));
    V3 a_view = xyz(mulmv(worldView, V4{a.x, a.y, a.z, 1.0f}));
    V3 b_view = xyz(mulmv(worldView, V4{b.x, b.y, b.z, 1.0f}));
    V3 c_view = xyz(mulmv(worldView, V4{c.x, c.y, c.z, 1.0f}));
    // worldToNDC (rasterization_helpers.cuh:86-93) == viewToNDC(worldView * v)
    V3 a_ndc = viewToNDC(a_view, f, aspect);
    V3 b_ndc = viewToNDC(b_view, f, aspect);
    V3 c_ndc = viewToNDC(c_view, f, aspect);

    float min_x, max_x, min_y, max_y;
    if (!stage2_bbox(a_ndc, b_ndc, c_ndc, a_view, b_view, c_view, f, aspect,
                     float(tg.W), float(tg.H), &min_x, &max_x, &min_y, &max_y)) {
        return;  // unreachable: stage 1 never queues triangles fully behind the camera
    }
    constexpr float epsilon = 0.5f;
    min_x -= epsilon; min_y -= epsilon;
    max_x += epsilon; max_y += epsilon;

    V2 a_screen = ndcToScreen(a_ndc, float(tg.W), float(tg.H));
    V2 b_screen = ndcToScreen(b_ndc, float(tg.W), float(tg.H));
    V2 c_screen = ndcToScreen(c_ndc, float(tg.W), float(tg.H));
    V2 v_ab = sub(b_screen, a_screen);
    V2 v_ac = sub(c_screen, a_screen);
    float inv_det = 1.0f / (v_ab.x * v_ac.y - v_ab.y * v_ac.x);
    int size_x = f2i(std::ceil(max_x)) - f2i(std::floor(min_x));
    int size_y = f2i(std::ceil(max_y)) - f2i(std::floor(min_y));
    int numFragments = (int)((uint32_t)size_x * (uint32_t)size_y);

    if (min_x >= max_x || min_y >= max_y) return;

    const int TILE = (int)tu.tile_size;
    if (numFragments >= TILE * TILE || a_ndc.z <= 0.0f || b_ndc.z <= 0.0f || c_ndc.z <= 0.0f) {
        int tile_x = f2i(min_x) / TILE;
        int tile_y = f2i(min_y) / TILE;
        int tiles_x = (f2i(max_x) + TILE - 1) / TILE - tile_x;
        int tiles_y = (f2i(max_y) + TILE - 1) / TILE - tile_y;
        uint32_t numTiles = (uint32_t)tiles_x * (uint32_t)tiles_y;
        for (uint32_t i = 0; i < numTiles; i++) {
            crx_huge_triangle t;
            t.meshIndex = (int32_t)meshIndex;
            t.triangleIndex = (int32_t)triangleIndex;
            t.tile_x = (int32_t)((uint32_t)tile_x + (i % (uint32_t)tiles_x));
            t.tile_y = (int32_t)((uint32_t)tile_y + (i / (uint32_t)tiles_x));
            huge.push_back(t);
        }
        return;
    }

    const float start_x = std::floor(min_x);
    const float start_y = std::floor(min_y);
    const uint64_t packed_id = (uint32_t)(cummulativeOffset + triangleIndex);
    float inv_z_a = 1.0f / a_ndc.z;
    float inv_z_b = 1.0f / b_ndc.z;
    float inv_z_c = 1.0f / c_ndc.z;
    for (int fragOffset = 0; fragOffset < numFragments; fragOffset++) {
        int fragX = fragOffset % size_x;
        int fragY = fragOffset / size_x;
        float px = start_x + (float)fragX;
        float py = start_y + (float)fragY;
        float sx = px - a_screen.x + 0.5f;
        float sy = py - a_screen.y + 0.5f;
        float s = (sx * v_ac.y - sy * v_ac.x) * inv_det;
        float t = (v_ab.x * sy - v_ab.y * sx) * inv_det;
        float v = 1.0f - (s + t);
        if (s >= 0.0f && t >= 0.0f && v >= 0.0f) {
            int ix = f2i(px), iy = f2i(py);
            // normalisation: samples outside the screen are discarded instead of
            // wrapping into the neighbouring row (or before the buffer).
            if (ix < 0 || ix >= tg.W || iy < 0 || iy >= tg.H) continue;
            int pixelID = ix + tg.W * iy;
            float inv_depth = v * inv_z_a + s * inv_z_b + t * inv_z_c;
            float depth = 1.0f / inv_depth;
            atomic_min_u64(&tg.fb[pixelID], pack_pixel(depth, packed_id));
            cnt.px++;
        }
    }
}

// intersectTriangle_mt, double-sided branch, rasterization_helpers.cuh:213-270
float intersect_mt(V3 dir, V3 v0, V3 v1, V3 v2) {
    V3 e1 = sub(v1, v0);
    V3 e2 = sub(v2, v0);
    V3 pvec = cross(dir, e2);
    float det = dot(e1, pvec);
    constexpr float EPSILON = 1e-7f;
    if (det > -EPSILON && det < EPSILON) return kInf;
    float inv_det = 1.0f / det;
    V3 tvec = V3{-v0.x, -v0.y, -v0.z};
    float u = dot(tvec, pvec) * inv_det;
    if (u < 0.0f || u > 1.0f) return kInf;
    V3 qvec = cross(tvec, e1);
    float v = dot(dir, qvec) * inv_det;
    if (v < 0.0f || u + v > 1.0f) return kInf;
    float t = dot(e2, qvec) * inv_det;
    return t > 0.0f ? t : kInf;
}

// stage3_drawHugeTriangles, RAYTRACE branch, triangles_visbuffer.cu:568-745 (one tile)
void stage3_entry(const Target& tg, const crx_tuning& tu, const crx_cmesh* instances,
                  const M4* transforms, const crx_huge_triangle& tri, Counters& cnt) {
    const float f = tg.f, aspect = tg.aspect;
    const crx_cmesh& mesh = instances[tri.meshIndex];
    const M4& worldView = transforms[mesh.instances.offset];
    V4 a_object = getVertex(mesh, 3 * (uint32_t)tri.triangleIndex + 0);
    V4 b_object = getVertex(mesh, 3 * (uint32_t)tri.triangleIndex + 1);
    V4 c_object = getVertex(mesh, 3 * (uint32_t)tri.triangleIndex + 2);
    V3 a_view = xyz(mulmv(worldView, a_object));
    V3 b_view = xyz(mulmv(worldView, b_object));
    V3 c_view = xyz(mulmv(worldView, c_object));
    const int TILE = (int)tu.tile_size;
    float min_x = float(tri.tile_x * TILE);
    float min_y = float(tri.tile_y * TILE);
    min_x = clampf(min_x, 0.0f, (float)tg.W);
    min_y = clampf(min_y, 0.0f, (float)tg.H);
    const float fx0 = std::floor(min_x), fy0 = std::floor(min_y);
    const uint64_t id = mesh.cummulativeTriangleCount + (uint64_t)(int64_t)tri.triangleIndex;
    const V3 viewDir{0.0f, 0.0f, -1.0f};
    for (int y = 0; y < TILE; y++) {
        for (int x = 0; x < TILE; x++) {
            float pfx = fx0 + float(x);
            float pfy = fy0 + float(y);
            int px = f2i(pfx), py = f2i(pfy);
            float u = 2.0f * (float(px) + 0.5f) / float(tg.W) - 1.0f;
            float v = 2.0f * (float(py) + 0.5f) / float(tg.H) - 1.0f;
            V3 rayDir = normalize(V3{(1.0f / (f / aspect)) * u, (1.0f / f) * v, -1.0f});
            float t = intersect_mt(rayDir, a_view, b_view, c_view);
            if (t == kInf) continue;
            // normalisation: columns past the right edge are discarded instead of
            // wrapping into the next row; rows past the bottom are discarded as in
            // the reference (pixelID >= width*height).
            if (px >= tg.W || py >= tg.H) continue;
            int pixelID = px + tg.W * py;
            float depth = dot(mul(t, rayDir), viewDir);
            atomic_min_u64(&tg.fb[pixelID], pack_pixel(depth, id));
            cnt.px++;
        }
    }
}

inline V3 computeRayDir(float x, float y, float width, float height, const M4& viewI,
                        const M4& proj, V3 origin) {
    // resolve.cu:662-676 (inverse(view) hoisted: it is the same value each call)
    float u = 2.0f * x / width - 1.0f;
    float v = 2.0f * y / height - 1.0f;
    V3 rayDir_view = normalize(V3{1.0f / proj.m[0] * u, 1.0f / proj.m[5] * v, -1.0f});
    V3 w = xyz(mulmv(viewI, V4{rayDir_view.x, rayDir_view.y, rayDir_view.z, 1.0f}));
    return normalize(sub(w, origin));
}

// intersectPlane, rasterization_helpers.cuh:318-335
inline float intersectPlane(V3 origin, V3 dir, V3 normal, float constant) {
    float denominator = dot(normal, dir);
    if (denominator == 0.0f) return kInf;
    float t = -(dot(origin, normal) + constant) / denominator;
    return (t >= 0.0f) ? t : kInf;
}

// sampleColor_linear, resolve.cu:148-195
inline uint32_t sampleColor_linear(const uint32_t* textureData, int width, int height, V2 uv) {
    if (textureData == nullptr) return 0;
    float ftx = (uv.x - std::floor(uv.x)) * float(width);
    float fty = (uv.y - std::floor(uv.y)) * float(height);
    auto getTexel = [&](float fx, float fy, float rgba[4]) {
        int tx = f2i(std::fmod(fx, float(width)));
        int ty = f2i(std::fmod(fy, float(height)));
        int texelID = tx + ty * width;
        // clamp(texelID, 0, width * height - 1), utils.cuh:205-212
        const int hi = width * height - 1;
        if (texelID < 0) texelID = 0;
        else if (texelID > hi) texelID = hi;
        uint32_t texel = textureData[texelID];
        for (int k = 0; k < 4; k++) rgba[k] = float((texel >> (8 * k)) & 0xffu);
    };
    float t00[4], t01[4], t10[4], t11[4];
    getTexel(ftx - 0.5f, fty - 0.5f, t00);
    getTexel(ftx - 0.5f, fty + 0.5f, t01);
    getTexel(ftx + 0.5f, fty - 0.5f, t10);
    getTexel(ftx + 0.5f, fty + 0.5f, t11);
    float wx = std::fmod(ftx + 0.5f, 1.0f);
    float wy = std::fmod(fty + 0.5f, 1.0f);
    uint32_t color = 0xff000000u;
    for (int k = 0; k < 3; k++) {
        float interpolated = (1.0f - wx) * (1.0f - wy) * t00[k] + wx * (1.0f - wy) * t10[k] +
                             (1.0f - wx) * wy * t01[k] + wx * wy * t11[k];
        color |= f2u8(interpolated) << (8 * k);
    }
    return color;
}

// Mip level of a TEXTURE sample, resolve.cu:849-856: int(0.5 * log2(m))
// clamped to [0, 7].  Taken as the exact value of the real function (log2
// in double is exact enough: for a float m its trunc never rounds across an
// integer); int(NaN) is 0 as in f2i.
inline int mipLevelOf(float m) {
    double ml = 0.5 * std::log2((double)m);
    if (std::isnan(ml)) return 0;
    ml = std::trunc(ml);
    ml = std::min(ml, 7.0);
    return ml < 0.0 ? 0 : (int)ml;
}

}  // namespace

extern "C" {

void crx_oracle_clear(uint64_t* fb, uint64_t numPixels) {
    for (uint64_t i = 0; i < numPixels; i++) fb[i] = 0xFFFFFFF000000000ull;
}

int crx_oracle_compute_mipmap(uint32_t* data, int32_t width, int32_t height) {
    // textureTools.cu:47: levels = int(log2(max(w, h))), here the exact floor
    // (the float log2 agrees below 2^21); levels 1 .. levels-1 are computed,
    // and (normalisation) at least levels 1 .. 7, the ones the resolve samples
    int64_t levels = 0;
    for (int64_t m = std::max(width, height); m > 1; m >>= 1) levels++;
    const int64_t last = std::max<int64_t>(levels - 1, 7);
    int64_t source_width = width, source_height = height, source_pixelOffset = 0;
    for (int64_t level = 1; level <= last; level++) {
        const int64_t target_width = (source_width + 2 - 1) / 2;
        const int64_t target_height = (source_height + 2 - 1) / 2;
        const int64_t target_pixelOffset = source_pixelOffset + source_width * source_height;
        for (int64_t target_y = 0; target_y < target_height; target_y++) {
            for (int64_t target_x = 0; target_x < target_width; target_x++) {
                // textureTools.cu:76-85: a debug colour per level 1-9, whose
                // alpha byte (0) survives; 0xff0000ff beyond
                uint32_t color = level <= 9 ? 0x00000000u : 0xff0000ffu;
                int32_t sR = 0, sG = 0, sB = 0;
                auto fetchAdd = [&](int64_t source_x, int64_t source_y) {
                    if (source_x >= source_width) return;
                    if (source_y >= source_height) return;
                    uint32_t sample = data[source_pixelOffset + source_x + source_width * source_y];
                    sR += sample & 0xff;
                    sG += (sample >> 8) & 0xff;
                    sB += (sample >> 16) & 0xff;
                };
                fetchAdd(2 * target_x + 0, 2 * target_y + 0);
                fetchAdd(2 * target_x + 0, 2 * target_y + 1);
                fetchAdd(2 * target_x + 1, 2 * target_y + 0);
                fetchAdd(2 * target_x + 1, 2 * target_y + 1);
                color = (color & 0xff000000u) | (uint32_t)(sR / 4) | ((uint32_t)(sG / 4) << 8) |
                        ((uint32_t)(sB / 4) << 16);
                data[target_pixelOffset + target_x + target_width * target_y] = color;
            }
        }
        source_width = target_width;
        source_height = target_height;
        source_pixelOffset = target_pixelOffset;
    }
    return (int)last;
}

}  // extern "C"

namespace {

void stage1_run(const crx_cmesh* meshes, uint32_t numMeshes, const crx_cmesh* instances,
                uint32_t numInstances, const crx_mat4* transforms, const crx_render_target* target,
                int instanced, const crx_tuning* tuning, int nthreads, std::vector<uint64_t>& out) {
    const Target tg = make_target(target);
    const crx_tuning tu = *tuning;
    // Walk order of stage1_drawSmallTriangles (triangles_visbuffer.cu:274-384):
    // non-instanced treats `instances` as meshes, instanced walks unique meshes
    // and loops over instances{offset,count} reusing the fetched vertices.
    const crx_cmesh* list = instanced ? meshes : instances;
    const uint32_t n = instanced ? numMeshes : numInstances;
    std::vector<uint64_t> firstTri(n + 1, 0);
    for (uint32_t k = 0; k < n; k++) firstTri[k + 1] = firstTri[k] + list[k].numTriangles;
    const uint64_t total = firstTri[n];
    const int nt = num_threads(nthreads);
    std::vector<std::vector<uint64_t>> queues(nt);
    const int64_t chunk = 4096;
    const int64_t numChunks = (int64_t)((total + chunk - 1) / chunk);
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
    for (int64_t ch = 0; ch < numChunks; ch++) {
#ifdef _OPENMP
        const int tid = omp_get_thread_num();
#else
        const int tid = 0;
#endif
        Counters cnt;
        uint64_t g0 = (uint64_t)ch * chunk;
        uint64_t g1 = std::min<uint64_t>(total, g0 + chunk);
        uint32_t k = (uint32_t)(std::upper_bound(firstTri.begin(), firstTri.end(), g0) - firstTri.begin()) - 1;
        for (uint64_t g = g0; g < g1; g++) {
            while (g >= firstTri[k + 1]) k++;
            const crx_cmesh& sh_mesh = list[k];
            int triangleIndex = (int)(g - firstTri[k]);
            V4 a = getVertex(sh_mesh, 3 * (uint32_t)triangleIndex + 0);
            V4 b = getVertex(sh_mesh, 3 * (uint32_t)triangleIndex + 1);
            V4 c = getVertex(sh_mesh, 3 * (uint32_t)triangleIndex + 2);
            if (!instanced) {
                const M4& worldView = transforms[sh_mesh.instances.offset];
                stage1_rasterize(tg, tu, sh_mesh, worldView, triangleIndex, 0, a, b, c, queues[tid], cnt);
            } else {
                for (int instanceIndex = 0; instanceIndex < sh_mesh.instances.count; instanceIndex++) {
                    const M4& worldView = transforms[sh_mesh.instances.offset + instanceIndex];
                    stage1_rasterize(tg, tu, sh_mesh, worldView, triangleIndex, instanceIndex, a, b, c, queues[tid], cnt);
                }
            }
        }
    }
    for (auto& q : queues) out.insert(out.end(), q.begin(), q.end());
}

void stage2_run(const crx_cmesh* instances, const crx_mat4* transforms, const crx_render_target* target,
                const crx_tuning* tuning, const uint64_t* nontrivial, uint64_t count, int nthreads,
                std::vector<crx_huge_triangle>& out) {
    const Target tg = make_target(target);
    const crx_tuning tu = *tuning;
    const int nt = num_threads(nthreads);
    std::vector<std::vector<crx_huge_triangle>> queues(nt);
#pragma omp parallel for schedule(dynamic, 64) num_threads(nt)
    for (int64_t i = 0; i < (int64_t)count; i++) {
#ifdef _OPENMP
        const int tid = omp_get_thread_num();
#else
        const int tid = 0;
#endif
        Counters cnt;
        stage2_entry(tg, tu, instances, transforms, nontrivial[i], queues[tid], cnt);
    }
    for (auto& q : queues) out.insert(out.end(), q.begin(), q.end());
}

}  // namespace

extern "C" {

uint64_t crx_oracle_stage1(const crx_cmesh* meshes, uint32_t numMeshes,
                           const crx_cmesh* instances, uint32_t numInstances,
                           const crx_mat4* transforms, const crx_render_target* target,
                           int instanced, const crx_tuning* tuning, int nthreads,
                           uint64_t* nontrivial, uint64_t capacity) {
    std::vector<uint64_t> q;
    stage1_run(meshes, numMeshes, instances, numInstances, transforms, target, instanced, tuning, nthreads, q);
    if (nontrivial) std::copy(q.begin(), q.begin() + std::min<uint64_t>(capacity, q.size()), nontrivial);
    return q.size();
}

uint64_t crx_oracle_stage2(const crx_cmesh* instances, const crx_mat4* transforms,
                           const crx_render_target* target, const crx_tuning* tuning,
                           const uint64_t* nontrivial, uint64_t count, int nthreads,
                           crx_huge_triangle* huge, uint64_t capacity) {
    std::vector<crx_huge_triangle> q;
    stage2_run(instances, transforms, target, tuning, nontrivial, count, nthreads, q);
    if (huge) std::copy(q.begin(), q.begin() + std::min<uint64_t>(capacity, q.size()), huge);
    return q.size();
}

void crx_oracle_stage3(const crx_cmesh* instances, const crx_mat4* transforms,
                       const crx_render_target* target, const crx_tuning* tuning,
                       const crx_huge_triangle* huge, uint64_t count, int nthreads) {
    const Target tg = make_target(target);
    const crx_tuning tu = *tuning;
    const int nt = num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 4) num_threads(nt)
    for (int64_t i = 0; i < (int64_t)count; i++) {
        Counters cnt;
        stage3_entry(tg, tu, instances, transforms, huge[i], cnt);
    }
}

int crx_oracle_draw(const crx_cmesh* meshes, uint32_t numMeshes,
                    const crx_cmesh* instances, uint32_t numInstances,
                    const crx_mat4* transforms, const crx_render_target* target,
                    int instanced, const crx_tuning* tuning, int nthreads,
                    crx_oracle_stats* stats) {
    if (!target || !tuning || target->width <= 0 || target->height <= 0) return -1;
    if (tuning->tile_size == 0) return -1;
    if (numInstances == 0) return 0;
    std::vector<uint64_t> q1;
    stage1_run(meshes, numMeshes, instances, numInstances, transforms, target, instanced, tuning, nthreads, q1);
    std::vector<crx_huge_triangle> q2;
    stage2_run(instances, transforms, target, tuning, q1.data(), q1.size(), nthreads, q2);
    crx_oracle_stage3(instances, transforms, target, tuning, q2.data(), q2.size(), nthreads);
    if (stats) {
        uint64_t tris = 0;
        if (instanced) {
            for (uint32_t k = 0; k < numMeshes; k++) tris += (uint64_t)meshes[k].numTriangles * (uint64_t)meshes[k].instances.count;
        } else {
            for (uint32_t k = 0; k < numInstances; k++) tris += instances[k].numTriangles;
        }
        stats->num_triangles = tris;
        stats->num_nontrivial = q1.size();
        stats->num_huge = q2.size();
        stats->num_stage1_px = stats->num_stage2_px = stats->num_stage3_px = 0;
    }
    return 0;
}

int crx_oracle_resolve(const crx_cmesh* meshes, uint32_t numMeshes, const uint64_t* triangleCountPrefixSum,
                       const crx_render_target* target, int attribute, int nthreads) {
    const bool texture = attribute == CRX_DISPLAY_TEXTURE;
    const int W = target->width, H = target->height;
    const M4 view = target->view;
    const M4 proj = target->proj;
    const M4 viewI = inverse(view);
    const V3 origin = xyz(mulmv(viewI, V4{0.0f, 0.0f, 0.0f, 1.0f}));
    const int nt = num_threads(nthreads);
#pragma omp parallel for schedule(static) num_threads(nt)
    for (int y = 0; y < H; y++) {
        for (int x = 0; x < W; x++) {
            int pixelID = x + W * y;
            V3 rayDir = computeRayDir(float(x) + 0.5f, float(y) + 0.5f, float(W), float(H), viewI, proj, origin);
            // the neighbours' rays enter only TEXTURE's mip level (resolve.cu:678-680)
            V3 rayDir_10{}, rayDir_01{}, rayDir_11{};
            if (texture) {
                rayDir_10 = computeRayDir(float(x) + 1.5f, float(y) + 0.5f, float(W), float(H), viewI, proj, origin);
                rayDir_01 = computeRayDir(float(x) + 0.5f, float(y) + 1.5f, float(W), float(H), viewI, proj, origin);
                rayDir_11 = computeRayDir(float(x) + 1.5f, float(y) + 1.5f, float(W), float(H), viewI, proj, origin);
            }
            uint64_t pixel = target->framebuffer[pixelID];
            uint64_t pixel_colorbuffer = target->colorbuffer[pixelID];
            float depth;
            uint64_t totalTriangleIndex = 0;
            unpack_pixel(pixel, &depth, &totalTriangleIndex);
            float depth_colorbuffer = bitsf((uint32_t)(pixel_colorbuffer >> 32));
            if (depth_colorbuffer < depth) continue;

            uint32_t meshIndex;
            {   // resolve.cu:697-721
                int left = 0;
                int right = (int)(numMeshes - 1);
                int searchResult = 0;
                while (left <= right) {
                    int mid = left + (right - left) / 2;
                    int64_t before = (int64_t)triangleCountPrefixSum[mid];
                    int64_t after = ((uint32_t)mid == (numMeshes - 1)) ? 0xffffffffffll : (int64_t)triangleCountPrefixSum[mid + 1];
                    if (totalTriangleIndex >= (uint64_t)before && totalTriangleIndex < (uint64_t)after) {
                        searchResult = mid;
                        break;
                    } else if ((uint64_t)after <= totalTriangleIndex) {
                        left = mid + 1;
                    } else {
                        right = mid - 1;
                    }
                    if (mid == left) break;
                }
                meshIndex = (uint32_t)searchResult;
            }

            uint32_t color = 0;
            if (!std::isinf(depth) && numMeshes > 0) {
                const crx_cmesh& mesh = meshes[meshIndex];
                uint32_t triangleIndex = (uint32_t)(totalTriangleIndex - mesh.cummulativeTriangleCount);
                triangleIndex = std::min(triangleIndex, mesh.numTriangles - 1);
                // normalisation: an empty mesh is shaded like an unloaded one
                // (the reference would fetch out of bounds)
                if (!mesh.isLoaded || mesh.numTriangles == 0) {
                    color = 0xffff00ff;
                } else {
                    uint32_t i0, i1, i2;
                    if (mesh.compressed) {
                        // resolve.cu:755-760 (bit width taken from the CMesh: the
                        // reference recomputes it with the formula the host used)
                        const uint64_t b = mesh.bitsPerIndex;
                        i0 = bitedit_read(mesh.indices, b * (3ull * triangleIndex + 0), b) + mesh.index_min;
                        i1 = bitedit_read(mesh.indices, b * (3ull * triangleIndex + 1), b) + mesh.index_min;
                        i2 = bitedit_read(mesh.indices, b * (3ull * triangleIndex + 2), b) + mesh.index_min;
                    } else if (mesh.indices) {
                        i0 = mesh.indices[3 * triangleIndex + 0];
                        i1 = mesh.indices[3 * triangleIndex + 1];
                        i2 = mesh.indices[3 * triangleIndex + 2];
                    } else {
                        i0 = 3 * triangleIndex + 0;
                        i1 = 3 * triangleIndex + 1;
                        i2 = 3 * triangleIndex + 2;
                    }
                    crx_vec3 pa = getVertex_resolved(mesh, i0), pb = getVertex_resolved(mesh, i1),
                             pc = getVertex_resolved(mesh, i2);
                    V3 a_world = xyz(mulmv(mesh.world, V4{pa.x, pa.y, pa.z, 1.0f}));
                    V3 b_world = xyz(mulmv(mesh.world, V4{pb.x, pb.y, pb.z, 1.0f}));
                    V3 c_world = xyz(mulmv(mesh.world, V4{pc.x, pc.y, pc.z, 1.0f}));
                    float t, t_10 = 0.0f, t_01 = 0.0f, t_11 = 0.0f;
                    {   // resolve.cu:790-801
                        V3 edge1 = sub(b_world, a_world);
                        V3 edge2 = sub(c_world, a_world);
                        V3 normal = normalize(cross(edge1, edge2));
                        float d = dot(a_world, normal);
                        t = intersectPlane(origin, rayDir, normal, -d);
                        if (texture) {
                            t_10 = intersectPlane(origin, rayDir_10, normal, -d);
                            t_01 = intersectPlane(origin, rayDir_01, normal, -d);
                            t_11 = intersectPlane(origin, rayDir_11, normal, -d);
                        }
                    }
                    V2 stv, stv_10{}, stv_01{}, stv_11{};
                    {   // resolve.cu:805-834
                        V3 v0 = sub(b_world, a_world);
                        V3 v1 = sub(c_world, a_world);
                        float d00 = dot(v0, v0);
                        float d01 = dot(v0, v1);
                        float d11 = dot(v1, v1);
                        float denom = d00 * d11 - d01 * d01;
                        float denomI = 1.0f / denom;
                        auto computeSTV = [&](float tt, V3 dir) {
                            V3 p = add(origin, mul(tt, dir));
                            V3 v2 = sub(p, a_world);
                            float d20 = dot(v2, v0);
                            float d21 = dot(v2, v1);
                            return V2{(d11 * d20 - d01 * d21) * denomI, (d00 * d21 - d01 * d20) * denomI};
                        };
                        stv = computeSTV(t, rayDir);
                        if (texture) {
                            stv_10 = computeSTV(t_10, rayDir_10);
                            stv_01 = computeSTV(t_01, rayDir_01);
                            stv_11 = computeSTV(t_11, rayDir_11);
                        }
                    }
                    color = 0xffff00ff;
                    uint32_t rgb[3];
                    bool setRGB = false;
                    if (attribute == CRX_DISPLAY_NONE) {
                        color = 0xffffffff;
                    } else if (attribute == CRX_DISPLAY_TRIANGLE_ID) {
                        color = triangleIndex * 12345678u;
                    } else if (attribute == CRX_DISPLAY_MESH_ID) {
                        color = ((uint32_t)mesh.id + 1u) * 12345678u;
                    } else if (attribute == CRX_DISPLAY_STAGE) {
                        // resolve.cu:974-1031: the stage the triangle would be drawn by,
                        // from its view-space vertices (resolve.cu:777-779)
                        V3 a_view = xyz(mulmv(view, V4{a_world.x, a_world.y, a_world.z, 1.0f}));
                        V3 b_view = xyz(mulmv(view, V4{b_world.x, b_world.y, b_world.z, 1.0f}));
                        V3 c_view = xyz(mulmv(view, V4{c_world.x, c_world.y, c_world.z, 1.0f}));
                        const float f = proj.m[5];
                        const float aspect = float(W) / float(H);
                        // viewToNDC, rasterization_helpers.cuh:95-101
                        auto viewToNDC = [&](V3 v) {
                            float depth = -v.z;
                            return V3{(f / aspect) * v.x / depth, f * v.y / depth, depth};
                        };
                        // ndcToScreen, rasterization_helpers.cuh:103-108
                        auto ndcToScreen = [&](V3 n) {
                            return V2{(n.x * 0.5f + 0.5f) * float(W), (n.y * 0.5f + 0.5f) * float(H)};
                        };
                        V3 a_ndc = viewToNDC(a_view), b_ndc = viewToNDC(b_view), c_ndc = viewToNDC(c_view);
                        bool isNontrivial = (a_ndc.z <= 0.0f || b_ndc.z <= 0.0f || c_ndc.z <= 0.0f);
                        V2 a_s = ndcToScreen(a_ndc), b_s = ndcToScreen(b_ndc), c_s = ndcToScreen(c_ndc);
                        // min / max of two floats are fminf / fmaxf in the reference's device code
                        float min_x = std::fmin(a_s.x, std::fmin(b_s.x, c_s.x));
                        float max_x = std::fmax(a_s.x, std::fmax(b_s.x, c_s.x));
                        float min_y = std::fmin(a_s.y, std::fmin(b_s.y, c_s.y));
                        float max_y = std::fmax(a_s.y, std::fmax(b_s.y, c_s.y));
                        min_x = std::fmax(min_x, 0.0f);
                        max_x = std::fmin(max_x, float(W - 1));
                        min_y = std::fmax(min_y, 0.0f);
                        max_y = std::fmin(max_y, float(H - 1));
                        int size_x = f2i(std::ceil(max_x) - std::floor(min_x));
                        int size_y = f2i(std::ceil(max_y) - std::floor(min_y));
                        // int numFragments compared with the uint32_t thresholds: unsigned
                        uint32_t numFragments = (uint32_t)size_x * (uint32_t)size_y;
                        int stage = 0;
                        if (numFragments <= 128u) stage = 0;          // THRESHOLD_SMALL
                        else if (numFragments <= 4096u) stage = 1;    // THRESHOLD_LARGE
                        else stage = 2;
                        if (isNontrivial) stage = 2;
                        // SCHEME_SPECTRAL (utils.cuh:21-33) entries 9, 7 and 3
                        static const uint32_t kSpectral[11] = {0xff42019e, 0xff4f3ed5, 0xff436df4, 0xff61aefd,
                                                               0xff8be0fe, 0xffbfffff, 0xff98f5e6, 0xffa4ddab,
                                                               0xffa5c266, 0xffbd8832, 0xffa24f5e};
                        color = stage == 0 ? kSpectral[9] : stage == 1 ? kSpectral[7] : kSpectral[3];
                    } else if (texture && mesh.uvs) {
                        // resolve.cu:836-882
                        const crx_vec2 uv_a = mesh.uvs[i0], uv_b = mesh.uvs[i1], uv_c = mesh.uvs[i2];
                        auto interp = [&](V2 b) {
                            float w = 1.0f - b.x - b.y;
                            return V2{(uv_a.x * w + uv_b.x * b.x) + uv_c.x * b.y,
                                      (uv_a.y * w + uv_b.y * b.x) + uv_c.y * b.y};
                        };
                        V2 uv = interp(stv), uv_10 = interp(stv_10), uv_01 = interp(stv_01), uv_11 = interp(stv_11);
                        // max / min of two floats are fmaxf / fminf in the reference's device code
                        V2 uvmax{std::fmax(std::fmax(uv.x, uv_10.x), std::fmax(uv_01.x, uv_11.x)),
                                 std::fmax(std::fmax(uv.y, uv_10.y), std::fmax(uv_01.y, uv_11.y))};
                        V2 uvmin{std::fmin(std::fmin(uv.x, uv_10.x), std::fmin(uv_01.x, uv_11.x)),
                                 std::fmin(std::fmin(uv.y, uv_10.y), std::fmin(uv_01.y, uv_11.y))};
                        uv.x = uv.x - std::floor(uv.x);
                        uv.y = uv.y - std::floor(uv.y);
                        const float tw = float(mesh.texture.width), th = float(mesh.texture.height);
                        V2 dx{(uvmax.x - uvmin.x) * tw, (uvmax.x - uvmin.x) * th};
                        V2 dy{(uvmax.y - uvmin.y) * tw, (uvmax.y - uvmin.y) * th};
                        int target_level = mipLevelOf(std::fmax(dx.x * dx.x + dx.y * dx.y, dy.x * dy.x + dy.y * dy.y));
                        // normalisation: without texture data every level samples 0
                        // (the reference steps the null pointer to level k and
                        // reads there)
                        if (!mesh.texture.data) target_level = 0;
                        const uint32_t* data = mesh.texture.data;
                        uint32_t width = (uint32_t)mesh.texture.width;
                        uint32_t height = (uint32_t)mesh.texture.height;
                        for (int i = 1; i <= target_level; i++) {
                            data = data + (size_t)width * height;
                            width = (width + 2 - 1) / 2;
                            height = (height + 2 - 1) / 2;
                        }
                        // a JPEG-coded texture stays magenta: the reference's
                        // resolve only queues its blocks for the decoder
                        if (!mesh.texture.huffmanTables) color = sampleColor_linear(data, (int)width, (int)height, uv);
                    } else if (attribute == CRX_DISPLAY_UV && mesh.uvs) {
                        crx_vec2 ua = mesh.uvs[i0], ub = mesh.uvs[i1], uc = mesh.uvs[i2];
                        float w = 1.0f - stv.x - stv.y;
                        float ux = (ua.x * w + ub.x * stv.x) + uc.x * stv.y;
                        float uy = (ua.y * w + ub.y * stv.x) + uc.y * stv.y;
                        ux = ux - std::floor(ux);
                        uy = uy - std::floor(uy);
                        rgb[0] = f2u8(std::fmod(ux - std::floor(ux), 1.0f) * 256.0f);
                        rgb[1] = f2u8(std::fmod(uy - std::floor(uy), 1.0f) * 256.0f);
                        rgb[2] = 0;
                        setRGB = true;
                    } else if (attribute == CRX_DISPLAY_NORMAL && mesh.normals) {
                        // resolve.cu:927-943
                        crx_vec3 q0 = mesh.normals[i0], q1 = mesh.normals[i1], q2 = mesh.normals[i2];
                        V3 n0 = xyz(mulmv(mesh.world, V4{q0.x, q0.y, q0.z, 0.0f}));
                        V3 n1 = xyz(mulmv(mesh.world, V4{q1.x, q1.y, q1.z, 0.0f}));
                        V3 n2 = xyz(mulmv(mesh.world, V4{q2.x, q2.y, q2.z, 0.0f}));
                        float w = 1.0f - stv.x - stv.y;
                        V3 N = add(add(mul(n0, w), mul(n1, stv.x)), mul(n2, stv.y));
                        N = normalize(N);
                        rgb[0] = f2u8(clampf(N.x * 255.0f, 0.0f, 255.0f));
                        rgb[1] = f2u8(clampf(N.y * 255.0f, 0.0f, 255.0f));
                        rgb[2] = f2u8(clampf(N.z * 255.0f, 0.0f, 255.0f));
                        setRGB = true;
                    } else if (attribute == CRX_DISPLAY_NORMAL && !mesh.normals) {
                        V3 N = normalize(cross(sub(b_world, a_world), sub(c_world, a_world)));
                        rgb[0] = f2u8(clampf(N.x * 255.0f, 0.0f, 255.0f));
                        rgb[1] = f2u8(clampf(N.y * 255.0f, 0.0f, 255.0f));
                        rgb[2] = f2u8(clampf(N.z * 255.0f, 0.0f, 255.0f));
                        setRGB = true;
                    } else if (attribute == CRX_DISPLAY_VERTEX_COLORS && mesh.colors) {
                        uint32_t C_a = mesh.colors[i0], C_b = mesh.colors[i1], C_c = mesh.colors[i2];
                        color = C_a;
                        float w = 1.0f - stv.x - stv.y;
                        float ch[3];
                        for (int k = 0; k < 3; k++) {
                            float ca = (float)((C_a >> (8 * k)) & 0xff);
                            float cb = (float)((C_b >> (8 * k)) & 0xff);
                            float cc = (float)((C_c >> (8 * k)) & 0xff);
                            ch[k] = (ca * w + cb * stv.x) + cc * stv.y;
                            rgb[k] = f2u8(clampf(ch[k], 0.0f, 255.0f));
                        }
                        setRGB = true;
                    } else {
                        color = 0xff0000ff;
                    }
                    if (setRGB) color = (color & 0xff000000u) | rgb[0] | (rgb[1] << 8) | (rgb[2] << 16);
                    color = color | 0xff000000u;
                }
            }
            if (depth != kInf) {
                target->colorbuffer[pixelID] = ((uint64_t)fbits(depth) << 32) | color;
            } else {
                target->colorbuffer[pixelID] = ((uint64_t)0x7f800000u << 32) | 0u;
            }
        }
    }
    return 0;
}

uint32_t crx_oracle_prepare_frame(const crx_cmesh* node_meshes, const crx_mat4* node_world,
                                  uint32_t numNodes, const crx_view* view,
                                  int frustum_culling, int disable_instancing,
                                  crx_cmesh* meshes_unique, uint32_t* numUnique,
                                  crx_cmesh* instances, crx_mat4* transforms,
                                  uint64_t* prefix) {
    // SNTriangles::update (SNTriangles.h:29-105): four normalised side planes of
    // proj * view, AABB of the node in world space, "positive vertex" test.
    const M4 viewProj = mulmm(view->proj, view->view);
    auto row = [&](int r) { return V4{viewProj.m[r], viewProj.m[4 + r], viewProj.m[8 + r], viewProj.m[12 + r]}; };
    V4 r0 = row(0), r1 = row(1), r3 = row(3);
    V4 P[4] = {{r3.x + r0.x, r3.y + r0.y, r3.z + r0.z, r3.w + r0.w},
               {r3.x - r0.x, r3.y - r0.y, r3.z - r0.z, r3.w - r0.w},
               {r3.x + r1.x, r3.y + r1.y, r3.z + r1.z, r3.w + r1.w},
               {r3.x - r1.x, r3.y - r1.y, r3.z - r1.z, r3.w - r1.w}};
    V3 pn[4];
    float pd[4];
    for (int i = 0; i < 4; i++) {
        V3 n{P[i].x, P[i].y, P[i].z};
        float len = std::sqrt(dot(n, n));
        pn[i] = V3{n.x / len, n.y / len, n.z / len};
        pd[i] = P[i].w / len;
    }
    std::vector<uint32_t> visible;
    for (uint32_t i = 0; i < numNodes; i++) {
        const crx_cmesh& nm = node_meshes[i];
        bool inside = true;
        if (frustum_culling) {
            // Box3::transform (HostDeviceInterface.h:84-104)
            const crx_box3 bx = nm.aabb;
            V3 corners[8] = {{bx.min.x, bx.min.y, bx.min.z}, {bx.max.x, bx.min.y, bx.min.z},
                             {bx.min.x, bx.max.y, bx.min.z}, {bx.max.x, bx.max.y, bx.min.z},
                             {bx.min.x, bx.min.y, bx.max.z}, {bx.max.x, bx.min.y, bx.max.z},
                             {bx.min.x, bx.max.y, bx.max.z}, {bx.max.x, bx.max.y, bx.max.z}};
            V3 wmin{kInf, kInf, kInf}, wmax{-kInf, -kInf, -kInf};
            for (auto& c : corners) {
                V3 w = xyz(mulmv(node_world[i], V4{c.x, c.y, c.z, 1.0f}));
                wmin.x = std::fmin(wmin.x, w.x); wmin.y = std::fmin(wmin.y, w.y); wmin.z = std::fmin(wmin.z, w.z);
                wmax.x = std::fmax(wmax.x, w.x); wmax.y = std::fmax(wmax.y, w.y); wmax.z = std::fmax(wmax.z, w.z);
            }
            for (int p = 0; p < 4; p++) {
                V3 n = pn[p];
                V3 v{(n.x >= 0.0f) ? wmax.x : wmin.x, (n.y >= 0.0f) ? wmax.y : wmin.y, (n.z >= 0.0f) ? wmax.z : wmin.z};
                if (dot(n, v) + pd[p] < 0.0f) { inside = false; break; }
            }
        }
        if (inside && nm.isLoaded) visible.push_back(i);
    }
    // CuRast_render.h:338-345: numTriangles descending, then mesh address; ties
    // (instances of one mesh) keep node order.
    std::stable_sort(visible.begin(), visible.end(), [&](uint32_t a, uint32_t b) {
        const crx_cmesh& A = node_meshes[a];
        const crx_cmesh& B = node_meshes[b];
        if (A.numTriangles == B.numTriangles) return A.address < B.address;
        return A.numTriangles > B.numTriangles;
    });
    // CuRast_render.h:356-435
    uint64_t sum = 0;
    uint32_t u = 0;
    int64_t lastUnique = -1;
    for (uint32_t i = 0; i < (uint32_t)visible.size(); i++) {
        uint32_t nidx = visible[i];
        crx_cmesh cm = node_meshes[nidx];
        cm.world = node_world[nidx];
        V3 c0{cm.world.m[0], cm.world.m[1], cm.world.m[2]};
        V3 c1{cm.world.m[4], cm.world.m[5], cm.world.m[6]};
        V3 c2{cm.world.m[8], cm.world.m[9], cm.world.m[10]};
        cm.flipTriangles = dot(cross(c0, c1), c2) < 0.0f;
        cm.cummulativeTriangleCount = sum;
        cm.instances.offset = (int32_t)i;
        transforms[i] = mulmm(view->view, cm.world);
        prefix[i] = sum;
        sum += cm.numTriangles;
        instances[i] = cm;
        if (lastUnique < 0 || meshes_unique[lastUnique].address != cm.address || disable_instancing) {
            meshes_unique[u] = cm;
            meshes_unique[u].instances.count = 0;
            lastUnique = u;
            u++;
        }
        meshes_unique[lastUnique].instances.count++;
    }
    *numUnique = u;
    return (uint32_t)visible.size();
}

uint64_t crx_oracle_pack_pixel(float depth, uint64_t trianglePrefix) { return pack_pixel(depth, trianglePrefix); }
void crx_oracle_unpack_pixel(uint64_t packed, float* depth, uint64_t* trianglePrefix) { unpack_pixel(packed, depth, trianglePrefix); }

void crx_oracle_transform(const crx_mat4* worldView, const float* p, float f, float aspect,
                          float* view_out, float* ndc_out) {
    V3 v = xyz(mulmv(*worldView, V4{p[0], p[1], p[2], 1.0f}));
    V3 n = viewToNDC(v, f, aspect);
    view_out[0] = v.x; view_out[1] = v.y; view_out[2] = v.z;
    ndc_out[0] = n.x; ndc_out[1] = n.y; ndc_out[2] = n.z;
}

void crx_oracle_mat4_mul(const crx_mat4* a, const crx_mat4* b, crx_mat4* out) { *out = mulmm(*a, *b); }
void crx_oracle_mat4_inverse(const crx_mat4* m, crx_mat4* out) { *out = inverse(*m); }
void crx_oracle_mat4_vec4(const crx_mat4* m, const float* v, float* out) {
    V4 r = mulmv(*m, V4{v[0], v[1], v[2], v[3]});
    out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = r.w;
}

void crx_oracle_stage2_bbox(const float* a, const float* b, const float* c,
                            float f, float aspect, float width, float height, float* box4) {
    V3 av{a[0], a[1], a[2]}, bv{b[0], b[1], b[2]}, cv{c[0], c[1], c[2]};
    float mnx = NAN, mxx = NAN, mny = NAN, mxy = NAN;
    stage2_bbox(viewToNDC(av, f, aspect), viewToNDC(bv, f, aspect), viewToNDC(cv, f, aspect),
                av, bv, cv, f, aspect, width, height, &mnx, &mxx, &mny, &mxy);
    box4[0] = mnx; box4[1] = mxx; box4[2] = mny; box4[3] = mxy;
}

float crx_oracle_intersect_mt(const float* dir, const float* v0, const float* v1, const float* v2) {
    return intersect_mt(V3{dir[0], dir[1], dir[2]}, V3{v0[0], v0[1], v0[2]},
                        V3{v1[0], v1[1], v1[2]}, V3{v2[0], v2[1], v2[2]});
}

uint32_t crx_oracle_bitedit_read(const uint32_t* data, uint64_t bitOffset, uint64_t bitCount) {
    return bitedit_read(data, bitOffset, bitCount);
}
void crx_oracle_bitedit_write(uint32_t* data, uint64_t bitOffset, uint64_t bitCount, uint32_t value) {
    bitedit_write(data, bitOffset, bitCount, value);
}

}  // extern "C"
