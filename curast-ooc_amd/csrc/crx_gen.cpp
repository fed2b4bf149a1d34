// crx_gen.cpp — seeded synthetic workloads (include/crx_gen.h).
#include "crx_gen.h"

#include <cmath>
#include <cstring>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;

inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

inline uint64_t splitmix_at(uint64_t seed, uint64_t i) { return mix64(seed + (i + 1) * kGolden); }

// uniform double in [0, 1) with 53 random bits
inline double unit(uint64_t r) { return (double)(r >> 11) * 0x1p-53; }

// lattice value in [0, 1) of integer point (ix, iy) at one octave
inline double lattice(uint64_t seed, int64_t ix, int64_t iy) {
    const uint64_t key = ((uint64_t)(uint32_t)ix << 32) | (uint64_t)(uint32_t)iy;
    return unit(mix64(seed ^ mix64(key + kGolden)));
}

inline double smooth(double t) { return t * t * (3.0 - 2.0 * t); }

}  // namespace

extern "C" {

uint64_t crx_gen_splitmix64(uint64_t seed, uint64_t i) { return splitmix_at(seed, i); }

uint64_t crx_gen_hash64(const void* data, size_t bytes) {
    const unsigned char* p = static_cast<const unsigned char*>(data);
    uint64_t h = 0x243F6A8885A308D3ull ^ (uint64_t)bytes;
    size_t i = 0;
    for (; i + 8 <= bytes; i += 8) {
        uint64_t w;
        std::memcpy(&w, p + i, 8);
        h = mix64(h ^ w) + kGolden;
    }
    uint64_t tail = 0;
    for (size_t k = 0; i + k < bytes; k++) tail |= (uint64_t)p[i + k] << (8 * k);
    h = mix64(h ^ tail ^ 0xFF);
    return h;
}

void crx_gen_sphere(uint32_t rings, uint32_t segments, float radius, float noise, uint64_t seed,
                    float* positions, uint32_t* indices, uint32_t* colors) {
    const double pi = 3.14159265358979323846;
    const uint64_t V = (uint64_t)(rings + 1) * segments;
    const uint64_t colorSeed = seed ^ 0xC0C0C0C0C0C0C0C0ull;
#pragma omp parallel for schedule(static)
    for (int64_t v = 0; v < (int64_t)V; v++) {
        const uint32_t r = (uint32_t)(v / segments), s = (uint32_t)(v % segments);
        const double theta = pi * (double)r / (double)rings;
        const double phi = 2.0 * pi * (double)s / (double)segments;
        const double u = 2.0 * unit(splitmix_at(seed, (uint64_t)v)) - 1.0;
        const double rad = (double)radius * (1.0 + (double)noise * u);
        if (positions) {
            positions[3 * v + 0] = (float)(rad * std::sin(theta) * std::cos(phi));
            positions[3 * v + 1] = (float)(rad * std::cos(theta));
            positions[3 * v + 2] = (float)(rad * std::sin(theta) * std::sin(phi));
        }
        if (colors) colors[v] = (uint32_t)(splitmix_at(colorSeed, (uint64_t)v) & 0x00ffffffull) | 0xff000000u;
    }
    if (!indices) return;
    const uint64_t Q = (uint64_t)rings * segments;
#pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < (int64_t)Q; q++) {
        const uint32_t r = (uint32_t)(q / segments), s = (uint32_t)(q % segments);
        const uint32_t s1 = (s + 1) % segments;
        const uint32_t v00 = r * segments + s, v01 = r * segments + s1;
        const uint32_t v10 = (r + 1) * segments + s, v11 = (r + 1) * segments + s1;
        uint32_t* t = indices + 6 * q;
        // outward normals: (v00, v11, v10) and (v00, v01, v11)
        t[0] = v00; t[1] = v11; t[2] = v10;
        t[3] = v00; t[4] = v01; t[5] = v11;
    }
}

void crx_gen_heightfield_tile(uint32_t n, uint32_t r0, uint32_t c0, uint32_t rows, uint32_t cols, float spacing,
                              float amplitude, float scale, uint32_t octaves, uint64_t seed, float* positions,
                              uint32_t* indices, uint32_t* colors, int nthreads) {
#ifdef _OPENMP
    const int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#else
    (void)nthreads;
    const int nt = 1;
#endif
    if (rows == 0 || cols == 0 || r0 >= n || c0 >= n || rows > n - r0 || cols > n - c0) return;
    if (octaves > 16) octaves = 16;
    uint64_t octSeed[16];
    for (uint32_t o = 0; o < octaves; o++) octSeed[o] = splitmix_at(seed, o);
    const uint64_t colorSeed = seed ^ 0xC0C0C0C0C0C0C0C0ull;
    if (positions || colors) {
#pragma omp parallel for schedule(dynamic, 16) num_threads(nt)
        for (int64_t lr = 0; lr < (int64_t)rows; lr++) {
            const uint64_t r = r0 + (uint64_t)lr;
            // per octave: the two lattice rows bracketing this grid row
            double fy[16];
            int64_t iy[16];
            double freq[16];
            for (uint32_t o = 0; o < octaves; o++) {
                freq[o] = std::ldexp(1.0, (int)o) / (double)scale;
                const double y = (double)r * freq[o];
                iy[o] = (int64_t)std::floor(y);
                fy[o] = smooth(y - (double)iy[o]);
            }
            int64_t curIx[16];
            double l00[16], l10[16], l01[16], l11[16];
            for (uint32_t o = 0; o < octaves; o++) curIx[o] = INT64_MIN;
            for (uint32_t lc = 0; lc < cols; lc++) {
                const uint64_t c = c0 + lc;
                double h = 0.0, amp = 1.0;
                for (uint32_t o = 0; o < octaves; o++) {
                    const double x = (double)c * freq[o];
                    const int64_t ix = (int64_t)std::floor(x);
                    if (ix != curIx[o]) {
                        curIx[o] = ix;
                        l00[o] = lattice(octSeed[o], ix, iy[o]);
                        l10[o] = lattice(octSeed[o], ix + 1, iy[o]);
                        l01[o] = lattice(octSeed[o], ix, iy[o] + 1);
                        l11[o] = lattice(octSeed[o], ix + 1, iy[o] + 1);
                    }
                    const double sx = smooth(x - (double)ix);
                    const double a = l00[o] + sx * (l10[o] - l00[o]);
                    const double b = l01[o] + sx * (l11[o] - l01[o]);
                    const double v = a + fy[o] * (b - a);
                    h += amp * (v - 0.5);
                    amp *= 0.5;
                }
                const double height = (double)amplitude * h;
                const uint64_t gv = r * n + c;                 // global vertex id (colour stream)
                const uint64_t lv = (uint64_t)lr * cols + lc;  // local vertex id
                if (positions) {
                    positions[3 * lv + 0] = (float)((double)c * (double)spacing);
                    positions[3 * lv + 1] = (float)height;
                    positions[3 * lv + 2] = (float)((double)r * (double)spacing);
                }
                if (colors) {
                    double t = 0.5 + h;   // h in about [-1, 1]
                    t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
                    const uint32_t rr = (uint32_t)(40.0 + 180.0 * t);
                    const uint32_t gg = (uint32_t)(200.0 - 120.0 * t);
                    const uint32_t bb = (uint32_t)(splitmix_at(colorSeed, gv) & 0x3f) + 40u;
                    colors[lv] = 0xff000000u | (bb << 16) | (gg << 8) | rr;
                }
            }
        }
    }
    if (indices && rows >= 2 && cols >= 2) crx_gen_heightfield_indices(cols, 0, rows - 1, 0, cols - 1, indices);
}

void crx_gen_heightfield(uint32_t n, float spacing, float amplitude, float scale, uint32_t octaves,
                         uint64_t seed, float* positions, uint32_t* indices, uint32_t* colors,
                         int nthreads) {
    crx_gen_heightfield_tile(n, 0, 0, n, n, spacing, amplitude, scale, octaves, seed, positions, nullptr, colors,
                             nthreads);
    if (indices) crx_gen_heightfield_indices(n, 0, n - 1, 0, n - 1, indices);
}

uint64_t crx_gen_heightfield_indices(uint32_t n, uint32_t r0, uint32_t r1, uint32_t c0, uint32_t c1,
                                     uint32_t* indices) {
    if (n < 2 || r1 <= r0 || c1 <= c0 || c1 > n - 1) return 0;
    const uint64_t w = c1 - c0;
    const uint64_t quads = (uint64_t)(r1 - r0) * w;
#pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < (int64_t)quads; q++) {
        const uint32_t r = r0 + (uint32_t)(q / w), c = c0 + (uint32_t)(q % w);
        const uint32_t v00 = r * n + c, v01 = v00 + 1, v10 = v00 + n, v11 = v10 + 1;
        uint32_t* t = indices + 6 * q;
        // +y facing: (v00, v10, v11) and (v00, v11, v01)
        t[0] = v00; t[1] = v10; t[2] = v11;
        t[3] = v00; t[4] = v11; t[5] = v01;
    }
    return 2 * quads;
}

void crx_gen_instances(uint32_t count, float extent, float smin, float smax, uint64_t seed, crx_mat4* out) {
    const double pi = 3.14159265358979323846;
    for (uint32_t i = 0; i < count; i++) {
        double u[7];
        for (int k = 0; k < 7; k++) u[k] = unit(splitmix_at(seed, 7ull * i + k));
        const double tx = u[0] * extent, ty = u[1] * extent, tz = u[2] * extent;
        // Shoemake uniform random unit quaternion
        const double a = std::sqrt(1.0 - u[3]), b = std::sqrt(u[3]);
        const double qx = a * std::sin(2.0 * pi * u[4]), qy = a * std::cos(2.0 * pi * u[4]);
        const double qz = b * std::sin(2.0 * pi * u[5]), qw = b * std::cos(2.0 * pi * u[5]);
        const double s = smin + u[6] * (smax - smin);
        const double R[3][3] = {
            {1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qz * qw), 2 * (qx * qz + qy * qw)},
            {2 * (qx * qy + qz * qw), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qx * qw)},
            {2 * (qx * qz - qy * qw), 2 * (qy * qz + qx * qw), 1 - 2 * (qx * qx + qy * qy)}};
        float* m = out[i].m;
        for (int c = 0; c < 3; c++)
            for (int r = 0; r < 3; r++) m[4 * c + r] = (float)(R[r][c] * s);
        m[3] = m[7] = m[11] = 0.0f;
        m[12] = (float)tx; m[13] = (float)ty; m[14] = (float)tz; m[15] = 1.0f;
    }
}

}  // extern "C"
