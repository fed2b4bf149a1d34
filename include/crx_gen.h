/*
 * crx_gen.h — seeded synthetic workloads for the visibility-buffer path.
 *
 * No counterpart in the reference: CuRast renders glTF / PLY datasets loaded
 * from disk (src/LargeGlbLoader.h, src/PlyLoader.h).  These generators make the
 * meshes, instance sets and camera sets that stand in for them, bit-for-bit
 * reproducibly (SplitMix64 streams, double-precision evaluation rounded once
 * to float).  All arrays are host memory owned by the caller.
 */
#ifndef CRX_GEN_H
#define CRX_GEN_H

#include <stdint.h>
#include <stddef.h>

#include "crx.h"

#ifdef __cplusplus
extern "C" {
#endif

/* i-th output of the SplitMix64 stream started at `seed`. */
uint64_t crx_gen_splitmix64(uint64_t seed, uint64_t i);

/* 64-bit hash of a byte range (word-wise multiply-xorshift), used to pin
 * generated arrays and rendered buffers in fixtures. */
uint64_t crx_gen_hash64(const void* data, size_t bytes);

/* Displaced lat-long sphere: (rings+1)*segments vertices (longitude wraps),
 * 2*rings*segments triangles, CCW seen from outside, radius
 * R*(1 + noise*U[-1,1]) per vertex, RGBA8 colours from the same seed.
 * positions: 3*V floats; indices: 3*T; colors: V (nullable). */
void crx_gen_sphere(uint32_t rings, uint32_t segments, float radius, float noise, uint64_t seed,
                    float* positions, uint32_t* indices, uint32_t* colors);

/* n x n regular grid heightfield, unit `spacing`, vertex (row r, col c) at
 * (c*spacing, h, r*spacing) with h = amplitude * sum_{o<octaves} 0.5^o *
 * (valueNoise(2^o * (c, r) / scale) - 0.5), value noise over a seeded
 * lattice with smoothstep interpolation.  2*(n-1)^2 triangles facing +y.
 * colors (nullable) shade by height.  nthreads <= 0 uses all cores. */
void crx_gen_heightfield(uint32_t n, float spacing, float amplitude, float scale, uint32_t octaves,
                         uint64_t seed, float* positions, uint32_t* indices, uint32_t* colors,
                         int nthreads);

/* Tile of the same heightfield: the rows x cols vertices starting at grid
 * row r0, column c0, bit-identical to those of crx_gen_heightfield(n, ...),
 * stored tile-locally (vertex (lr, lc) at lr*cols + lc), and the
 * 2*(rows-1)*(cols-1) triangles of the tile's quads with tile-local indices
 * (same orientation).  Neighbouring tiles that share a row or column of
 * vertices stitch without cracks.  Any output may be NULL. */
void crx_gen_heightfield_tile(uint32_t n, uint32_t r0, uint32_t c0, uint32_t rows, uint32_t cols, float spacing,
                              float amplitude, float scale, uint32_t octaves, uint64_t seed, float* positions,
                              uint32_t* indices, uint32_t* colors, int nthreads);

/* Index buffer of a sub-grid of a heightfield whose vertex rows hold n
 * vertices: quads with rows in [r0, r1) and columns in [c0, c1), c1 <= n-1,
 * 2 triangles each, indices r*n + c into the vertex array.  Returns the
 * number of triangles written. */
uint64_t crx_gen_heightfield_indices(uint32_t n, uint32_t r0, uint32_t r1, uint32_t c0, uint32_t c1,
                                     uint32_t* indices);

/* `count` world transforms T*R*S: translation uniform in [0, extent)^3,
 * uniformly random rotation (Shoemake quaternion), uniform scale in
 * [smin, smax]. */
void crx_gen_instances(uint32_t count, float extent, float smin, float smax, uint64_t seed, crx_mat4* out);

#ifdef __cplusplus
}
#endif

#endif
