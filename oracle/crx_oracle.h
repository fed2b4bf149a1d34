/*
 * crx_oracle.h — CPU oracle for the visibility-buffer path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or as the timed CPU baseline).  The product library never links it.
 *
 * A scalar C++17 restatement of CuRast's triangle visibility-buffer kernels
 * (src/kernels/triangles_visbuffer.cu, rasterization_helpers.cuh, resolve.cu)
 * and of the frame preparation in CuRast::draw (src/CuRast_render.h:266-519),
 * compiled with -ffp-contract=off: IEEE binary32, correctly rounded division
 * and sqrt, GLM 0.9.9 evaluation order.  Structures are the ones of crx.h, so
 * pointer members of crx_cmesh / crx_render_target are HOST pointers here.
 * OpenMP parallelises over triangles; the framebuffer is updated with a
 * compare-exchange minimum, which is order independent, so every thread count
 * gives the same bits.
 */
#ifndef CRX_ORACLE_H
#define CRX_ORACLE_H

#include "../include/crx.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct crx_oracle_stats {
    uint64_t num_triangles;    /* triangle-instances visited by stage 1 */
    uint64_t num_nontrivial;   /* stage-2 queue entries */
    uint64_t num_huge;         /* stage-3 tile entries */
    uint64_t num_stage1_px;    /* atomicMin calls per stage */
    uint64_t num_stage2_px;
    uint64_t num_stage3_px;
} crx_oracle_stats;

/* kernel_clearFramebuffer, resolve.cu:213-232 */
void crx_oracle_clear(uint64_t* fb, uint64_t numPixels);

/* drawTrianglesVisbuffer: stage 1 over all triangles, then stage 2 over the
 * queued ones, then stage 3 over the queued tiles.  Returns 0 on success. */
int crx_oracle_draw(const crx_cmesh* meshes, uint32_t numMeshes,
                    const crx_cmesh* instances, uint32_t numInstances,
                    const crx_mat4* transforms, const crx_render_target* target,
                    int instanced, const crx_tuning* tuning, int nthreads,
                    crx_oracle_stats* stats);

/* Stage entry points for queue-level checks.  Queues are caller arrays with a
 * capacity; the functions return the number of entries produced (which may
 * exceed the capacity, in which case the tail is not stored). */
uint64_t crx_oracle_stage1(const crx_cmesh* meshes, uint32_t numMeshes,
                           const crx_cmesh* instances, uint32_t numInstances,
                           const crx_mat4* transforms, const crx_render_target* target,
                           int instanced, const crx_tuning* tuning, int nthreads,
                           uint64_t* nontrivial, uint64_t capacity);
uint64_t crx_oracle_stage2(const crx_cmesh* instances, const crx_mat4* transforms,
                           const crx_render_target* target, const crx_tuning* tuning,
                           const uint64_t* nontrivial, uint64_t count, int nthreads,
                           crx_huge_triangle* huge, uint64_t capacity);
void crx_oracle_stage3(const crx_cmesh* instances, const crx_mat4* transforms,
                       const crx_render_target* target, const crx_tuning* tuning,
                       const crx_huge_triangle* huge, uint64_t count, int nthreads);

/* computeMipMap's kernel (textureTools.cu:41-133) over the chain `data`
 * whose level 0 is in place: levels 1 .. max(L - 1, 7), L = floor(log2(max(w,
 * h))).  Returns the last level written. */
int crx_oracle_compute_mipmap(uint32_t* data, int32_t width, int32_t height);

/* kernel_resolve_visbuffer_to_colorbuffer2D, resolve.cu:640-1141 */
int crx_oracle_resolve(const crx_cmesh* instances, uint32_t numInstances,
                       const uint64_t* prefix, const crx_render_target* target,
                       int attribute, int nthreads);

/* Frame preparation of CuRast::draw (CuRast_render.h:266-519) for a list of
 * nodes.  node_meshes[i] describes node i's mesh (numTriangles, buffers, aabb,
 * address used for grouping), node_world[i] its global transform.  Outputs
 * (capacity numNodes each): unique meshes, all visible instances, transforms,
 * prefix.  Returns the number of visible instances; *numUnique receives the
 * number of unique meshes. */
uint32_t crx_oracle_prepare_frame(const crx_cmesh* node_meshes, const crx_mat4* node_world,
                                  uint32_t numNodes, const crx_view* view,
                                  int frustum_culling, int disable_instancing,
                                  crx_cmesh* meshes_unique, uint32_t* numUnique,
                                  crx_cmesh* instances, crx_mat4* transforms,
                                  uint64_t* prefix);

/* Building blocks exposed for unit tests. */
uint64_t crx_oracle_pack_pixel(float depth, uint64_t trianglePrefix);
void crx_oracle_unpack_pixel(uint64_t packed, float* depth, uint64_t* trianglePrefix);
/* worldView * (x,y,z,1) -> view xyz, and viewToNDC -> ndc (x, y, depth). */
void crx_oracle_transform(const crx_mat4* worldView, const float* xyz, float f, float aspect,
                          float* view_out, float* ndc_out);
void crx_oracle_mat4_mul(const crx_mat4* a, const crx_mat4* b, crx_mat4* out);
void crx_oracle_mat4_inverse(const crx_mat4* m, crx_mat4* out);
void crx_oracle_mat4_vec4(const crx_mat4* m, const float* v, float* out);
/* Screen-space bounding box of stage 2 (rasterization_helpers.cuh:350-460). */
void crx_oracle_stage2_bbox(const float* a_view, const float* b_view, const float* c_view,
                            float f, float aspect, float width, float height, float* box4);
/* Moller-Trumbore of stage 3 (rasterization_helpers.cuh:213-270), double-sided. */
float crx_oracle_intersect_mt(const float* dir, const float* v0, const float* v1, const float* v2);
/* BitEdit::readU32 / writeU32 (src/BitEdit.h:8-49). */
uint32_t crx_oracle_bitedit_read(const uint32_t* data, uint64_t bitOffset, uint64_t bitCount);
void crx_oracle_bitedit_write(uint32_t* data, uint64_t bitOffset, uint64_t bitCount, uint32_t value);

#ifdef __cplusplus
}
#endif

#endif
