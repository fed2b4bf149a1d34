/*
 * crx.h — C-ABI of the MI355X (gfx950) visibility-buffer rasterizer.
 *
 * This is the drop-in boundary for CuRast's triangle visibility-buffer path:
 * the three-stage `triangles_visbuffer` pipeline, the framebuffer clear and the
 * visibility-buffer resolve, plus the scene / mesh / instance / camera surface
 * that feeds them.  Every entry point is `extern "C"`, takes plain pointers and
 * sizes, returns a crx_status and never exits the process.
 *
 * Reference interfaces replaced (paths relative to the CuRast-OOC source tree):
 *   crx_cmesh              CMesh          src/kernels/HostDeviceInterface.h:196-234 (byte-compatible, 280 B)
 *   crx_render_target      RenderTarget   src/kernels/HostDeviceInterface.h:131-143 (byte-compatible, 240 B)
 *   crx_huge_triangle      HugeTriangle   src/kernels/HostDeviceInterface.h:254-259
 *   crx_tuning             TILE_SIZE / TRIANGLES_PER_SWEEP / THRESHOLD_SMALL / MAX_* constants,
 *                          src/kernels/HostDeviceInterface.h:261-266 (compile-time there, runtime here)
 *   crx_clear_framebuffer  kernel_clearFramebuffer launch, src/CuRast_render.h:491-512,
 *                          kernel src/kernels/resolve.cu:213-232
 *   crx_draw_triangles_visbuffer
 *                          drawTrianglesVisbuffer(...), src/CuRast_render.h:211-264, which launches
 *                          kernel_stage{1,2,3}_* src/kernels/triangles_visbuffer.cu:755-835
 *   crx_raster_args        RasterArgs     src/kernels/HostDeviceInterface.h:309-323 (same field order,
 *                          + two queue capacities)
 *   crx_raster_stage1/2/3  kernel_stage1_drawSmallTriangles_* / kernel_stage2_drawMediumTriangles_* /
 *                          kernel_stage3_drawHugeTriangles_*, src/kernels/triangles_visbuffer.cu:755-835,
 *                          as launched by name in src/CuRast_render.h:235-264
 *   crx_resolve_visbuffer  kernel_resolve_visbuffer_to_colorbuffer2D launch, src/CuRast_render.h:580-592,
 *                          kernel src/kernels/resolve.cu:640-1141
 *   crx_scene_* / crx_mesh_* / crx_render
 *                          Scene / SNTriangles / Mesh and the frame preparation of CuRast::draw,
 *                          src/scene/Scene.h:9-51, src/scene/SNTriangles.h:18-105, src/scene/Mesh.h:4-25,
 *                          src/CuRast_render.h:266-519
 *   crx_camera_projection  Camera::createProjectionMatrix, src/VKRenderer.h:141-148
 *
 * Arithmetic contract.  All device arithmetic is IEEE-754 binary32 with
 * round-to-nearest, correctly rounded division and square root, denormals
 * preserved and no contraction of a*b+c into fused multiply-adds.  Operations
 * follow GLM 0.9.9's evaluation order (mat4*vec4 = (c0*x + c1*y) + (c2*z + c3*w),
 * dot3 = (x+y)+z, cross, normalize = v * (1/sqrt(dot(v,v)))).  Where the
 * reference uses __fdividef or fast-math division, the contract uses the IEEE
 * quotient.  Under this contract the visibility buffer is a pure function of the
 * inputs and the crx_tuning routing parameters (small_threshold, tile_size);
 * scheduling, grid sizes and queue order cannot change a single bit because
 * every write is a 64-bit atomic minimum.
 *
 * Deliberate normalisations of reference behaviour (every other quirk is kept):
 *   - stage 2 and stage 3 discard samples outside [0,W)x[0,H) instead of
 *     writing them to a wrapped or negative framebuffer index
 *     (triangles_visbuffer.cu:548-549 and :694-731);
 *   - queues are bounded; on overflow the draw grows its queues and redraws the
 *     frame, so no triangle is ever dropped (triangles_visbuffer.cu:192-197, :503-511);
 *   - a mesh may hold at most 1431655765 triangles (3*T must fit in uint32);
 *   - IX_PU16: a draw's meshes are all compressed or all uncompressed (the
 *     reference picks the kernel by meshes[0] and would misread the others),
 *     compressed meshes need an index buffer, and the resolve takes the bit
 *     width from CMesh::bitsPerIndex (the reference recomputes it with the
 *     same formula the host used, CuRast_render.h:358-359).
 * Kept quirks: stage 1 never samples column W-1 or row H-1; the instanced
 * stage-1 triangle ID adds (instance*numTriangles) mod 2^32; stage 2 forms its
 * triangle ID in 32-bit arithmetic; stage 1 uses strict edge tests, stage 2
 * inclusive ones and stage 3 the Moller-Trumbore [0,1] range.
 */
#ifndef CRX_H
#define CRX_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CRX_ABI_VERSION 1

typedef enum crx_status {
    CRX_OK = 0,
    CRX_INVALID_ARG = 1,
    CRX_OUT_OF_MEMORY = 2,
    CRX_QUEUE_OVERFLOW = 3,
    CRX_HIP_ERROR = 4,
    CRX_NO_DEVICE = 5,
    CRX_UNSUPPORTED = 6,
    CRX_REJECTED = 7
} crx_status;

/* ------------------------------------------------------------------ */
/* Plain-old-data types, GLM memory layout (tightly packed floats).    */
/* ------------------------------------------------------------------ */

typedef struct crx_vec2 { float x, y; } crx_vec2;
typedef struct crx_vec3 { float x, y, z; } crx_vec3;
/* Column-major 4x4 like glm::mat4: element (col c, row r) is m[4*c + r]. */
typedef struct crx_mat4 { float m[16]; } crx_mat4;

/* Box3, src/kernels/HostDeviceInterface.h:57-105 */
typedef struct crx_box3 { crx_vec3 min, max; } crx_box3;

/* Texture, src/kernels/HostDeviceInterface.h:177-185.  data (device) holds
 * the RGBA8 mip chain, level 0 first, each next level ((w + 1) / 2) x
 * ((h + 1) / 2) texels, up to level 7 (the deepest CRX_DISPLAY_TEXTURE
 * samples).  huffmanTables != NULL marks a JPEG-coded texture, which the
 * resolve shades magenta (0xffff00ff) as the reference's resolve pass does
 * before its decoder runs; the remaining members are carried, not read. */
typedef struct crx_texture {
    int32_t width;
    int32_t height;
    uint32_t* data;
    void* huffmanTables;
    void* quanttables;
    uint32_t* mcuPositions;
    uint32_t handle;
} crx_texture;

/* CMesh, src/kernels/HostDeviceInterface.h:196-234.  Pointer members are
 * device pointers.  Only the fields named in the comments are read. */
typedef struct crx_cmesh {
    uint32_t numTriangles;              /* read */
    uint32_t* indices;                  /* read; NULL = non-indexed (vertex 3t+k); IX_PU16: bit-packed */
    crx_vec3* positions;                /* read; xyz float, 12-byte stride; IX_PU16: 3 x uint16 */
    uint64_t cummulativeTriangleCount;  /* read: ID of triangle 0 */
    crx_box3 aabb;                      /* read by the host frustum cull; IX_PU16 decode */
    crx_vec3 compressionFactor;         /* read (IX_PU16) */
    uint32_t index_min;                 /* read (IX_PU16) */
    uint32_t bitsPerIndex;              /* read (IX_PU16), 0..32 */
    struct { int32_t offset; int32_t count; } instances;  /* read */
    crx_mat4 world;                     /* read by resolve */
    int32_t id;                         /* read by resolve (MESH_ID) */
    crx_vec2* uvs;                      /* read by resolve (TEXTURE, UV) */
    crx_vec3* normals;                  /* read by resolve (NORMAL; NULL: the face normal) */
    uint32_t* colors;                   /* read by resolve (VERTEX_COLORS), RGBA8 */
    uint32_t firstTriangle;
    uint32_t numVertices;
    uint32_t index_max;
    uint64_t address;
    crx_texture texture;                /* read by resolve (TEXTURE): width, height, data, huffmanTables */
    uint8_t isLoaded;                   /* read by resolve */
    uint8_t flipTriangles;
    uint8_t compressed;                 /* read: 1 = IX_PU16 (crx_compress_mesh) */
    struct { crx_vec3* positions; uint32_t* colors; uint32_t numPoints; } impostor;
} crx_cmesh;

/* RenderTarget, src/kernels/HostDeviceInterface.h:131-143 */
typedef struct crx_render_target {
    uint64_t* framebuffer;   /* device, width*height packed depth|triangleID */
    uint64_t* colorbuffer;   /* device, width*height depth<<32 | RGBA8 */
    int32_t width;
    int32_t height;
    crx_mat4 view;
    crx_mat4 viewI;
    crx_mat4 proj;           /* proj[1][1] (m[5]) is the focal factor f */
    crx_vec3 cameraPos;
    float f;
    float aspect;
    uint8_t debug;
} crx_render_target;

/* HugeTriangle, src/kernels/HostDeviceInterface.h:254-259 */
typedef struct crx_huge_triangle {
    int32_t meshIndex;
    int32_t triangleIndex;
    int32_t tile_x;
    int32_t tile_y;
} crx_huge_triangle;

/* DisplayAttribute, src/kernels/HostDeviceInterface.h:277-286 */
typedef enum crx_display_attribute {
    CRX_DISPLAY_NONE = 0,
    CRX_DISPLAY_TEXTURE = 1,       /* bilinear sample of CMesh::texture at the mip level of the pixel's uv footprint */
    CRX_DISPLAY_UV = 2,
    CRX_DISPLAY_NORMAL = 3,
    CRX_DISPLAY_VERTEX_COLORS = 4,
    CRX_DISPLAY_TRIANGLE_ID = 5,
    CRX_DISPLAY_MESH_ID = 6,
    CRX_DISPLAY_STAGE = 7          /* the stage the triangle's size and near-plane test route it to (debug colours) */
} crx_display_attribute;

/* Runtime form of the reference's compile-time constants.
 * small_threshold and tile_size route triangles between stages and so fix the
 * bits of the result (the oracle takes the same struct); every other field only
 * changes scheduling.  crx_tuning_default() returns the reference values. */
typedef struct crx_tuning {
    uint32_t small_threshold;     /* THRESHOLD_SMALL = 128: stage-1 fragment limit */
    uint32_t tile_size;           /* TILE_SIZE = 64: stage-2 huge limit (tile^2) and stage-3 tile edge; 16..256, multiple of 16 */
    uint32_t max_nontrivial;      /* MAX_NONTRIVIAL_TRIANGLES = 5e6: initial stage-2 queue capacity */
    uint32_t max_huge;            /* MAX_HUGE_TRIANGLES = 5e6: initial stage-3 queue capacity */
    uint32_t stage1_block;        /* threads per stage-1 workgroup: 128, 256 or 512 */
    uint32_t stage1_waves_per_cu; /* stage-1 grid = CUs * this / (stage1_block/64), at most 64;
                                     0 = one workgroup per batch */
    uint32_t stage2_waves_per_cu; /* stage-2 persistent grid size in waves per CU */
    uint32_t stage3_waves_per_cu; /* stage-3 persistent grid size in waves per CU */
    uint32_t stage1_cull_order;   /* 0 = backface test before the frustum test, 1 = frustum first
                                     testing y before x, 2 = frustum first testing x before y,
                                     3 = per item: 0 for meshes drawn over >= 4 instances and for
                                     draws of several views, else 1 (all are exact early-outs;
                                     the order only changes cost) */
    uint32_t stage1_order;        /* batch-to-workgroup order: 0 = contiguous share per XCD,
                                     1 = grid-stride, 2 = contiguous share per workgroup */
    uint32_t stage1_fast_tiny;    /* 1 = decide sure tiny-triangle culls from v_rcp_f32 screen
                                     coordinates with a proven error margin before the exact
                                     path (exact either way; default 0) */
    uint32_t stage1_tris_per_lane; /* triangles each stage-1 lane takes per batch: 1, 2 or 4 (a batch
                                     is stage1_block * this consecutive triangles of one item);
                                     0 = per draw: a draw with a mesh drawn over >= 4 instances
                                     (every instanced crx_raster_stage1 draw) takes 1 and a grid
                                     of 64 waves per CU, so does a draw of several views, an
                                     IX_PU16 draw 2 and 64 waves per CU, others 2 and
                                     stage1_waves_per_cu */
    uint32_t stage1_layout;       /* draws of one view over meshes drawn over < 4 instances (IX_PU16
                                     meshes only with stage1_tris_per_lane = 2), with 2 triangles
                                     per lane, a block of 256 and at least 16 batches of 128
                                     triangles per wave of the grid:
                                     0 = lane t takes triangles t and t + stage1_block;
                                     1 = lane t takes triangles 2t and 2t + 1, and a wave whose pairs
                                     all share two vertices as the two triangles of a quad do gathers
                                     and transforms each shared vertex once; waves claim units of
                                     stage1_strip_rows batches of 128 triangles;
                                     2 (default) = 1, and a mesh stored in rows (a grid: the row
                                     length detected from the index buffer by crx_mesh_create) is
                                     walked in units of stage1_strip_rows consecutive rows of a
                                     strip of stage1_strip_batches batches of 128 triangles, so the
                                     vertex row two rows share is read once */
    uint32_t stage1_strip_rows;   /* rows (or batches) per unit of layouts 1 and 2, 1..256 (default 8) */
    uint32_t stage1_strip_batches; /* layout 2: batches of 128 triangles side by side in a strip, 1..8
                                      (default 1); a unit's wave draws each row's batches left to right */
} crx_tuning;

/* Per-draw counters read back after a draw. */
typedef struct crx_draw_stats {
    uint64_t num_triangles;        /* triangle-instances submitted to stage 1 */
    uint32_t num_nontrivial;       /* entries in the stage-2 queue */
    uint32_t num_huge;             /* tile entries in the stage-3 queue */
    uint32_t redraws;              /* redraws caused by queue growth */
    uint32_t stage1_kernel;        /* the stage-1 kernel the draw ran: 0 = k_stage1 (lanes take
                                      triangles a block apart), 1 = k_stage1_views (several views),
                                      2 = k_stage1_pairs (consecutive pairs, tuning.stage1_layout) */
} crx_draw_stats;

typedef struct crx_context crx_context;

/* ------------------------------------------------------------------ */
/* Library / context                                                   */
/* ------------------------------------------------------------------ */

int crx_abi_version(void);
const char* crx_status_string(crx_status s);
/* Message of the last failing call on this thread ("" if none). */
const char* crx_last_error(void);
void crx_tuning_default(crx_tuning* out);

/* One context per device.  Owns queues, counters and staging for the draws
 * issued through it; calls on one context are serialised by the caller. */
crx_status crx_context_create(int device, const crx_tuning* tuning, crx_context** out);
crx_status crx_context_destroy(crx_context* ctx);
crx_status crx_context_set_tuning(crx_context* ctx, const crx_tuning* tuning);
crx_status crx_context_get_tuning(const crx_context* ctx, crx_tuning* out);
/* Number of compute units of the context's device. */
crx_status crx_context_num_cus(const crx_context* ctx, int* out);

/* ------------------------------------------------------------------ */
/* Kernel-level entry points (replace the reference's launches)        */
/* `stream` is a hipStream_t (NULL = default stream).                  */
/* ------------------------------------------------------------------ */

/* kernel_clearFramebuffer (resolve.cu:213-232): fb[i] = 0xFFFFFFF0'00000000. */
crx_status crx_clear_framebuffer(uint64_t* framebuffer, uint32_t numPixels, void* stream);

/* drawTrianglesVisbuffer (CuRast_render.h:211-264).  `meshes` / `instances`
 * are HOST arrays of CMesh whose buffers are device pointers; `transforms` is a
 * HOST array of numInstances world-to-view matrices (view * world);
 * `target` is a HOST RenderTarget whose framebuffer was cleared.  With
 * instanced != 0 stage 1 walks the unique `meshes` and loops over their
 * instances{offset,count} (the _instanced kernels); otherwise it walks
 * `instances` as meshes.  Enqueues on `stream`; the call returns after
 * the stage-2 and stage-3 queue sizes are known (one small readback). */
crx_status crx_draw_triangles_visbuffer(crx_context* ctx,
                                        const crx_cmesh* meshes, uint32_t numMeshes,
                                        const crx_cmesh* instances, uint32_t numInstances,
                                        const crx_mat4* transforms,
                                        const crx_render_target* target,
                                        int instanced, void* stream,
                                        crx_draw_stats* stats /* nullable */);

/* ---- Kernel-level entry points: the three stages as separate calls over
 * device data, for callers that launch kernel_stage{1,2,3}_* themselves. ---- */

/* RasterArgs (HostDeviceInterface.h:309-323), same fields in the same order;
 * every pointer is a DEVICE address.  meshes / instances are device arrays of
 * CMesh, transforms a device array of numInstances view * world matrices;
 * the five counters and the two queues are caller-owned device memory (the
 * queues hold nontrivialCapacity / hugeCapacity entries: the two fields
 * after `target` have no counterpart in the reference, whose queues are sized
 * by MAX_NONTRIVIAL_TRIANGLES / MAX_HUGE_TRIANGLES).  `target` itself is
 * passed by value, as in RasterArgs. */
typedef struct crx_raster_args {
    const crx_cmesh* meshes;
    uint32_t numMeshes;
    const crx_cmesh* instances;
    uint32_t numInstances;
    const crx_mat4* transforms;
    uint32_t* numProcessedBatches;
    uint32_t* numProcessedBatches_nontrivial;
    crx_huge_triangle* hugeTriangles;
    uint32_t* hugeTrianglesCounter;
    uint32_t* numProcessedHugeTriangles;
    uint32_t* nontrivialTrianglesCounter;
    uint64_t* nontrivialTrianglesList;
    crx_render_target target;
    uint32_t nontrivialCapacity;
    uint32_t hugeCapacity;
} crx_raster_args;

#define CRX_RASTER_INSTANCED 1u    /* the _instanced stage-1 kernels: walk `meshes`, loop over instances{offset,count} */
#define CRX_RASTER_COMPRESSED 2u   /* the _compressed kernels: IX_PU16 meshes */

/* kernel_stage1_drawSmallTriangles_indexbuffer{_compressed,_uncompressed}{_instanced,}
 * (the flags pick the variant as the kernel name does; non-indexed meshes,
 * the _noindexbuffer kernels, are handled by every variant).  Zeroes the five
 * counters, as the reference's stage 1 does before its grid barrier, builds
 * the draw records from the device CMesh arrays on the device and launches
 * stage 1 on `stream`.  Asynchronous: it reads no device memory.  Meshes whose
 * instance range falls outside the instance list, or whose format differs
 * from the flags, draw nothing.  Entries past a queue's capacity are dropped;
 * a counter above its capacity after the stage reports that. */
crx_status crx_raster_stage1(crx_context* ctx, const crx_raster_args* args, uint32_t flags, void* stream);
/* kernel_stage2_drawMediumTriangles_* and kernel_stage3_drawHugeTriangles_*:
 * each consumes its queue on `stream`.  Like the reference's stage kernels
 * they depend only on their args: each call builds the instance and view
 * records it reads from the args' device arrays, so draws with different
 * args may interleave on one context (each with its own counters and
 * queues).  Calls on one context that use different streams are ordered on
 * the device: a call waits for the launches of the call before it. */
crx_status crx_raster_stage2(crx_context* ctx, const crx_raster_args* args, void* stream);
crx_status crx_raster_stage3(crx_context* ctx, const crx_raster_args* args, void* stream);

/* crx_raster_stage1, crx_raster_stage2 and crx_raster_stage3 of `args` in one
 * call on `stream`, allocating nothing.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, hosts that forbid hipMalloc).  The draw
 * records live in a static device array of 4 MiB per GPU, held until the
 * process ends: 344 bytes per mesh drawn over one instance (about 12,000 of
 * them), or 256 bytes per instance plus 88 per mesh with CRX_RASTER_INSTANCED
 * (about 16,000 instances of one mesh); larger args fail with
 * CRX_INVALID_ARG.  The caller owns
 * counters, queues and framebuffer as for crx_raster_stage1.  Every call on
 * one GPU shares that array: calls that overlap on one GPU (on different
 * streams, or from different threads) race; calls on one stream do not. */
crx_status crx_raster_visbuffer_noalloc(crx_context* ctx, const crx_raster_args* args, uint32_t flags,
                                        void* stream);

/* kernel_resolve_visbuffer_to_colorbuffer2D (resolve.cu:640-1141).
 * `instances` is a HOST array of CMesh, `prefix` a HOST array of the per-
 * instance cumulative triangle counts (CuRast_render.h:412-435).  Writes
 * target->colorbuffer. */
crx_status crx_resolve_visbuffer(crx_context* ctx,
                                 const crx_cmesh* instances, uint32_t numInstances,
                                 const uint64_t* prefix,
                                 const crx_render_target* target,
                                 crx_display_attribute attribute, void* stream);

/* ---- Batched multi-view (no single reference counterpart: it generalises the
 * instanced stage-1 loop of triangles_visbuffer.cu:373-379, which fetches a
 * triangle once and loops over transforms, to a loop over views). ---- */

#define CRX_MAX_VIEWS 256

/* drawTrianglesVisbuffer for numViews (1..CRX_MAX_VIEWS) render targets that
 * share one mesh / instance list, in one pass over the geometry.
 * `transforms` holds numViews * numInstances matrices, view-major:
 * transforms[v * numInstances + i] = view_v * world_i.  targets[v] gives view
 * v's cleared framebuffer, size and projection; sizes may differ per view.
 * The visibility buffer of every view is bit-identical to a single-view
 * crx_draw_triangles_visbuffer of the same lists into that target.  A
 * multi-view draw takes at most 2^24 instances.  stats->num_triangles counts
 * triangle-instances times views. */
crx_status crx_draw_triangles_visbuffer_views(crx_context* ctx,
                                              const crx_cmesh* meshes, uint32_t numMeshes,
                                              const crx_cmesh* instances, uint32_t numInstances,
                                              const crx_mat4* transforms,
                                              const crx_render_target* targets, uint32_t numViews,
                                              int instanced, void* stream,
                                              crx_draw_stats* stats /* nullable */);

/* Resolve of numViews targets sharing one instance list, in one launch. */
crx_status crx_resolve_visbuffer_views(crx_context* ctx,
                                       const crx_cmesh* instances, uint32_t numInstances,
                                       const uint64_t* prefix,
                                       const crx_render_target* targets, uint32_t numViews,
                                       crx_display_attribute attribute, void* stream);

/* Clear every target's framebuffer (and, if colorbuffers != 0, colorbuffer)
 * to 0xFFFFFFF0'00000000 in one launch. */
crx_status crx_clear_views(crx_context* ctx, const crx_render_target* targets, uint32_t numViews,
                           int colorbuffers, void* stream);

/* Per-stage device time of the last draw / resolve issued on a context with
 * timing enabled (hipEvent pairs around each launch, the role of the
 * reference's Timer, src/Timer.h:61-113).  Milliseconds; -1 if not measured. */
typedef struct crx_timings {
    float clear_ms;
    float stage1_ms;
    float stage2_ms;
    float stage3_ms;
    float resolve_ms;
    float pipeline_ms;   /* first stage-1 event to last stage-3 event ("<triangles visbuffer pipeline>") */
} crx_timings;

crx_status crx_context_enable_timing(crx_context* ctx, int enable);
/* Synchronises on the recorded events and returns their durations. */
crx_status crx_context_timings(crx_context* ctx, crx_timings* out);

/* ------------------------------------------------------------------ */
/* Device memory plumbing for C hosts (hipMalloc / hipMemcpy wrappers). */
/* ------------------------------------------------------------------ */

crx_status crx_device_alloc(crx_context* ctx, size_t bytes, void** out);
crx_status crx_device_free(crx_context* ctx, void* ptr);
crx_status crx_memcpy_htod(crx_context* ctx, void* dst, const void* src, size_t bytes, void* stream);
crx_status crx_memcpy_dtoh(crx_context* ctx, void* dst, const void* src, size_t bytes, void* stream);
crx_status crx_stream_synchronize(crx_context* ctx, void* stream);
/* Plain device-to-device copy kernel (16 bytes per lane, nontemporal).  No counterpart in
 * the reference: it is the known-size kernel that calibrates HBM traffic
 * counters (FETCH_SIZE / WRITE_SIZE) for the roofline measurement. */
crx_status crx_copy_buffer(crx_context* ctx, void* dst, const void* src, size_t bytes, void* stream);
/* Check of stage 1's short reciprocal and quotient sequences against the
 * IEEE division on the device (no counterpart in the reference: stage 1
 * divides with v_rcp_f32 refined by FMAs, and this shows the results are the
 * IEEE ones).  Every float b with 2^-125 <= |b| < 2^125 is checked for the
 * reciprocal; `pairs` hashed operand pairs with exponents in [-60, 63] for
 * the quotient, and as many with a numerator below 2^-60 for the screen
 * coordinate term q * 0.5 + 0.5.  The three mismatch counts are returned.
 * Blocking. */
crx_status crx_check_division(crx_context* ctx, uint64_t pairs, uint64_t* rcpMismatches, uint64_t* divMismatches,
                              uint64_t* smallNumeratorMismatches);
/* Streaming read of `bytes` (16-byte aligned) of device memory with
 * nontemporal loads (no counterpart in the reference): the measured read
 * bandwidth against which stage 1's traffic is read.  sink is a device word
 * the kernel writes only for one fixed XOR of the data. */
crx_status crx_probe_read(crx_context* ctx, const void* src, size_t bytes, uint32_t* sink, void* stream);
/* 64-bit atomic-minimum probe (no counterpart in the reference): every lane
 * of a CUs * 8 x 256 grid performs atomicsPerLane atomicMin on the device
 * buffer of 2^logEntries entries, scattered (hashed positions) or coalesced
 * (a wave's lanes on 64 consecutive entries).  *issued receives the number
 * of atomics enqueued; the caller times the stream.  The measured peak
 * against which the stages' atomic rates are read. */
crx_status crx_probe_atomic_min(crx_context* ctx, uint64_t* buffer, uint32_t logEntries, uint32_t atomicsPerLane,
                                int scattered, uint64_t* issued, void* stream);
/* VALU issue-rate probe (no counterpart in the reference): a CUs * 8 x 256
 * grid whose lanes each run `iters` rounds of 8 independent dependency
 * chains, mode 0 of v_fma_f32 and mode 1 of 32-bit integer add and xor;
 * modes 2 and 3 run the fma chains with 32 lanes or 1 lane of each wave
 * active; mode 4 interleaves 4 fma and 4 integer chains in every wave and
 * mode 5 runs fma chains on even waves and integer chains on odd ones.
 * *waves receives the number of waves launched; the caller times the stream
 * and reads the instruction count from the SQ_INSTS_VALU counter.  The
 * measured issue peak against which the instanced stage 1 (an
 * ALU-bound kernel) is read.  sink is a device word written only for a
 * result no input produces. */
crx_status crx_probe_valu(crx_context* ctx, uint32_t iters, int mode, uint32_t* sink, uint64_t* waves, void* stream);

/* ------------------------------------------------------------------ */
/* Camera helpers                                                       */
/* ------------------------------------------------------------------ */

/* Camera::createProjectionMatrix (VKRenderer.h:141-148): infinite far plane,
 * f = 1/tan(fovy/2) in float; fovy in radians. */
void crx_camera_projection(float near_, float fovy, float aspect, crx_mat4* out);
/* GLM lookAt (right-handed), float arithmetic. */
void crx_camera_look_at(const float eye[3], const float center[3], const float up[3], crx_mat4* out);
/* GLM inverse / product in GLM's evaluation order (float). */
void crx_mat4_inverse(const crx_mat4* m, crx_mat4* out);
void crx_mat4_mul(const crx_mat4* a, const crx_mat4* b, crx_mat4* out);

/* ------------------------------------------------------------------ */
/* Scene API (replaces Scene / SNTriangles / Mesh and CuRast::draw)     */
/* ------------------------------------------------------------------ */

typedef struct crx_mesh crx_mesh;
typedef struct crx_scene crx_scene;

/* Mesh (Mesh.h:4-25).  Copies host arrays into device memory owned by the
 * mesh: positions (3*numVertices floats), indices (3*numTriangles, or NULL for
 * a non-indexed mesh with numVertices == 3*numTriangles), colors (numVertices
 * RGBA8 words, nullable).  The AABB is computed from the positions. */
crx_status crx_mesh_create(crx_context* ctx, const float* positions, uint32_t numVertices,
                           const uint32_t* indices, uint32_t numTriangles,
                           const uint32_t* colors, crx_mesh** out);
/* IX_PU16 compression of an indexed mesh, as CuRast's loader does it
 * (LargeGlbLoader.h:482-504 for indices, :640-680 for positions):
 * indices become index - index_min packed at bitsPerIndex =
 * ceil(log2f(index_max - index_min + 1)) bits each (BitEdit layout, word
 * count rounded up to whole 32-bit words); each position becomes 3 x uint16
 * clamp(65536 * (p - aabb.min) / (aabb.max - aabb.min), 0, 65535) evaluated in
 * double (0 on a flat axis).  qpositions holds 3*numVertices uint16,
 * packedIndices at least 3*numTriangles + 1 words; *packedWords receives the
 * words used.  meta receives numTriangles, numVertices, aabb,
 * compressionFactor = (aabb.max - aabb.min) / 65536, index_min, index_max,
 * bitsPerIndex and compressed = 1 (CuRast_render.h:358-378); its pointers are
 * left to the caller. */
crx_status crx_compress_mesh(const float* positions, uint32_t numVertices, const uint32_t* indices,
                             uint32_t numTriangles, uint16_t* qpositions, uint32_t* packedIndices,
                             uint64_t* packedWords, crx_cmesh* meta);

/* Mesh stored IX_PU16-compressed in device memory (crx_compress_mesh of the
 * host arrays, then upload); draws through the _compressed path. */
crx_status crx_mesh_create_compressed(crx_context* ctx, const float* positions, uint32_t numVertices,
                                      const uint32_t* indices, uint32_t numTriangles, const uint32_t* colors,
                                      crx_mesh** out);

/* Same, but wraps caller-owned device buffers (not freed by the mesh). */
crx_status crx_mesh_create_device(crx_context* ctx, const float* d_positions, uint32_t numVertices,
                                  const uint32_t* d_indices, uint32_t numTriangles,
                                  const uint32_t* d_colors, const crx_box3* aabb, crx_mesh** out);
crx_status crx_mesh_destroy(crx_mesh* mesh);
crx_status crx_mesh_info(const crx_mesh* mesh, uint32_t* numTriangles, uint32_t* numVertices, crx_box3* aabb);
/* Triangles per row that crx_mesh_create found in the index buffer of a mesh
 * stored in rows (0 = none; an even count: a row of quads holds two
 * triangles per quad): the stage-1 walk of tuning.stage1_layout 2
 * uses it to read the vertices two rows share once.  No counterpart in the
 * reference, whose stage 1 walks every mesh contiguously; it orders the
 * walk and changes no result. */
crx_status crx_mesh_row_triangles(const crx_mesh* mesh, uint32_t* out);
/* Texels a w x h texture's mip chain occupies in crx_texture.data: levels
 * 0 .. max(L - 1, 7), L = floor(log2(max(w, h))), each ((w + 1) / 2) x
 * ((h + 1) / 2) texels of the level before (createMipMappedTexture,
 * LargeGlbLoader.h:97-111, sizes levels 0 .. L - 1; levels up to 7 are
 * what CRX_DISPLAY_TEXTURE may sample).  w * h must be below 2^31. */
crx_status crx_texture_chain_texels(int32_t width, int32_t height, uint64_t* texels);
/* computeMipMap (textureTools.cu:138-178, kernel :41-133): with level 0 of
 * the RGBA8 chain `data` (device) in place, computes levels 1 .. max(L - 1, 7)
 * on `stream`, one launch per level.  A texel is the per-channel sum of the
 * source level's 2 x 2 block (texels past the source's edge add 0) divided by
 * 4 and truncated; alpha is 0 for levels 1-9 and 255 beyond, as the
 * reference leaves it.  data holds crx_texture_chain_texels texels; levels
 * 1-7 of a texture smaller than 256 texels on its longer side are computed
 * here although the reference neither allocates nor computes them. */
crx_status crx_texture_compute_mipmap(crx_context* ctx, uint32_t* data, int32_t width, int32_t height, void* stream);
/* Vertex attributes of a mesh (Mesh.h:17-19, cptr_uv / cptr_normal): uvs
 * (2*numVertices floats) and normals (3*numVertices floats) are copied from
 * host arrays into device memory owned by the mesh; a NULL array leaves that
 * attribute as it is.  Only the resolve reads them (TEXTURE and UV the uvs,
 * NORMAL the normals).  Replacing an attribute waits for the device before
 * freeing the old buffer. */
crx_status crx_mesh_set_attributes(crx_mesh* mesh, const float* uvs, const float* normals);
/* Device pointers of the mesh buffers (colors may be NULL). */
crx_status crx_mesh_buffers(const crx_mesh* mesh, const float** d_positions, const uint32_t** d_indices,
                            const uint32_t** d_colors);

crx_status crx_scene_create(crx_context* ctx, crx_scene** out);
crx_status crx_scene_destroy(crx_scene* scene);
/* SNTriangles node: mesh drawn with a world transform (column-major). */
crx_status crx_scene_add_node(crx_scene* scene, crx_mesh* mesh, const crx_mat4* world, uint32_t* node_id);
crx_status crx_scene_set_transform(crx_scene* scene, uint32_t node_id, const crx_mat4* world);
/* SNTriangles::texture (SNTriangles.h:20) of a node, copied into the node's
 * CMesh as toCMesh does (CuRast_render.h:374-376).  texture->data is a
 * caller-owned device mip chain (crx_texture); NULL clears the texture. */
crx_status crx_scene_set_texture(crx_scene* scene, uint32_t node_id, const crx_texture* texture);
crx_status crx_scene_num_nodes(const crx_scene* scene, uint32_t* out);

/* View (VKRenderer.h:78-82): view matrix, projection and framebuffer size. */
typedef struct crx_view {
    crx_mat4 view;
    crx_mat4 proj;
    int32_t width;
    int32_t height;
} crx_view;

/* The node cull of SNTriangles::update (SNTriangles.h:29-105) over n
 * world-space boxes: the four normalised side planes of proj * view and the
 * positive-vertex test.  Writes the kept box indices in ascending order to ids
 * (at most `capacity`) and their number to *count.  Host only: the boxes go
 * through the same range hierarchy the server culls its nodes with, whose
 * answer equals each box's own test. */
crx_status crx_frustum_cull(const crx_view* view, const crx_box3* boxes, uint32_t n, uint32_t* ids,
                            uint32_t capacity, uint32_t* count);

/* Scene-level switches (CuRastSettings.h:31,49,55,60). */
typedef struct crx_render_settings {
    int32_t instanced;             /* RASTERIZER_VISBUFFER_INSTANCED vs RASTERIZER_VISBUFFER */
    int32_t frustum_culling;       /* enableFrustumCulling (host AABB cull per node) */
    int32_t resolve;               /* run the resolve pass */
    int32_t attribute;             /* crx_display_attribute */
} crx_render_settings;

void crx_render_settings_default(crx_render_settings* out);

/* CuRast::draw for one view: host node cull, sort/group instances,
 * per-instance transforms and triangle prefix, clear, the three stages and
 * (optionally) resolve.  framebuffer / colorbuffer are caller-owned device
 * buffers of width*height uint64 (colorbuffer may be NULL when !resolve).
 * Everything is enqueued on `stream`; the call returns once the stage queue
 * sizes are read back, while the resolve may still run: order later work on
 * the buffers after `stream`.  Frames that keep the node set, the instancing
 * switch and the culled nodes reuse the instance records already on the
 * device and upload only the per-view transforms. */
crx_status crx_render(crx_context* ctx, crx_scene* scene, const crx_view* view,
                      const crx_render_settings* settings,
                      uint64_t* framebuffer, uint64_t* colorbuffer, void* stream,
                      crx_draw_stats* stats /* nullable */);

/* crx_render for numViews (1..CRX_MAX_VIEWS) views in one batched pass.  The
 * node cull keeps a node that is visible in ANY view, so all views share one
 * instance list (and one triangle-ID numbering); each view's buffers equal a
 * single-view draw of that list.  framebuffers[v] / colorbuffers[v] are
 * device buffers of views[v].width * views[v].height uint64.  Returns as
 * crx_render does: the resolve may still run on `stream`. */
crx_status crx_render_views(crx_context* ctx, crx_scene* scene, const crx_view* views, uint32_t numViews,
                            const crx_render_settings* settings,
                            uint64_t* const* framebuffers, uint64_t* const* colorbuffers, void* stream,
                            crx_draw_stats* stats /* nullable */);

/* ------------------------------------------------------------------ */
/* Frames captured in a hipGraph (no reference counterpart: CuRast      */
/* launches each kernel from the host every frame, CuRast_render.h:     */
/* 491-592).                                                           */
/* ------------------------------------------------------------------ */

typedef struct crx_frame crx_frame;

/* Records clear -> stages 1-3 -> (resolve) of the scene for numViews views of
 * fixed sizes and buffers into one hipGraph, once.  The frame draws every
 * node of the scene (triangle ids numbered over all nodes in CuRast::draw's
 * sort order); each view skips the nodes its own cull drops.  The frame owns
 * its metadata, queues (sized by the context's tuning) and counters.  The
 * scene's node set must not change while the frame lives. */
crx_status crx_frame_create(crx_context* ctx, crx_scene* scene, const crx_view* views, uint32_t numViews,
                            const crx_render_settings* settings, uint64_t* const* framebuffers,
                            uint64_t* const* colorbuffers, crx_frame** out);
/* New cameras (same sizes) and the scene's current node transforms for the
 * next launch; waits for the previous launch to finish reading them. */
crx_status crx_frame_set_views(crx_frame* frame, const crx_view* views);
/* Enqueues the captured graph on `stream`. */
crx_status crx_frame_launch(crx_frame* frame, void* stream);
/* Waits for the last launch; CRX_QUEUE_OVERFLOW if a stage queue overflowed
 * (the result is then incomplete: recreate the frame with larger
 * crx_tuning queue capacities). */
crx_status crx_frame_wait(crx_frame* frame, crx_draw_stats* stats);
crx_status crx_frame_destroy(crx_frame* frame);

/* ------------------------------------------------------------------ */
/* Serving with continuous batching over an out-of-core node store.    */
/* No single reference counterpart: it combines CuRast::render's       */
/* per-frame draw (CuRast_render.h:873-904), the node culling of        */
/* SNTriangles::update (SNTriangles.h:29-105) and the residency /       */
/* async-upload scheme of the point-cloud octree (globals.h:433-501,    */
/* structureUpdate.cpp:256, SNCOctree.h:32,102-119) applied to          */
/* triangle nodes.                                                     */
/* ------------------------------------------------------------------ */

typedef struct crx_server crx_server;

typedef struct crx_server_config {
    uint64_t node_cache_bytes;   /* device memory for resident node geometry (LRU cache of slots) */
    uint64_t frame_bytes;        /* device memory for in-flight result buffers: requests are
                                    admitted while their W*H*8 (x2 with resolve) bytes fit */
    uint32_t max_batch_views;    /* views per batched draw, 1..CRX_MAX_VIEWS */
    uint32_t lanes;              /* batches in flight at once (each its own stream, queues and host
                                    thread), 1..8 */
    int32_t resolve;             /* also resolve every request to RGBA8 vertex colours */
    uint32_t prefetch;           /* 1 (default): after each batch, upload the nodes whose boxes touch
                                    the drawn nodes' boxes while the cache has slots not used in the
                                    last 64 batches (own stream, one node at a time); 0: upload on
                                    demand only */
    uint32_t shared_stream;      /* 0 (default): each lane its own stream, and batches run side by
                                    side; 1: the lanes admit, cull, upload and enqueue their batches
                                    concurrently but launch them on one stream, so batches run on
                                    the device in the order they were enqueued */
    uint32_t lane_backlog;       /* with every lane idle the first lane takes the work; while a lane
                                    draws, another idle lane starts a batch only when at least this
                                    many requests wait (default 4: a short queue waits for the
                                    drawing lane rather than slow its batch down by sharing the
                                    device with it); 1 = an idle lane starts at once */
    uint32_t merge_backlog;      /* a batch takes several requests only when at least this many wait
                                    (default 4), else one: the views of a batch finish together, so
                                    merging a short queue delays its oldest request */
    uint32_t reserved;
} crx_server_config;

void crx_server_config_default(crx_server_config* out);

/* Page-locked host memory (hipHostMalloc) for node geometry the server
 * streams from. */
crx_status crx_host_alloc(crx_context* ctx, size_t bytes, void** out);
crx_status crx_host_free(crx_context* ctx, void* ptr);

crx_status crx_server_create(crx_context* ctx, const crx_server_config* config, crx_server** out);
/* Stops the lanes after the submitted requests are drawn and frees the
 * server's device memory (result buffers included). */
crx_status crx_server_destroy(crx_server* server);

/* A scene node: a triangle mesh in HOST memory (page-locked, from
 * crx_host_alloc, valid while the server lives; colors nullable), its
 * object-space bounding box and world transform.  resident != 0 places it in
 * the device cache at crx_server_start (while slots remain); the others stay
 * host-resident until a request needs them.  Node ids are consecutive from 0;
 * node i's triangle ids start at the sum of the triangle counts of nodes
 * 0..i-1, whatever else a request draws. */
crx_status crx_server_add_node(crx_server* server, const float* positions, uint32_t numVertices,
                               const uint32_t* indices, uint32_t numTriangles, const uint32_t* colors,
                               const crx_box3* aabb, const crx_mat4* world, int resident, uint32_t* node_id);

/* Uploads the resident nodes and starts the lanes.  No nodes can be added
 * afterwards. */
crx_status crx_server_start(crx_server* server);

/* Queues a render request for one view; returns at once with a ticket.
 * Each lane, when free, takes the oldest queued requests whose buffers fit
 * the frame budget (up to max_batch_views), streams in the nodes they see
 * that are not resident, and draws them in one multi-view pass on its own
 * stream.  A request's visibility buffer depends only on its view and the
 * scene: it is bit-identical whether drawn alone or batched with others. */
crx_status crx_server_submit(crx_server* server, const crx_view* view, uint64_t* ticket);

typedef struct crx_request_result {
    const uint64_t* framebuffer;   /* device, width*height; valid until crx_server_release */
    const uint64_t* colorbuffer;   /* device, or NULL without resolve */
    double submit_time;            /* seconds on the server's clock */
    double start_time;             /* its batch was taken by a lane */
    double done_time;              /* its batch completed on the device */
    uint32_t batch_views;          /* views in that batch */
    uint32_t uploads;              /* nodes streamed in for that batch */
    crx_status status;             /* CRX_OK, or why the request failed */
    uint32_t reserved;
} crx_request_result;

/* Blocks until the request is drawn. */
crx_status crx_server_wait(crx_server* server, uint64_t ticket, crx_request_result* out);
/* Returns the request's buffers to the frame budget. */
crx_status crx_server_release(crx_server* server, uint64_t ticket);
/* Seconds on the server's clock (the time base of crx_request_result). */
double crx_server_clock(const crx_server* server);

typedef struct crx_server_stats {
    uint64_t requests;         /* completed */
    uint64_t batches;
    uint64_t uploads;          /* nodes streamed host -> device */
    uint64_t upload_bytes;
    uint64_t evictions;
    uint32_t resident_nodes;
    uint32_t cache_slots;
    uint32_t max_batch;        /* largest batch drawn */
    uint32_t prefetches;       /* of the uploads, nodes streamed ahead of the requests that drew them */
} crx_server_stats;

crx_status crx_server_get_stats(const crx_server* server, crx_server_stats* out);

/* Nodes a view draws (the node cull of that view), ascending ids; returns
 * their number in *count (at most `capacity` ids are written). */
crx_status crx_server_view_nodes(const crx_server* server, const crx_view* view, uint32_t* ids, uint32_t capacity,
                                 uint32_t* count);

/* ------------------------------------------------------------------ */
/* Autotuner (no reference counterpart: the reference fixes these as   */
/* compile-time constants, HostDeviceInterface.h:261-266, and a        */
/* blockDim-coupled tile size, triangles_visbuffer.cu:683-684).        */
/* ------------------------------------------------------------------ */

#define CRX_TUNE_SCHEDULE 1u   /* grid sizes, block size, batch order, cull order, fast tiny cull:
                                  the result is bit-identical under every choice */
#define CRX_TUNE_ROUTING 2u    /* small_threshold and tile_size: they move triangles between stages
                                  and so change result bits (the oracle follows the same values) */
#define CRX_TUNE_QUEUES 4u     /* size max_nontrivial / max_huge from the observed queue counts */

typedef struct crx_autotune_result {
    crx_tuning tuning;      /* best tuning found (also installed on the context) */
    float start_ms;         /* median device time of stages 1-3 with the starting tuning */
    float best_ms;          /* the same with the best tuning */
    uint32_t trials;        /* candidate configurations measured */
    uint32_t reserved;
} crx_autotune_result;

/* Coordinate descent over the knobs selected by `flags`, starting from the
 * context's tuning: each candidate draws the scene into all views `reps`
 * times (after one warm-up draw) and is scored by the median device time of
 * stages 1-3.  The buffers are overwritten.  On return the context uses the
 * best tuning. */
crx_status crx_autotune(crx_context* ctx, crx_scene* scene, const crx_view* views, uint32_t numViews,
                        uint64_t* const* framebuffers, uint32_t flags, uint32_t reps,
                        crx_autotune_result* out);

#ifdef __cplusplus
}
#endif

#endif /* CRX_H */
