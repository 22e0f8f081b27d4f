/*
 * voxelrt.h -- C ABI of libvoxelrt_hip, the MI355X (gfx950) implementation of
 * VoxelRT's base frame pipeline: packed-u32 voxel grid + one word per 8^3
 * brick, two-level Amanatides-Woo DDA, 1-spp path tracer with sun-cone NEE,
 * temporal reprojection and a 3-pass a-trous filter.
 *
 * Every entry point returns int (VRT_OK = 0, < 0 = error) and never throws;
 * vrt_last_error() gives the message of the most recent failure on a context.
 * Host arrays are owned by the caller and are read or written before the call
 * returns. Device memory is owned by the context and its scene handles.
 * A context is externally synchronised: one caller thread at a time (the
 * serving loop's internal worker thread is the only other user).
 *
 * Reference interfaces replaced (paths into astrilux1/VoxelRT):
 *   - the WebGPU frame: src/main.js:680-803 (tick: pathtrace -> temporal ->
 *     atrous x3), uniforms src/shaders/common.wgsl:13-35, per-pass bindings
 *     pathtrace.wgsl:19-30, temporal.wgsl:7-12, atrous.wgsl:8-18;
 *   - the scene buffers: src/main.js:187-203 (voxels, bricks, brickMasks
 *     written once from generateScene, src/scene.js:232-476);
 *   - the traversal: src/shaders/voxel.wgsl:73-157 (trace / traceShadow);
 *   - the HDR capture: src/main.js:628-659;
 *   - VoxelBench's Technique::frame (VoxelBench/pipeline/technique.h:8-15),
 *     through the adapter in voxelrt_technique.h.
 */
#ifndef VOXELRT_H
#define VOXELRT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes --------------------------------------------------------- */
#define VRT_OK              0
#define VRT_ERR_INVALID    -1  /* bad argument, shape or handle */
#define VRT_ERR_DEVICE     -2  /* HIP runtime or kernel launch failure */
#define VRT_ERR_NO_MEMORY  -3  /* device allocation failed */
#define VRT_ERR_BUDGET     -4  /* request exceeds the context's HBM budget */
#define VRT_ERR_NOT_READY  -5  /* serving: ticket not finished yet */
#define VRT_ERR_CANCELLED  -6  /* serving: request cancelled / rejected */
#define VRT_ERR_NO_DEVICE  -7  /* no HIP device visible */

/* ---- constants shared with the reference shaders ------------------------- */
#define VRT_BRICK            8          /* scene.js:6, BRICK */
#define VRT_MASK_WORDS      16          /* 512-bit occupancy mask per brick */
#define VRT_VOXEL_SIZE   0.0625f        /* scene.js:5, 1/16 m */
#define VRT_EMISSIVE_SCALE 24.0f        /* common.wgsl:11 */

/* Face ids follow VoxelBench (core.h:106): 0..5 = +x,-x,+y,-y,+z,-z, meaning
 * the outward normal of the face that was hit.  VRT_FACE_NONE marks a hit
 * whose normal is the zero vector (ray started inside a solid voxel while
 * moving parallel to the entry axis); VRT_FACE_MISS marks a miss. */
#define VRT_FACE_NONE  6
#define VRT_FACE_MISS  255
#define VRT_VOXEL_MISS 0xFFFFFFFFu

/* uniforms flag bits (params0.z, common.wgsl:20-22, main.js:413-416) */
#define VRT_FLAG_TEMPORAL  1u
#define VRT_FLAG_DENOISE   2u
/* Indirect light only (no reference counterpart; for VoxelBench techniques,
 * which estimate indirect light, technique.h:2-3): the primary vertex adds
 * no emission and no sun NEE, and a primary miss returns 0 instead of the
 * sky.  The RNG stream is consumed exactly as without the flag. */
#define VRT_FLAG_INDIRECT  4u

/* ---- opaque handles ------------------------------------------------------- */
typedef struct vrt_ctx vrt_ctx;
typedef struct vrt_scene vrt_scene;

typedef struct vrt_create_info {
  int32_t  device;        /* HIP device ordinal */
  uint32_t flags;         /* reserved, 0 */
  uint64_t hbm_budget;    /* bytes for scenes + per-view state; 0 = 90% of free */
} vrt_create_info;

/* One traced ray's result (16 B).  Output of vrt_trace_rays; the WGSL Hit
 * (voxel.wgsl:59-63) extended with the voxel index, face id and the DDA's two
 * step counts. */
typedef struct vrt_hit {
  float    t;             /* world metres, -1 on miss (voxel.wgsl:125) */
  uint32_t voxel;         /* x + y*nx + z*nx*ny, VRT_VOXEL_MISS on miss */
  uint16_t brick_steps;   /* brick-loop iterations (one brick word each) */
  uint16_t voxel_steps;   /* voxel-loop iterations (one mask word each) */
  uint8_t  face;          /* VRT_FACE_*, see above */
  uint8_t  _pad[3];
} vrt_hit;

/* Per-view uniforms, byte-for-byte the 304-byte Uniforms block of
 * common.wgsl:13-35 (column-major matrices).  Fields the base path does not
 * read (ReSTIR params2/3/5/6) are carried for layout compatibility. */
typedef struct vrt_uniforms {
  float    inv_view_proj[16];   /* 0   */
  float    prev_view_proj[16];  /* 64  */
  float    cam_pos[4];          /* 128 xyz metres, w = time */
  float    prev_cam_pos[4];     /* 144 */
  float    sun_dir[4];          /* 160 xyz toward the sun, w = cos(radius) */
  float    sun_radiance[4];     /* 176 rgb irradiance, w = sky intensity */
  uint32_t params0[4];          /* 192 x = frame (+seed), y = bounces, z = flags */
  float    params1[4];          /* 208 x = width, y = height, z = exposure */
  uint32_t params2[4];          /* 224 restir (unused by the base path) */
  float    params3[4];          /* 240 */
  float    params4[4];          /* 256 y = max temporal history */
  float    params5[4];          /* 272 */
  float    params6[4];          /* 288 */
} vrt_uniforms;

/* Accumulation (temporal history) storage format, main.js:148-149. */
#define VRT_ACCUM_RGBA16F 0
#define VRT_ACCUM_RGBA32F 1

/* One view to render.  Per-view device state (G-buffer and accumulation
 * ping-pong, denoise targets) is kept resident in the context, keyed by
 * view_id, and reallocated when width/height/accum_format change. */
typedef struct vrt_view {
  const vrt_scene* scene;
  vrt_uniforms     u;
  float    firefly_clamp;       /* FCLAMP_VALUE (pathtrace.wgsl:35); <=0 off */
  uint32_t view_id;             /* key for resident history */
  uint32_t spp;                 /* samples per pixel, >=1 (1 = reference) */
  uint32_t accum_format;        /* VRT_ACCUM_* */
} vrt_view;

/* Optional host destinations for one view's outputs (NULL = not copied).
 * Layouts are row-major, width*height pixels, the reference's texture
 * formats (main.js:341-349). */
typedef struct vrt_frame {
  uint16_t* radiance;   /* rgba16f demodulated radiance (pathtrace out)   */
  uint8_t*  albedo;     /* rgba8unorm sqrt(albedo)                        */
  float*    gbuf;       /* rgba32f (normal.xyz, t) of this frame          */
  void*     accum;      /* accumulation, rgba16f or rgba32f per format    */
  uint16_t* denoised;   /* rgba16f final a-trous output (denoise flag)    */
  float*    hdr;        /* rgb f32 illum*(alb8/255)^2, main.js:628-659    */
  /* filled on return */
  uint64_t  rays;       /* rays traced for this view (primary+shadow+bounce) */
  uint64_t  brick_steps;
  uint64_t  voxel_steps;
  uint8_t*  ldr;        /* rgba8 present.wgsl:40-55 output: the same illum
                           times albedo, times exposure (params1.z), ACES fit,
                           gamma 1/2.2 (the canvas capture, main.js:571-626) */
} vrt_frame;

/* Aggregate counters since vrt_create or the last vrt_stats_reset. */
typedef struct vrt_stats {
  uint64_t rays;
  uint64_t brick_steps;
  uint64_t voxel_steps;
  uint64_t hits;
  uint64_t frames;
  double   gpu_ms_pathtrace;   /* HIP-event time of the last render call */
  double   gpu_ms_temporal;
  double   gpu_ms_atrous;
  double   gpu_ms_frame;       /* pathtrace start .. last a-trous end (with
                                  batch_mode 2 the passes run as one graph:
                                  pathtrace = frame, temporal = a-trous = 0) */
  uint64_t hbm_resident;       /* bytes held by scenes + view state */
  uint64_t hbm_budget;
  /* sums over every vrt_render call since create / reset (HIP events on the
   * context's stream, read back lazily so rendering never waits on them) */
  uint64_t timed_calls;
  double   sum_ms_pathtrace;
  double   sum_ms_temporal;
  double   sum_ms_atrous;
  double   sum_ms_frame;
  uint64_t scene_evictions;  /* VRT_SCENE_EVICTABLE scenes moved to host memory */
  uint64_t scene_restores;   /* ... and brought back */
} vrt_stats;

/* ---- context -------------------------------------------------------------- */
int  vrt_create(const vrt_create_info* info, vrt_ctx** out);
void vrt_destroy(vrt_ctx* ctx);
const char* vrt_last_error(const vrt_ctx* ctx);
int  vrt_device_count(int32_t* count);

/* ---- scenes --------------------------------------------------------------- *
 * vrt_scene_upload copies a dense packed-u32 grid (x fastest, scene.js:234;
 * dims multiples of 8) to the device and builds the brick words and 512-bit
 * masks there (the build scene.js:411-433 does on the CPU). */
int vrt_scene_upload(vrt_ctx* ctx, const uint32_t* vox, uint32_t nx, uint32_t ny,
                     uint32_t nz, uint64_t scene_key, vrt_scene** out);
/* Editable scenes (VRT_SCENE_EDITABLE) also keep the dense grid resident
 * (4 B per voxel, counted in the HBM budget); vrt_scene_edit writes a box of
 * voxels (sx*sy*sz words, x fastest) into it and updates the scene on the
 * device: when no brick the box touches changes occupancy (empty / not),
 * only those bricks' masks and pool entries are rewritten in place;
 * otherwise the bricks, masks, pool and occupancy bitmaps are rebuilt (slots
 * follow brick order, so a change of occupancy anywhere moves later slots).  Work
 * already enqueued that reads the scene completes first; frames launched
 * after the call see the edit.  Replaces VoxelBench's setVox + brick
 * rebuild on a SceneEvent geomEdit (core.h:78-82, scenes.h:8-12). */
#define VRT_SCENE_EDITABLE 1u
/* Evictable scenes (no reference counterpart) may leave the device when the
 * context's HBM budget is needed -- by a scene upload, a view's state, a
 * served request -- and nothing uses them: no queued or running served
 * request, no call in progress.  The least recently used goes first; its
 * device buffers are copied to host memory and the budget is released.  Any
 * later use restores it (copied back, not rebuilt), so results do not change;
 * vrt_trace_rays_device_noalloc, which must not allocate, rejects an evicted
 * scene (vrt_scene_restore brings it back first).  Scenes without the flag
 * stay resident until released.  vrt_stats counts evictions and restores. */
#define VRT_SCENE_EVICTABLE 2u
int vrt_scene_upload_ex(vrt_ctx* ctx, const uint32_t* vox, uint32_t nx, uint32_t ny,
                        uint32_t nz, uint64_t scene_key, uint32_t flags, vrt_scene** out);
int vrt_scene_edit(vrt_ctx* ctx, vrt_scene* scene, const uint32_t* vox, uint32_t x0,
                   uint32_t y0, uint32_t z0, uint32_t sx, uint32_t sy, uint32_t sz);
int vrt_scene_release(vrt_ctx* ctx, vrt_scene* scene);
/* Bring an evicted scene back to the device (no-op when resident); *resident
 * of vrt_scene_resident says where it is now.  No reference counterpart. */
int vrt_scene_restore(vrt_ctx* ctx, vrt_scene* scene);
int vrt_scene_resident(const vrt_scene* scene, int32_t* resident);
/* How the scene's vrt_scene_edit calls were applied so far: in place (only
 * the touched bricks) or by a rebuild.  No reference counterpart. */
int vrt_scene_edit_counts(const vrt_scene* scene, uint64_t* in_place, uint64_t* rebuilt);
/* Read back the device-built brick words (nbricks) and masks (nbricks*16). */
int vrt_scene_bricks(vrt_ctx* ctx, const vrt_scene* scene, uint32_t* bricks,
                     uint32_t* masks);
int vrt_scene_dims(const vrt_scene* scene, uint32_t dims[3], uint64_t* bytes);

/* ---- traversal (voxel.wgsl:73-153) ---------------------------------------- *
 * rays: n records of 6 floats (origin xyz in metres, direction xyz).
 * Host arrays in/out. */
int vrt_trace_rays(vrt_ctx* ctx, const vrt_scene* scene, const float* rays,
                   uint64_t n, float tmax, vrt_hit* hits);
/* Same on device-resident buffers (hipMalloc'd by the caller), enqueued on
 * the context's stream; *ms receives the kernel's HIP-event time. */
int vrt_trace_rays_device(vrt_ctx* ctx, const vrt_scene* scene,
                          const float* d_rays, uint64_t n, float tmax,
                          vrt_hit* d_hits, float* ms);

/* Allocation-free form of vrt_trace_rays_device.  No reference counterpart:
 * for callers that cannot allocate device memory -- code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  The same
 * kernels and the same hits as vrt_trace_rays_device with the context's
 * tuning; their scratch (the ray / step / hit counters and the persistent
 * kernels' ray queue) is a fixed 1280-byte device array inside the
 * library, one per GPU, held until the process ends and cleared by a kernel
 * at the start of each call (no hipMalloc, no hipFree, no memset node).
 * n is at most VRT_NOALLOC_MAX_RAYS (the queue head is 32-bit); larger n is
 * rejected with VRT_ERR_INVALID.  Every call on one GPU shares that array:
 * calls that overlap on the device (other contexts or streams) race and
 * must be serialised by the caller.  The launch is enqueued on the
 * context's stream and the call returns without waiting; its rays are not
 * added to vrt_stats. */
#define VRT_NOALLOC_MAX_RAYS 0x7fffffffull
int vrt_trace_rays_device_noalloc(vrt_ctx* ctx, const vrt_scene* scene,
                                  const float* d_rays, uint64_t n, float tmax,
                                  vrt_hit* d_hits);

/* Device buffers for the device-resident entry points (rays, hits), owned by
 * the context and counted in its HBM budget; freed by vrt_device_free or
 * vrt_destroy.  vrt_copy moves bytes between host and device (or device and
 * device) on the context's stream and waits for the copy.  No reference
 * counterpart (the WebGPU host creates its buffers with
 * device.createBuffer, main.js:187-203). */
#define VRT_COPY_TO_DEVICE 1u
#define VRT_COPY_TO_HOST   2u
#define VRT_COPY_DEVICE    3u
int vrt_device_alloc(vrt_ctx* ctx, uint64_t bytes, void** dptr);
int vrt_device_free(vrt_ctx* ctx, void* dptr);
int vrt_copy(vrt_ctx* ctx, void* dst, const void* src, uint64_t bytes, uint32_t kind);

/* ---- frames ---------------------------------------------------------------- *
 * Render nviews views: pathtrace -> temporal -> (if VRT_FLAG_DENOISE)
 * a-trous steps 1/2/4, per view.  nviews > 1 runs the batched form: all
 * views' pixels share one persistent launch per pass.  outs may be NULL. */
int vrt_render(vrt_ctx* ctx, const vrt_view* views, uint32_t nviews,
               vrt_frame* outs);
/* Drop the resident state of one view (history restarts). */
int vrt_view_release(vrt_ctx* ctx, uint32_t view_id);
/* Page-locked host memory for frame outputs (vrt_frame destinations,
 * vrt_wait's hdr): read-backs into it run as direct DMA at full link speed,
 * where a pageable destination goes through driver staging buffers.  Plays
 * the part of the MAP_READ staging buffer the WebGPU host maps after
 * copyTextureToBuffer (main.js:571-626).  Free with vrt_host_free (ctx may
 * then be NULL, e.g. after vrt_destroy). */
int vrt_host_alloc(vrt_ctx* ctx, uint64_t bytes, void** ptr);
int vrt_host_free(vrt_ctx* ctx, void* ptr);

/* Single passes on host arrays, for the WebGPU pass interface
 * (temporal.wgsl:7-12, atrous.wgsl:8-18). */
int vrt_temporal_pass(vrt_ctx* ctx, const vrt_uniforms* u,
                      const uint16_t* radiance, const float* gbuf_cur,
                      const void* accum_prev, const float* gbuf_prev,
                      uint32_t accum_format, void* accum_out);
int vrt_atrous_pass(vrt_ctx* ctx, const vrt_uniforms* u, const void* src,
                    uint32_t src_format, const float* gbuf, int32_t step,
                    uint16_t* dst);

/* ---- launch tuning ----------------------------------------------------------- *
 * Kernel variants and launch shapes.  Every variant computes identical
 * results; only speed differs.  vrt_autotune measures candidates on the
 * given scene and keeps the fastest in the context.  No reference
 * counterpart: WebGPU dispatches fixed 8x8 workgroups (main.js:728-734). */
typedef struct vrt_tuning {
  uint32_t trace_kernel;      /* traversal: 0 = one ray per lane, nested brick / voxel
                                 loops on global memory,
                                 4 = one ray per lane, single loop, both levels in LDS,
                                 5 = persistent ray queue, single loop in LDS,
                                 6 = one ray per lane, single loop on global memory
                                     (the form 4 and 5 fall back to when their LDS
                                     does not fit, 512^3 and up),
                                 7 = step-level refill on persistent waves (default):
                                     lanes wait at brick boundaries and at finished
                                     rays; the wave switches levels for
                                     trace_min_refill waiting lanes and writes
                                     records / starts rays for trace_steps finished
                                     lanes at once; rays come from a queue split over
                                     eight counters; the DDA in LDS where it fits,
                                     else on global memory.
                                 1, 2 and 3 are rejected. */
  uint32_t trace_blocks;      /* trace 5 and 7: persistent grid (256-thread blocks), 0 = 1 */
  uint32_t trace_steps;       /* trace 5: DDA iterations between queue refill checks;
                                 trace 7: finished lanes (of 64) collected per pass */
  uint32_t trace_min_refill;  /* trace 5: idle lanes (of 64) that trigger a refill;
                                 trace 7: lanes waiting at a brick boundary per pass */
  uint32_t pt_kernel;         /* path tracer:
                                 0 = one pixel per lane, nested loops on global memory,
                                 4 = one pixel per lane, single loop, both levels in LDS,
                                 7 = as 4 with the single loop on global memory,
                                 8 = step-level refill on persistent waves (default):
                                     pt_blocks waves (0 = 16 per CU) take 8x8 pixel
                                     patches from a queue; lanes wait at brick
                                     boundaries and at finished rays, and the wave
                                     switches levels for pt_min_refill waiting lanes
                                     and shades pt_shade_min finished rays at once;
                                     the single-loop DDA in LDS where it fits (four
                                     waves per block share it), pt_tile 1 = on global
                                     memory always.
                                 1, 2, 3, 5 and 6 are rejected.  The batched forms
                                 run pt 0, 4, 7 or 8 (see batch_tile). */
  uint32_t pt_blocks;         /* pt 8: persistent waves, 0 = 16 per CU */
  uint32_t unused0;           /* not read; keeps the layout */
  uint32_t pt_min_refill;     /* pt 8: lanes (of 64) waiting at a brick boundary
                                 before the wave switches their levels (0 = 1) */
  uint32_t unused1;           /* not read; keeps the layout */
  uint32_t pt_waves;          /* minimum waves per SIMD the register allocation must
                                 allow: 0 = compiler's choice, 4 or 5; 6 for pt 8 */
  uint32_t batch_mode;        /* batched form (nviews > 1, serving):
                                 0 = one grid launch per pass over all tiles,
                                 1 = persistent path tracer on a shared tile
                                     queue + fixed-grid filter passes,
                                 2 = as 1, replayed from a captured hipGraph */
  uint32_t batch_blocks;      /* persistent path-tracer grid, 0 = 8 per CU */
  uint32_t dda_batch;         /* traversal kernels 4, 5 and 6: lanes of a wave that wait
                                 at occupied bricks before they enter them together
                                 (1..64; 0 = 1, enter at once).  The path tracers
                                 other than pt 8 enter at once. */
  uint32_t empty_skip;        /* traversal kernels 4 and 5: 1 = take the coarse steps
                                 through runs of empty brick rows in bulk (same
                                 results, fewer instructions on sparse scenes),
                                 0 = step by step.  The path tracers do not read it. */
  uint32_t pt_order;          /* order in which blocks take tiles (pt 8: patches):
                                 0 row-major, 1 one band of tiles per XCD (not pt 8),
                                 2 row-major from the last tile (default: the bottom
                                 rows, usually the costliest, start first), 3 bands
                                 from their last tile (pt 8: as 2) */
  uint32_t pt_tile;           /* pt 0, 4, 7: 0 = a 16x16 tile per 256-thread block,
                                 1 = 8x8 tiles of one wave (pt 7), 2 / 3 = 32x32 /
                                 32x16 tiles whose lanes take the next pixel when
                                 theirs is done, one ray per lane per round (pt 4
                                 and 7), 4 / 5 = the same on one-wave blocks of
                                 16x16 / 16x8 tiles (pt 7).
                                 pt 8: 0 = DDA in LDS where it fits, 1 = DDA on
                                 global memory */
  uint32_t trace_waves;       /* traversal kernels 4 to 7: minimum waves per SIMD
                                 the register allocation must allow: 0 = compiler's
                                 choice, 6 (pays on sparse scenes) */
  uint32_t pt_shade_min;      /* pt 8: finished rays (of 64) a wave collects before it
                                 shades them together (0 = 1: at once) */
  uint32_t batch_tile;        /* pt 8 in the batched forms: 0 = its batched form (one-wave
                                 blocks, the DDA on global memory, each wave on one
                                 view at a time: chunks of 1..8 patches of 8x8 pixels,
                                 two or more chunks per wave, from a queue split over
                                 eight counters; pt_blocks waves, 0 = 16 per CU);
                                 2..5 = pt 7 with 4 waves per SIMD on that pt_tile
                                 form */
  uint32_t reserved;
} vrt_tuning;
int vrt_tuning_get(vrt_ctx* ctx, vrt_tuning* out);
int vrt_tuning_set(vrt_ctx* ctx, const vrt_tuning* in);
/* Time the candidate variants on `scene` (traversal on `nrays` random rays,
 * path tracing on a width x height view at the scene's spawn-like pose
 * given by u) and keep the fastest.  Either part may be skipped with 0.
 * The best traversal launch shape depends on the launch size: pass the ray
 * count of the launches the tuning will serve. */
int vrt_autotune(vrt_ctx* ctx, const vrt_scene* scene, uint64_t nrays,
                 const vrt_uniforms* u, vrt_tuning* chosen);
/* Time the batched form's launch shapes on `views` (2 or more: their scenes,
 * sizes, poses, spp, bounces and flags, rendered as one vrt_render call
 * would) and keep the fastest: batch_mode (among those whose bit is set in
 * `modes`, bit m = batch_mode m; 0 = all), pt_tile (the path tracer's tile
 * form) and batch_blocks (the persistent grid), with the path-tracer variant
 * and waves of the current tuning (e.g. from vrt_autotune).  The sweep
 * renders into view states of its own: the views' histories are untouched,
 * and its rays do not count in the statistics.  No reference counterpart. */
int vrt_autotune_batch(vrt_ctx* ctx, const vrt_view* views, uint32_t nviews,
                       uint32_t modes, vrt_tuning* chosen);

/* ---- stats ------------------------------------------------------------------ */
int vrt_stats_get(vrt_ctx* ctx, vrt_stats* out);
int vrt_stats_reset(vrt_ctx* ctx);

/* ---- serving: continuous batching ------------------------------------------- *
 * A request is one frame of one view (scene, pose, size, spp, bounces, flags).
 * vrt_submit queues it and returns at once.  A worker thread inside the
 * context groups queued requests into batched launches (the batched form of
 * vrt_render) on `streams` device streams: a request that arrives while other
 * batches are on the GPU joins the next launch as soon as a stream and HBM
 * free up, and that launch overlaps the running ones on the device.
 *   - Admission is by HBM footprint: the request's view state (when it is not
 *     resident yet) plus its result buffer must fit the context's budget.
 *     View state of views with no queued or running request is evicted
 *     least-recently-used first (its temporal history restarts, as after
 *     vrt_view_release).  A request that cannot fit even with everything else
 *     evicted fails with VRT_ERR_BUDGET.
 *   - Frames of one view_id run in submission order on one stream, one frame
 *     of a view per launch (temporal history ping-pong).
 *   - Results stay on the device until vrt_poll / vrt_wait copy them out:
 *     hdr (main.js:628-659) and the request's own ray / step counters.
 *   - Every pixel's result is the same as vrt_render of that view alone.
 * Scenes referenced by queued or running requests cannot be released.
 * No reference counterpart: the reference renders one view per
 * requestAnimationFrame tick (main.js:680-803). */
typedef uint64_t vrt_ticket;

typedef struct vrt_request {
  vrt_view view;
  uint32_t flags;           /* reserved, 0 */
  uint32_t _pad;
} vrt_request;

typedef struct vrt_serve_config {
  uint32_t streams;         /* device streams batches overlap on, 1..4 (default 4) */
  uint32_t max_batch_views; /* requests per launch, 1..1024 (default 8) */
  uint64_t max_batch_pixels;/* pixels x spp per launch (default 4M) */
} vrt_serve_config;

/* Host steady-clock times in ms since vrt_create. */
typedef struct vrt_timing {
  double   submit_ms;
  double   launch_ms;       /* the batch holding the request was enqueued */
  double   done_ms;         /* the worker saw that batch complete */
  double   latency_ms;      /* done_ms - submit_ms */
  double   gpu_ms;          /* HIP-event span of that batch */
  uint32_t batch_views;     /* requests in that batch */
  uint32_t stream;          /* stream index it ran on */
  uint64_t batch_id;
} vrt_timing;

typedef struct vrt_serve_stats {
  uint64_t submitted, completed, cancelled, failed;
  uint64_t batches, batched_requests;
  uint64_t evictions;
  uint64_t queued, running;        /* right now */
  uint64_t max_running_batches;    /* most batches seen on the device at once */
} vrt_serve_stats;

/* Configure before the first vrt_submit (NULL = defaults). */
int vrt_serve_configure(vrt_ctx* ctx, const vrt_serve_config* cfg);
int vrt_submit(vrt_ctx* ctx, const vrt_request* req, vrt_ticket* out);
/* VRT_ERR_NOT_READY while queued or running.  On VRT_OK the results are
 * copied to `out` (hdr when non-NULL, counters always; other vrt_frame
 * fields are ignored), `timing` is filled when non-NULL, and the ticket is
 * retired.  A failed request returns its error once and is retired. */
int vrt_poll(vrt_ctx* ctx, vrt_ticket t, vrt_frame* out, vrt_timing* timing);
/* As vrt_poll, blocking up to timeout_ms (0 = no limit). */
int vrt_wait(vrt_ctx* ctx, vrt_ticket t, uint32_t timeout_ms, vrt_frame* out,
             vrt_timing* timing);
/* Cancel a request that has not been launched yet (VRT_ERR_INVALID once it
 * runs or finished).  The ticket is retired. */
int vrt_cancel(vrt_ctx* ctx, vrt_ticket t);
int vrt_serve_stats_get(vrt_ctx* ctx, vrt_serve_stats* out);

/* ---- host helpers (no device work) ------------------------------------------ *
 * Camera matrices exactly as src/camera.js:47-56 + src/math.js (double math,
 * rounded to f32): outputs column-major viewProj and invViewProj. */
void vrt_camera_matrices(const double pos[3], double yaw, double pitch,
                         double fov_y, double aspect, double near_z,
                         double far_z, float view_proj[16],
                         float inv_view_proj[16]);
/* Fill a vrt_uniforms the way main.js:399-426 does for the base preset. */
/* Camera: fov 70 deg, near 0.05, far 400 (camera.js:10-12).  prev_view_proj
 * / prev_cam_pos may be NULL (first frame: current pose, main.js:401). */
void vrt_uniforms_init(vrt_uniforms* u, const double pos[3], double yaw,
                       double pitch, const float prev_view_proj[16],
                       const double prev_cam_pos[3], uint32_t width,
                       uint32_t height, uint32_t frame, uint32_t bounces,
                       uint32_t flags, double sun_azimuth,
                       double sun_elevation, double sun_cos_radius,
                       float max_history);

/* ---- seeded scene generators (host, no device work) ------------------------ */
#define VRT_SCENE_RANDOM_BRICKS 1  /* traversal microbench grid */
#define VRT_SCENE_WORLD         2  /* terrain + Cornell room + boxes + trees + lanterns */
#define VRT_SCENE_VOXELRT       3  /* restatement of generateScene(N, variant), scene.js:232-409 */

typedef struct vrt_scene_desc {
  uint32_t kind;               /* VRT_SCENE_* */
  uint32_t nx, ny, nz;         /* VOXELRT: cubic, nx used */
  uint64_t seed;
  float    p_brick, p_voxel;   /* RANDOM_BRICKS: brick / voxel occupancy probability */
  float    h_min, h_max;       /* WORLD: terrain height range (voxels) */
  uint32_t room, room_h;       /* WORLD: Cornell room interior extent / height, 0 = none */
  uint32_t boxes, trees, lanterns;  /* WORLD: object counts */
  uint32_t variant;            /* VOXELRT: 0 = 'default', 1 = 'lamps' */
} vrt_scene_desc;

/* Fill vox (nx*ny*nz packed words, x fastest). */
int vrt_generate_scene(const vrt_scene_desc* desc, uint32_t* vox);
/* Spawn pose of a generated scene (scene.js:463-472 for VRT_SCENE_VOXELRT;
 * outside the room's doorway, looking in, for VRT_SCENE_WORLD). */
int vrt_scene_spawn(const vrt_scene_desc* desc, double pos[3], double* yaw, double* pitch);

#ifdef __cplusplus
}
#endif
#endif /* VOXELRT_H */
