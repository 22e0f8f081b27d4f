/*
 * vrt_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU FP32 restatement of VoxelRT's WGSL base frame pipeline, used as the
 * parity oracle for libvoxelrt_hip.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load it, and only as the checker or the timed
 * CPU baseline; the product never links or calls it.
 *
 * Each function names the reference file:line it restates.  Pinning (what
 * checks each function against something other than itself) is listed per
 * function; tests/test_oracle_*.py hold those checks.
 *
 * Arithmetic conventions (identical on the device side, so results are
 * bit-comparable):
 *   - FP32, no FMA contraction (-ffp-contract=off), IEEE div and sqrt;
 *   - min(a,b) = b<a ? b : a, max(a,b) = a<b ? b : a (first operand on ties);
 *   - dot(a,b) = (ax*bx + ay*by) + az*bz; normalize(v) = v / sqrt(dot(v,v));
 *   - mix(a,b,t) = a*(1-t) + b*t (WGSL definition);
 *   - sin/cos/exp2/log2 are this project's own (Cody-Waite reduction + Taylor
 *     series, see vrt_oracle.cpp); pow(x,8) and pow(x,32) by squaring; the
 *     a-trous weights' e^x is exp_w (rint split, Taylor, ldexp);
 *   - bilinear history sampling with 8 fractional bits of texel position.
 */
#ifndef VRT_ORACLE_H
#define VRT_ORACLE_H
#include <stdint.h>
#include "../include/voxelrt.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_grid {
  const uint32_t* vox;     /* nx*ny*nz packed words, x fastest */
  const uint32_t* bricks;  /* (nx/8)*(ny/8)*(nz/8) words, 0 = empty */
  const uint32_t* masks;   /* 16 words per brick */
  int32_t nx, ny, nz;
  float voxel_size;        /* metres per voxel (1/16 in the reference) */
} orc_grid;

/* per-pixel counters of one render */
typedef struct orc_counters {
  uint64_t rays, brick_steps, voxel_steps, hits;
} orc_counters;

/* scene.js:411-433 (brick occupancy + 512-bit masks), generalised to
 * non-cubic dims that are multiples of 8.  Returns 0, or -1 on bad dims. */
int orc_build_bricks(const uint32_t* vox, int nx, int ny, int nz,
                     uint32_t* bricks, uint32_t* masks);

/* voxel.wgsl:73-153 trace(): one ray, world units.  n_out (3 floats) and
 * mat_out may be NULL. */
void orc_trace(const orc_grid* g, const float ro[3], const float rd[3],
               float max_t_world, vrt_hit* hit, float* n_out,
               uint32_t* mat_out);
/* batched, OpenMP over rays; rays = 6 floats per ray */
void orc_trace_rays(const orc_grid* g, const float* rays, uint64_t n,
                    float max_t_world, vrt_hit* hits, int threads);

/* common.wgsl:69-84 */
uint32_t orc_pcg_hash(uint32_t v);
/* first n rand() outputs of the stream of pixel (px,py) at frame */
void orc_rand_stream(uint32_t px, uint32_t py, uint32_t frame, int n,
                     uint32_t* states, float* values);

/* this project's transcendental functions (shared definition with the
 * device code; accuracy checked against libm in the tests) */
float orc_sin(float x);
float orc_cos(float x);
float orc_exp2(float x);
float orc_log2(float x);
float orc_exp(float x);
float orc_pow(float x, float y);
/* e^x as the a-trous weights compute it (x <= 0; NaN and x <= -160 ln 2
 * give 0) */
float orc_exp_w(float x);
/* orc_exp_w of the float bit patterns start + i*stride (mod 2^32), i < n */
void orc_exp_w_table(uint64_t start, uint64_t stride, uint64_t n, float* out);
/* bilinear filter position of texel coordinate tc, 8 fractional bits */
void orc_sub_texel(float tc, float* fl, float* fr);

/* f32 -> f16 bits, round to nearest even (denormals, overflow to inf) */
uint16_t orc_f32_to_f16(float f);
float orc_f16_to_f32(uint16_t h);

/* math.js:3-79 + camera.js:47-56 in double, rounded to f32 like
 * Float32Array stores */
void orc_camera_matrices(const double pos[3], double yaw, double pitch,
                         double fov_y, double aspect, double near_z,
                         double far_z, float view_proj[16],
                         float inv_view_proj[16]);

/* common.wgsl:90-95, 101-139 pieces (for unit tests) */
void orc_camera_ray(const vrt_uniforms* u, float uvx, float uvy, float out[3]);
void orc_sky_color(const vrt_uniforms* u, const float rd[3], int disc,
                   float out[3]);
void orc_sample_sun_dir(const vrt_uniforms* u, uint32_t* rng, float out[3]);
void orc_cosine_hemisphere(const float n[3], uint32_t* rng, float out[3]);

/* pathtrace.wgsl:248-330 (base branch, RF_RESTIR off) over pixel rows
 * [y0, y1).  Outputs are full-image arrays (row-major, width*height px):
 * radiance rgba16f bits, albedo rgba8unorm, gbuf rgba32f.  counters may be
 * NULL.  spp > 1 averages spp streams keyed by frame*spp+s. */
void orc_pathtrace(const orc_grid* g, const vrt_uniforms* u, float fclamp,
                   uint32_t spp, int y0, int y1, uint16_t* radiance,
                   uint8_t* albedo, float* gbuf, orc_counters* counters,
                   int threads);
/* same, but only the listed pixel indices (y*w+x); outputs still indexed by
 * pixel index in full-image arrays. */
void orc_pathtrace_pixels(const orc_grid* g, const vrt_uniforms* u,
                          float fclamp, uint32_t spp, const uint32_t* pix,
                          uint64_t npix, uint16_t* radiance, uint8_t* albedo,
                          float* gbuf, orc_counters* counters, int threads);

/* temporal.wgsl:17-77.  accum_prev/accum_out are rgba16f bits
 * (format 0) or rgba32f (format 1). */
void orc_temporal(const vrt_uniforms* u, const uint16_t* radiance,
                  const float* gbuf_cur, const void* accum_prev,
                  const float* gbuf_prev, uint32_t format, void* accum_out,
                  int threads);

/* atrous.wgsl:26-83, one pass of the given step.  src is rgba16f bits
 * (format 0) or rgba32f (format 1); dst is rgba16f bits. */
void orc_atrous(const vrt_uniforms* u, const void* src, uint32_t src_format,
                const float* gbuf, int step, uint16_t* dst, int threads);

/* main.js:628-659: hdr = illum * (alb8/255)^2, rgb f32 */
void orc_hdr(int w, int h, const void* illum, uint32_t illum_format,
             const uint8_t* albedo, float* hdr);

/* present.wgsl:40-55: c = illum * (alb8/255)^2 * exposure, ACES fit,
 * pow(c, 1/2.2), stored rgba8unorm (alpha 255), FP32 like the shader. */
void orc_present(int w, int h, const void* illum, uint32_t illum_format,
                 const uint8_t* albedo, float exposure, uint8_t* rgba);

#ifdef __cplusplus
}
#endif
#endif
