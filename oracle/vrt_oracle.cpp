// vrt_oracle.cpp -- TEST INFRASTRUCTURE ONLY (parity oracle, CPU baseline).
//
// Plain, scalar FP32 restatement of VoxelRT's WGSL base frame pipeline.  Each
// function follows the cited reference lines operation by operation; the
// conventions shared with the device code are listed in vrt_oracle.h.  Built
// with -ffp-contract=off so no multiply-add is fused.
#include "vrt_oracle.h"

#include <omp.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

struct v3 {
  float x, y, z;
};
inline v3 mk(float x, float y, float z) { return v3{x, y, z}; }
inline v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
inline v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
inline v3 mul(v3 a, v3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
inline v3 muls(v3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
inline v3 divv(v3 a, v3 b) { return mk(a.x / b.x, a.y / b.y, a.z / b.z); }
inline v3 divs(v3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
inline float fmin_(float a, float b) { return b < a ? b : a; }
inline float fmax_(float a, float b) { return a < b ? b : a; }
inline float fclamp_(float x, float lo, float hi) { return fmin_(fmax_(x, lo), hi); }
inline float dot3(v3 a, v3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline v3 normalize3(v3 v) {
  float l = std::sqrt(dot3(v, v));
  return divs(v, l);
}
inline v3 cross3(v3 a, v3 b) {
  return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
inline float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }
inline v3 mix3(v3 a, v3 b, float t) {
  return mk(mixf(a.x, b.x, t), mixf(a.y, b.y, t), mixf(a.z, b.z, t));
}

// common.wgsl:7-11
const float PI_F = 3.14159265358979f;
const float TWO_PI_F = 2.0f * PI_F;  // const-folded (2*PI) as in WGSL
const float INV_PI_F = 0.31830988618f;
const float EMISSIVE_SCALE = 24.0f;

// Bilinear filter position of a texel coordinate tc = uv * size - 0.5
// (temporal.wgsl:63, textureSampleLevel with a linear sampler).  WGSL leaves
// the filter's precision to the implementation; texture units filter in
// fixed point with 8 fractional bits of sub-texel position, so tc is rounded
// to a multiple of 1/256 (nearest, ties to even) before it is split into the
// floor and the fraction.  (An exact-FP32 split instead blends ~1e-4 of a
// neighbour into the history at every static-camera frame, where the
// reprojected position misses the texel centre by rounding alone; over
// hundreds of accumulated frames that leaks several percent of the
// neighbouring pixels -- e.g. sky -- into the accumulation.)
void sub_texel(float tc, float* fl, float* fr) {
  const float q = std::rint(tc * 256.0f);
  *fl = std::floor(q * (1.0f / 256.0f));
  *fr = (q - *fl * 256.0f) * (1.0f / 256.0f);
}

// ---------------------------------------------------------------------------
// Transcendentals.  The reference calls WGSL cos/sin/pow/exp, whose results
// are implementation-defined (WGSL allows 2^-11 absolute error for cos/sin).
// This project defines its own: Cody-Waite argument reduction and Taylor
// series (coefficients 1/n!), evaluated in a fixed order with no fused ops, so
// CPU and GPU agree bit for bit.  Accuracy vs libm is checked in the tests.
// ---------------------------------------------------------------------------
const float TWO_OVER_PI = 0x1.45f306p-1f;
const float PIO2_HI = 0x1.92p+0f;        // 8 significant bits: k*HI exact
const float PIO2_MID = 0x1.fb6p-12f;     // next 12 bits
const float PIO2_LO = -0x1.777a5cp-25f;  // remainder
const float LN2_F = 0x1.62e43p-1f;
const float LOG2E_F = 0x1.715476p+0f;
const float SQRT2_F = 1.41421356f;

void sincos_own(float x, float* s, float* c) {
  float k = std::floor(x * TWO_OVER_PI + 0.5f);
  float r = ((x - k * PIO2_HI) - k * PIO2_MID) - k * PIO2_LO;
  float r2 = r * r;
  float sp = (((float)(1.0 / 362880.0) * r2 + (float)(-1.0 / 5040.0)) * r2 +
              (float)(1.0 / 120.0)) * r2 + (float)(-1.0 / 6.0);
  float sn = r + (r * r2) * sp;
  float cp = ((((float)(-1.0 / 3628800.0) * r2 + (float)(1.0 / 40320.0)) * r2 +
               (float)(-1.0 / 720.0)) * r2 + (float)(1.0 / 24.0)) * r2 - 0.5f;
  float cs = 1.0f + r2 * cp;
  int q = ((int)k) & 3;
  switch (q) {
    case 0: *s = sn; *c = cs; break;
    case 1: *s = cs; *c = -sn; break;
    case 2: *s = -sn; *c = -cs; break;
    default: *s = -cs; *c = sn; break;
  }
}

// 2^x = 2^k * e^(f ln2), k = round(x): Taylor series to degree 7, Horner
// steps as correctly rounded fused multiply-adds (the device's exp2_r).
// target("fma"): the fmaf calls as FMA3 instructions rather than libm calls
// (the same correctly rounded results).
__attribute__((target("fma"))) float exp2_own(float x) {
  if (x != x) return x;
  if (!(x > -126.0f)) return 0.0f;
  if (x >= 128.0f) return INFINITY;
  float k = std::floor(x + 0.5f);
  float f = x - k;
  float y = f * LN2_F;
  float p = std::fmaf((float)(1.0 / 5040.0), y, (float)(1.0 / 720.0));
  p = std::fmaf(p, y, (float)(1.0 / 120.0));
  p = std::fmaf(p, y, (float)(1.0 / 24.0));
  p = std::fmaf(p, y, (float)(1.0 / 6.0));
  p = std::fmaf(p, y, 0.5f);
  p = std::fmaf(p, y, 1.0f);
  p = std::fmaf(p, y, 1.0f);
  int ki = (int)k;
  if (ki > 127) {
    p = p * 2.0f;
    ki -= 1;
  }
  uint32_t bits = (uint32_t)(ki + 127) << 23;
  float sc;
  std::memcpy(&sc, &bits, 4);
  return p * sc;
}

float log2_own(float x) {
  if (x != x || x < 0.0f) return NAN;
  if (x == 0.0f) return -INFINITY;
  if (x == INFINITY) return INFINITY;
  int e = 0;
  uint32_t b;
  std::memcpy(&b, &x, 4);
  if (b < 0x00800000u) {  // denormal: scale into the normal range
    x = x * 8388608.0f;
    std::memcpy(&b, &x, 4);
    e = -23;
  }
  e += (int)((b >> 23) & 0xffu) - 127;
  uint32_t mb = (b & 0x007fffffu) | 0x3f800000u;
  float m;
  std::memcpy(&m, &mb, 4);
  if (m > SQRT2_F) {
    m = m * 0.5f;
    e += 1;
  }
  float s = (m - 1.0f) / (m + 1.0f);
  float s2 = s * s;
  float p = ((((float)(2.0 / 11.0) * s2 + (float)(2.0 / 9.0)) * s2 +
              (float)(2.0 / 7.0)) * s2 + (float)(2.0 / 5.0)) * s2 + (float)(2.0 / 3.0);
  p = p * s2 + 2.0f;
  float lnm = s * p;
  return (float)e + lnm * LOG2E_F;
}

inline float pow_own(float x, float y) {
  if (x == 0.0f) return y > 0.0f ? 0.0f : (y == 0.0f ? 1.0f : INFINITY);
  return exp2_own(y * log2_own(x));
}
inline float exp_own(float x) { return exp2_own(x * LOG2E_F); }

// e^x for the a-trous weights (atrous.wgsl:74-75; WGSL allows exp 3 + 2|x|
// ulps), whose arguments are <= 0 or NaN: z = x log2(e), floored at -160 (a
// NaN argument becomes -160, so the weight is 0), k = rint(z) (nearest, ties
// to even), f = z - k in [-1/2, 1/2] (exact), then 2^k * 2^f with a degree-5
// polynomial for 2^f fitted for relative error on [-1/2, 1/2] (constant term
// 1; within 2.1 ulps) in fused multiply-add steps, scaled by ldexp
// (correctly rounded, subnormal results included).  Every step is exact or
// correctly rounded, so the device's exp_w (vrt_device.h) gives the same bits.
__attribute__((target("fma"))) float exp_w(float x) {
  const float z = std::fmax(x * LOG2E_F, -160.0f);
  const float k = std::rint(z);
  const float f = z - k;  // exact, in [-1/2, 1/2]
  float p = std::fmaf(0x1.5bb9fep-10f, f, 0x1.3cea88p-7f);
  p = std::fmaf(p, f, 0x1.c6b754p-5f);
  p = std::fmaf(p, f, 0x1.ebf9bcp-3f);
  p = std::fmaf(p, f, 0x1.62e42ap-1f);
  p = std::fmaf(p, f, 1.0f);
  return std::ldexp(p, (int)k);
}
inline float pow8(float x) {
  float x2 = x * x;
  float x4 = x2 * x2;
  return x4 * x4;
}
inline float pow32(float x) {
  float a = x * x;
  a = a * a;
  a = a * a;
  a = a * a;
  return a * a;
}

// ---------------------------------------------------------------------------
// f16 / unorm8 storage (texture formats, main.js:341-349)
// ---------------------------------------------------------------------------
uint16_t f32_to_f16(float f) {
  uint32_t x;
  std::memcpy(&x, &f, 4);
  uint32_t sign = (x >> 16) & 0x8000u;
  uint32_t ax = x & 0x7fffffffu;
  if (ax >= 0x7f800000u) return (uint16_t)(sign | (ax > 0x7f800000u ? 0x7e00u : 0x7c00u));
  if (ax >= 0x38800000u) {  // normal half range (or overflow)
    uint32_t h = (ax >> 13) - (112u << 10);
    uint32_t low = ax & 0x1fffu;
    if (low > 0x1000u || (low == 0x1000u && (h & 1u))) h++;
    if (h >= 0x7c00u) h = 0x7c00u;
    return (uint16_t)(sign | h);
  }
  uint32_t e = ax >> 23;
  if (e == 0) return (uint16_t)sign;
  uint32_t mant = (ax & 0x7fffffu) | 0x800000u;
  uint32_t shift = 126u - e;  // value in units of 2^-24 is mant >> shift
  if (shift > 24u) return (uint16_t)sign;
  uint32_t q = mant >> shift;
  uint32_t rem = mant & ((1u << shift) - 1u);
  uint32_t half = 1u << (shift - 1u);
  if (rem > half || (rem == half && (q & 1u))) q++;
  return (uint16_t)(sign | q);
}

float f16_to_f32(uint16_t h) {
  uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
  uint32_t e = (h >> 10) & 0x1fu;
  uint32_t m = h & 0x3ffu;
  uint32_t bits;
  if (e == 0) {
    if (m == 0) {
      bits = sign;
    } else {
      float v = (float)m * 0x1p-24f;
      std::memcpy(&bits, &v, 4);
      bits |= sign;
    }
  } else if (e == 31) {
    bits = sign | 0x7f800000u | (m << 13);
  } else {
    bits = sign | ((e + 112u) << 23) | (m << 13);
  }
  float f;
  std::memcpy(&f, &bits, 4);
  return f;
}

inline uint8_t f32_to_unorm8(float x) {
  float c = fclamp_(x, 0.0f, 1.0f);
  return (uint8_t)std::floor(c * 255.0f + 0.5f);
}

// ---------------------------------------------------------------------------
// PCG, common.wgsl:67-84
// ---------------------------------------------------------------------------
inline uint32_t pcg_hash(uint32_t v_in) {
  uint32_t v = v_in * 747796405u + 2891336453u;
  uint32_t w = ((v >> ((v >> 28u) + 4u)) ^ v) * 277803737u;
  return (w >> 22u) ^ w;
}
inline uint32_t init_rng(uint32_t px, uint32_t py, uint32_t frame) {
  return pcg_hash(px + pcg_hash(py + pcg_hash(frame)));
}
inline float rand01(uint32_t* s) {
  *s = pcg_hash(*s);
  return (float)(*s) * (1.0f / 4294967296.0f);
}

// ---------------------------------------------------------------------------
// Material, voxel.wgsl:41-51
// ---------------------------------------------------------------------------
struct Mat {
  v3 albedo;
  float emissive;
};
inline Mat unpack_material(uint32_t m) {
  const float inv255 = 1.0f / 255.0f;
  Mat r;
  r.albedo = muls(mk((float)(m & 0xffu), (float)((m >> 8) & 0xffu),
                     (float)((m >> 16) & 0xffu)), inv255);
  r.albedo = mul(r.albedo, r.albedo);
  r.emissive = (float)((m >> 24) & 0xffu) * inv255;
  return r;
}

// ---------------------------------------------------------------------------
// Two-level DDA, voxel.wgsl:65-153.  Non-cubic dims generalise GRID per axis;
// the brick-loop bound 3*BGRID+2 becomes bgx+bgy+bgz+2 (equal for cubes).
// ---------------------------------------------------------------------------
struct TraceOut {
  float t;
  v3 n;
  uint32_t mat;
  uint32_t voxel;
  uint32_t brick_steps, voxel_steps;
  uint8_t face;
};

inline float safe_inv1(float v) {  // voxel.wgsl:65-71
  float s = v >= 0.0f ? 1.0f : -1.0f;
  return s / fmax_(std::fabs(v), 1e-8f);
}
inline int sign_i(float v) { return v > 0.0f ? 1 : (v < 0.0f ? -1 : 0); }

// float-domain clamp of floor(x) into [lo, hi], then to int (WGSL converts
// with saturation first; the result is the same for every finite x)
inline int floor_clamp(float x, int lo, int hi) {
  float f = std::floor(x);
  f = fclamp_(f, (float)lo, (float)hi);
  return (int)f;
}

void trace_impl(const orc_grid* g, v3 roW, v3 rd, float maxTWorld, TraceOut* o) {
  o->t = -1.0f;
  o->mat = 0;
  o->n = mk(-rd.x, -rd.y, -rd.z);
  o->voxel = VRT_VOXEL_MISS;
  o->face = VRT_FACE_MISS;
  o->brick_steps = 0;
  o->voxel_steps = 0;

  const float inv_voxel = 1.0f / g->voxel_size;
  const int gx = g->nx, gy = g->ny, gz = g->nz;
  const int bgx = gx / VRT_BRICK, bgy = gy / VRT_BRICK, bgz = gz / VRT_BRICK;

  v3 ro = muls(roW, inv_voxel);
  float maxT = maxTWorld * inv_voxel;
  v3 invRd = mk(safe_inv1(rd.x), safe_inv1(rd.y), safe_inv1(rd.z));
  int sd[3] = {sign_i(rd.x), sign_i(rd.y), sign_i(rd.z)};
  float ge[3] = {rd.x >= 0.0f ? 1.0f : 0.0f, rd.y >= 0.0f ? 1.0f : 0.0f,
                 rd.z >= 0.0f ? 1.0f : 0.0f};
  float rov[3] = {ro.x, ro.y, ro.z};
  float rdv[3] = {rd.x, rd.y, rd.z};
  float inv[3] = {invRd.x, invRd.y, invRd.z};
  float G[3] = {(float)gx, (float)gy, (float)gz};
  int BG[3] = {bgx, bgy, bgz};

  float tmin[3], tmax[3];
  for (int c = 0; c < 3; ++c) {
    float t0 = (0.0f - rov[c]) * inv[c];
    float t1 = (G[c] - rov[c]) * inv[c];
    tmin[c] = fmin_(t0, t1);
    tmax[c] = fmax_(t0, t1);
  }
  float tEnter = fmax_(fmax_(tmin[0], tmin[1]), fmax_(tmin[2], 0.0f));
  float tExit = fmin_(fmin_(tmax[0], tmax[1]), fmin_(tmax[2], maxT));
  if (tEnter > tExit) return;

  int ax;  // entry face axis (voxel.wgsl:96-98)
  if (tmin[0] >= tmin[1] && tmin[0] >= tmin[2]) ax = 0;
  else if (tmin[1] >= tmin[2]) ax = 1;
  else ax = 2;

  float t = tEnter;
  int bpos[3];
  float tDeltaB[3], tMaxB[3];
  {
    float ts = t + 1e-4f;
    for (int c = 0; c < 3; ++c) {
      float sp = rov[c] + rdv[c] * ts;
      bpos[c] = floor_clamp(sp / 8.0f, 0, BG[c] - 1);
      tDeltaB[c] = std::fabs(inv[c]) * 8.0f;
      tMaxB[c] = (((float)bpos[c] + ge[c]) * 8.0f - rov[c]) * inv[c];
    }
  }

  const int maxIter = bgx + bgy + bgz + 2;
  for (int iter = 0; iter < maxIter; ++iter) {
    if (t > tExit) break;
    uint64_t bi = (uint64_t)bpos[0] + (uint64_t)bpos[1] * bgx +
                  (uint64_t)bpos[2] * (uint64_t)bgx * bgy;
    o->brick_steps++;
    if (g->bricks[bi] != 0u) {
      int bmin[3] = {bpos[0] * 8, bpos[1] * 8, bpos[2] * 8};
      int bmax[3] = {bmin[0] + 7, bmin[1] + 7, bmin[2] + 7};
      int vpos[3];
      float tMaxV[3];
      float ts = t + 1e-4f;
      for (int c = 0; c < 3; ++c) {
        float p = rov[c] + rdv[c] * ts;
        vpos[c] = floor_clamp(p, bmin[c], bmax[c]);
        tMaxV[c] = (((float)vpos[c] + ge[c]) - rov[c]) * inv[c];
      }
      float tCur = t;
      int vax = ax;
      for (int j = 0; j < 3 * VRT_BRICK; ++j) {
        o->voxel_steps++;
        int lx = vpos[0] - bmin[0], ly = vpos[1] - bmin[1], lz = vpos[2] - bmin[2];
        uint32_t bit = (uint32_t)(lx + ly * 8 + lz * 64);
        uint32_t word = g->masks[bi * 16u + (bit >> 5)];
        if ((word & (1u << (bit & 31u))) != 0u) {
          uint64_t vi = (uint64_t)vpos[0] + (uint64_t)vpos[1] * gx +
                        (uint64_t)vpos[2] * (uint64_t)gx * gy;
          uint32_t m = g->vox[vi];
          if (tCur > maxT) return;
          o->t = tCur * g->voxel_size;
          float nn[3];
          for (int c = 0; c < 3; ++c) nn[c] = (c == vax ? -1.0f : -0.0f) * (float)sd[c];
          o->n = mk(nn[0], nn[1], nn[2]);
          o->mat = m;
          o->voxel = (uint32_t)vi;
          if (nn[vax] > 0.0f) o->face = (uint8_t)(2 * vax);
          else if (nn[vax] < 0.0f) o->face = (uint8_t)(2 * vax + 1);
          else o->face = VRT_FACE_NONE;
          return;
        }
        if (tMaxV[0] <= tMaxV[1] && tMaxV[0] <= tMaxV[2]) {
          tCur = tMaxV[0]; tMaxV[0] += std::fabs(inv[0]); vpos[0] += sd[0]; vax = 0;
        } else if (tMaxV[1] <= tMaxV[2]) {
          tCur = tMaxV[1]; tMaxV[1] += std::fabs(inv[1]); vpos[1] += sd[1]; vax = 1;
        } else {
          tCur = tMaxV[2]; tMaxV[2] += std::fabs(inv[2]); vpos[2] += sd[2]; vax = 2;
        }
        if (vpos[0] < bmin[0] || vpos[1] < bmin[1] || vpos[2] < bmin[2] ||
            vpos[0] > bmax[0] || vpos[1] > bmax[1] || vpos[2] > bmax[2])
          break;
      }
    }
    if (tMaxB[0] <= tMaxB[1] && tMaxB[0] <= tMaxB[2]) {
      t = tMaxB[0]; tMaxB[0] += tDeltaB[0]; bpos[0] += sd[0]; ax = 0;
    } else if (tMaxB[1] <= tMaxB[2]) {
      t = tMaxB[1]; tMaxB[1] += tDeltaB[1]; bpos[1] += sd[1]; ax = 1;
    } else {
      t = tMaxB[2]; tMaxB[2] += tDeltaB[2]; bpos[2] += sd[2]; ax = 2;
    }
    if (bpos[0] < 0 || bpos[1] < 0 || bpos[2] < 0 || bpos[0] >= bgx ||
        bpos[1] >= bgy || bpos[2] >= bgz)
      break;
  }
}

// ---------------------------------------------------------------------------
// Camera / sky / sampling, common.wgsl:90-139
// ---------------------------------------------------------------------------
inline void mat_vec4(const float* m, const float v[4], float out[4]) {
  for (int r = 0; r < 4; ++r)
    out[r] = ((m[0 * 4 + r] * v[0] + m[1 * 4 + r] * v[1]) + m[2 * 4 + r] * v[2]) +
             m[3 * 4 + r] * v[3];
}

v3 camera_ray(const vrt_uniforms* u, float uvx, float uvy) {
  float ndcx = uvx * 2.0f - 1.0f;
  float ndcy = 1.0f - uvy * 2.0f;
  float v[4] = {ndcx, ndcy, 1.0f, 1.0f};
  float far[4];
  mat_vec4(u->inv_view_proj, v, far);
  v3 p = mk(far[0] / far[3], far[1] / far[3], far[2] / far[3]);
  return normalize3(sub(p, mk(u->cam_pos[0], u->cam_pos[1], u->cam_pos[2])));
}

v3 sky_color(const vrt_uniforms* u, v3 rd, bool disc) {
  float up = fclamp_(rd.y, -1.0f, 1.0f);
  float tt = pow_own(fmax_(up, 0.0f), 0.55f);
  v3 sky = mix3(mk(0.55f, 0.68f, 0.85f), mk(0.15f, 0.32f, 0.62f), tt);
  if (up < 0.0f) sky = mix3(sky, mk(0.28f, 0.26f, 0.24f), fmin_(1.0f, -up * 4.0f));
  v3 sd = mk(u->sun_dir[0], u->sun_dir[1], u->sun_dir[2]);
  float sunAmt = fmax_(dot3(rd, sd), 0.0f);
  float p8 = pow8(sunAmt);
  sky = add(sky, muls(muls(mk(1.0f, 0.75f, 0.5f), p8), 0.2f));
  if (disc && sunAmt > u->sun_dir[3]) {
    float solidAngle = TWO_PI_F * (1.0f - u->sun_dir[3]);
    sky = add(sky, divs(mk(u->sun_radiance[0], u->sun_radiance[1], u->sun_radiance[2]),
                        fmax_(solidAngle, 1e-4f)));
  }
  return muls(sky, u->sun_radiance[3]);
}

v3 sample_sun_dir(const vrt_uniforms* u, uint32_t* rng) {
  float r0 = rand01(rng);
  float r1 = rand01(rng);
  float cosT = mixf(1.0f, u->sun_dir[3], r0);
  float sinT = std::sqrt(fmax_(0.0f, 1.0f - cosT * cosT));
  float phi = TWO_PI_F * r1;
  v3 w = mk(u->sun_dir[0], u->sun_dir[1], u->sun_dir[2]);
  v3 a = std::fabs(w.x) > 0.9f ? mk(0.0f, 1.0f, 0.0f) : mk(1.0f, 0.0f, 0.0f);
  v3 tang = normalize3(cross3(a, w));
  v3 bitan = cross3(w, tang);
  float s, c;
  sincos_own(phi, &s, &c);
  v3 d = add(add(muls(tang, c * sinT), muls(bitan, s * sinT)), muls(w, cosT));
  return normalize3(d);
}

v3 cosine_hemisphere(v3 n, uint32_t* rng) {
  float r0 = rand01(rng);
  float r1 = rand01(rng);
  float phi = TWO_PI_F * r0;
  float sr = std::sqrt(r1);
  v3 a = std::fabs(n.x) > 0.9f ? mk(0.0f, 1.0f, 0.0f) : mk(1.0f, 0.0f, 0.0f);
  v3 tang = normalize3(cross3(a, n));
  v3 bitan = cross3(n, tang);
  float s, c;
  sincos_own(phi, &s, &c);
  v3 d = add(add(muls(tang, c * sr), muls(bitan, s * sr)),
             muls(n, std::sqrt(fmax_(0.0f, 1.0f - r1))));
  return normalize3(d);
}

// ---------------------------------------------------------------------------
// Base path tracer, pathtrace.wgsl:248-330, one pixel, one sample.
// ---------------------------------------------------------------------------
struct PixelResult {
  v3 radiance;      // demodulated (or sky)
  v3 albedo_sqrt;   // value written to the rgba8unorm albedo target
  float gbuf[4];
};

inline void count(orc_counters* c, const TraceOut& h) {
  c->rays++;
  c->brick_steps += h.brick_steps;
  c->voxel_steps += h.voxel_steps;
  if (h.t >= 0.0f) c->hits++;
}

PixelResult shade_pixel(const orc_grid* g, const vrt_uniforms* u, float fclamp,
                        uint32_t px, uint32_t py, uint32_t frame,
                        orc_counters* cnt) {
  PixelResult res;
  const float w = u->params1[0], h = u->params1[1];
  uint32_t rng = init_rng(px, py, frame);
  float jx = rand01(&rng) - 0.5f;
  float jy = rand01(&rng) - 0.5f;
  float uvx = (((float)px + 0.5f) + jx) / (float)(uint32_t)w;
  float uvy = (((float)py + 0.5f) + jy) / (float)(uint32_t)h;
  v3 ro = mk(u->cam_pos[0], u->cam_pos[1], u->cam_pos[2]);
  v3 rd = camera_ray(u, uvx, uvy);

  TraceOut h0;
  trace_impl(g, ro, rd, 1e4f, &h0);
  count(cnt, h0);
  const bool indirect = (u->params0[2] & VRT_FLAG_INDIRECT) != 0u;
  if (h0.t < 0.0f) {
    res.radiance = indirect ? mk(0.0f, 0.0f, 0.0f) : sky_color(u, rd, true);
    res.albedo_sqrt = mk(1.0f, 1.0f, 1.0f);
    res.gbuf[0] = -rd.x; res.gbuf[1] = -rd.y; res.gbuf[2] = -rd.z; res.gbuf[3] = -1.0f;
    return res;
  }
  Mat mat1 = unpack_material(h0.mat);
  v3 x1 = add(ro, muls(rd, h0.t));
  v3 n1 = h0.n;
  res.albedo_sqrt = mk(std::sqrt(mat1.albedo.x), std::sqrt(mat1.albedo.y),
                       std::sqrt(mat1.albedo.z));
  res.gbuf[0] = n1.x; res.gbuf[1] = n1.y; res.gbuf[2] = n1.z; res.gbuf[3] = h0.t;

  v3 radiance = mk(0.0f, 0.0f, 0.0f);
  v3 tp = mk(1.0f, 1.0f, 1.0f);
  v3 pos = x1, n = n1;
  Mat mat = mat1;
  const uint32_t nb = u->params0[1];
  const v3 sunRad = mk(u->sun_radiance[0], u->sun_radiance[1], u->sun_radiance[2]);
  for (uint32_t bounce = 0; bounce <= nb; ++bounce) {
    const bool direct = !(indirect && bounce == 0u);
    if (mat.emissive > 0.0f && direct)
      radiance = add(radiance, muls(mul(tp, mat.albedo), mat.emissive * EMISSIVE_SCALE));
    v3 surf = add(pos, muls(n, 1e-3f));
    v3 l = sample_sun_dir(u, &rng);
    float ndl = dot3(n, l);
    if (ndl > 0.0f && l.y > 0.0f && direct) {
      TraceOut hs;
      trace_impl(g, surf, l, 1e4f, &hs);
      count(cnt, hs);
      if (!(hs.t >= 0.0f))
        radiance = add(radiance, mul(muls(muls(mul(tp, mat.albedo), INV_PI_F), ndl), sunRad));
    }
    if (bounce == nb) break;
    rd = cosine_hemisphere(n, &rng);
    tp = mul(tp, mat.albedo);
    if (bounce >= 2u) {
      float p = fclamp_(fmax_(tp.x, fmax_(tp.y, tp.z)), 0.05f, 0.95f);
      if (rand01(&rng) > p) break;
      tp = divs(tp, p);
    }
    TraceOut hb;
    trace_impl(g, surf, rd, 1e4f, &hb);
    count(cnt, hb);
    if (hb.t < 0.0f) {
      radiance = add(radiance, mul(tp, sky_color(u, rd, false)));
      break;
    }
    pos = add(surf, muls(rd, hb.t));
    n = hb.n;
    mat = unpack_material(hb.mat);
  }
  if (!(fclamp <= 0.0f))
    radiance = mk(fmin_(radiance.x, fclamp), fmin_(radiance.y, fclamp), fmin_(radiance.z, fclamp));
  v3 den = mk(fmax_(mat1.albedo.x, 1e-3f), fmax_(mat1.albedo.y, 1e-3f), fmax_(mat1.albedo.z, 1e-3f));
  res.radiance = divv(radiance, den);
  return res;
}

void pathtrace_pixel(const orc_grid* g, const vrt_uniforms* u, float fclamp,
                     uint32_t spp, uint32_t px, uint32_t py, uint16_t* radiance,
                     uint8_t* albedo, float* gbuf, orc_counters* cnt) {
  const uint32_t w = (uint32_t)u->params1[0];
  const uint64_t i = (uint64_t)py * w + px;
  if (spp < 1) spp = 1;
  v3 acc = mk(0.0f, 0.0f, 0.0f);
  PixelResult first{};
  for (uint32_t s = 0; s < spp; ++s) {
    PixelResult r = shade_pixel(g, u, fclamp, px, py, u->params0[0] * spp + s, cnt);
    if (s == 0) first = r;
    acc = add(acc, r.radiance);
  }
  v3 out = divs(acc, (float)spp);
  if (radiance) {
    radiance[i * 4 + 0] = f32_to_f16(out.x);
    radiance[i * 4 + 1] = f32_to_f16(out.y);
    radiance[i * 4 + 2] = f32_to_f16(out.z);
    radiance[i * 4 + 3] = f32_to_f16(1.0f);
  }
  if (albedo) {
    albedo[i * 4 + 0] = f32_to_unorm8(first.albedo_sqrt.x);
    albedo[i * 4 + 1] = f32_to_unorm8(first.albedo_sqrt.y);
    albedo[i * 4 + 2] = f32_to_unorm8(first.albedo_sqrt.z);
    albedo[i * 4 + 3] = 255;
  }
  if (gbuf)
    for (int c = 0; c < 4; ++c) gbuf[i * 4 + c] = first.gbuf[c];
}

inline float load_tex(const void* p, uint32_t fmt, uint64_t idx) {
  if (fmt == VRT_ACCUM_RGBA32F) return ((const float*)p)[idx];
  return f16_to_f32(((const uint16_t*)p)[idx]);
}
inline void store_tex(void* p, uint32_t fmt, uint64_t idx, float v) {
  if (fmt == VRT_ACCUM_RGBA32F) ((float*)p)[idx] = v;
  else ((uint16_t*)p)[idx] = f32_to_f16(v);
}

inline void add_counters(orc_counters* dst, const orc_counters& c) {
  if (!dst) return;
#pragma omp atomic
  dst->rays += c.rays;
#pragma omp atomic
  dst->brick_steps += c.brick_steps;
#pragma omp atomic
  dst->voxel_steps += c.voxel_steps;
#pragma omp atomic
  dst->hits += c.hits;
}

inline int nthreads(int t) { return t > 0 ? t : omp_get_max_threads(); }

}  // namespace

extern "C" {

int orc_build_bricks(const uint32_t* vox, int nx, int ny, int nz, uint32_t* bricks,
                     uint32_t* masks) {
  if (nx <= 0 || ny <= 0 || nz <= 0 || nx % 8 || ny % 8 || nz % 8) return -1;
  const int bx = nx / 8, by = ny / 8, bz = nz / 8;
  const uint64_t nb = (uint64_t)bx * by * bz;
  std::memset(masks, 0, nb * 16 * sizeof(uint32_t));
  for (int z = 0; z < bz; ++z)
    for (int y = 0; y < by; ++y)
      for (int x = 0; x < bx; ++x) {
        uint32_t occ = 0;
        uint64_t bi = (uint64_t)x + (uint64_t)y * bx + (uint64_t)z * bx * by;
        for (int lz = 0; lz < 8; ++lz)
          for (int ly = 0; ly < 8; ++ly)
            for (int lx = 0; lx < 8; ++lx) {
              uint64_t vi = (uint64_t)(x * 8 + lx) + (uint64_t)(y * 8 + ly) * nx +
                            (uint64_t)(z * 8 + lz) * (uint64_t)nx * ny;
              if (vox[vi] != 0u) {
                occ = 1;
                uint32_t bit = (uint32_t)(lx + ly * 8 + lz * 64);
                masks[bi * 16 + (bit >> 5)] |= 1u << (bit & 31u);
              }
            }
        bricks[bi] = occ;
      }
  return 0;
}

void orc_trace(const orc_grid* g, const float ro[3], const float rd[3],
               float max_t_world, vrt_hit* hit, float* n_out, uint32_t* mat_out) {
  TraceOut o;
  trace_impl(g, mk(ro[0], ro[1], ro[2]), mk(rd[0], rd[1], rd[2]), max_t_world, &o);
  if (hit) {
    std::memset(hit, 0, sizeof(*hit));
    hit->t = o.t;
    hit->voxel = o.voxel;
    hit->brick_steps = (uint16_t)o.brick_steps;
    hit->voxel_steps = (uint16_t)o.voxel_steps;
    hit->face = o.face;
  }
  if (n_out) { n_out[0] = o.n.x; n_out[1] = o.n.y; n_out[2] = o.n.z; }
  if (mat_out) *mat_out = o.mat;
}

void orc_trace_rays(const orc_grid* g, const float* rays, uint64_t n, float max_t_world,
                    vrt_hit* hits, int threads) {
#pragma omp parallel for schedule(dynamic, 4096) num_threads(nthreads(threads))
  for (int64_t i = 0; i < (int64_t)n; ++i)
    orc_trace(g, rays + i * 6, rays + i * 6 + 3, max_t_world, hits + i, nullptr, nullptr);
}

uint32_t orc_pcg_hash(uint32_t v) { return pcg_hash(v); }

void orc_rand_stream(uint32_t px, uint32_t py, uint32_t frame, int n, uint32_t* states,
                     float* values) {
  uint32_t s = init_rng(px, py, frame);
  for (int i = 0; i < n; ++i) {
    float v = rand01(&s);
    if (states) states[i] = s;
    if (values) values[i] = v;
  }
}

float orc_sin(float x) { float s, c; sincos_own(x, &s, &c); return s; }
float orc_cos(float x) { float s, c; sincos_own(x, &s, &c); return c; }
float orc_exp2(float x) { return exp2_own(x); }
float orc_log2(float x) { return log2_own(x); }
float orc_exp(float x) { return exp_own(x); }
float orc_pow(float x, float y) { return pow_own(x, y); }
uint16_t orc_f32_to_f16(float f) { return f32_to_f16(f); }
float orc_f16_to_f32(uint16_t h) { return f16_to_f32(h); }

void orc_camera_matrices(const double pos[3], double yaw, double pitch, double fov_y,
                         double aspect, double near_z, double far_z, float view_proj[16],
                         float inv_view_proj[16]) {
  // math.js:96-107 cameraBasis
  double cp = std::cos(pitch), sp = std::sin(pitch);
  double cy = std::cos(yaw), sy = std::sin(yaw);
  double fw[3] = {sy * cp, sp, cy * cp};
  double rt[3] = {-cy, 0.0, sy};
  double up[3] = {rt[1] * fw[2] - rt[2] * fw[1], rt[2] * fw[0] - rt[0] * fw[2],
                  rt[0] * fw[1] - rt[1] * fw[0]};
  double z[3] = {-fw[0], -fw[1], -fw[2]};
  auto dotd = [](const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
  // Float32Array contents (math.js:73-131)
  float proj[16] = {0}, iproj[16] = {0}, c2w[16], w2c[16];
  double f = 1.0 / std::tan(fov_y / 2.0);
  proj[0] = (float)(f / aspect);
  proj[5] = (float)f;
  proj[10] = (float)(far_z / (near_z - far_z));
  proj[11] = -1.0f;
  proj[14] = (float)((near_z * far_z) / (near_z - far_z));
  iproj[0] = (float)(aspect / f);
  iproj[5] = (float)(1.0 / f);
  iproj[11] = (float)((near_z - far_z) / (near_z * far_z));
  iproj[14] = -1.0f;
  iproj[15] = (float)(1.0 / near_z);
  double c2wd[16] = {rt[0], rt[1], rt[2], 0, up[0], up[1], up[2], 0,
                     -fw[0], -fw[1], -fw[2], 0, pos[0], pos[1], pos[2], 1};
  double w2cd[16] = {rt[0], up[0], z[0], 0, rt[1], up[1], z[1], 0, rt[2], up[2], z[2], 0,
                     -dotd(rt, pos), -dotd(up, pos), -dotd(z, pos), 1};
  for (int i = 0; i < 16; ++i) { c2w[i] = (float)c2wd[i]; w2c[i] = (float)w2cd[i]; }
  // math.js:60-70 mat4Mul: double accumulation of f32 inputs, f32 store
  auto mul4 = [](const float* a, const float* b, float* out) {
    for (int c = 0; c < 4; ++c)
      for (int r = 0; r < 4; ++r) {
        double s = 0.0;
        for (int k = 0; k < 4; ++k) s += (double)a[k * 4 + r] * (double)b[c * 4 + k];
        out[c * 4 + r] = (float)s;
      }
  };
  mul4(proj, w2c, view_proj);
  mul4(c2w, iproj, inv_view_proj);
}

float orc_exp_w(float x) { return exp_w(x); }
void orc_exp_w_table(uint64_t start, uint64_t stride, uint64_t n, float* out) {
  for (uint64_t i = 0; i < n; ++i) {
    const uint32_t b = (uint32_t)(start + i * stride);
    float x;
    std::memcpy(&x, &b, 4);
    out[i] = exp_w(x);
  }
}
void orc_sub_texel(float tc, float* fl, float* fr) { sub_texel(tc, fl, fr); }

void orc_camera_ray(const vrt_uniforms* u, float uvx, float uvy, float out[3]) {
  v3 r = camera_ray(u, uvx, uvy);
  out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

void orc_sky_color(const vrt_uniforms* u, const float rd[3], int disc, float out[3]) {
  v3 r = sky_color(u, mk(rd[0], rd[1], rd[2]), disc != 0);
  out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

void orc_sample_sun_dir(const vrt_uniforms* u, uint32_t* rng, float out[3]) {
  v3 r = sample_sun_dir(u, rng);
  out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

void orc_cosine_hemisphere(const float n[3], uint32_t* rng, float out[3]) {
  v3 r = cosine_hemisphere(mk(n[0], n[1], n[2]), rng);
  out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

void orc_pathtrace(const orc_grid* g, const vrt_uniforms* u, float fclamp, uint32_t spp,
                   int y0, int y1, uint16_t* radiance, uint8_t* albedo, float* gbuf,
                   orc_counters* counters, int threads) {
  const int w = (int)u->params1[0];
#pragma omp parallel num_threads(nthreads(threads))
  {
    orc_counters local{0, 0, 0, 0};
#pragma omp for schedule(dynamic, 8)
    for (int y = y0; y < y1; ++y)
      for (int x = 0; x < w; ++x)
        pathtrace_pixel(g, u, fclamp, spp, (uint32_t)x, (uint32_t)y, radiance, albedo, gbuf,
                        &local);
    add_counters(counters, local);
  }
}

void orc_pathtrace_pixels(const orc_grid* g, const vrt_uniforms* u, float fclamp,
                          uint32_t spp, const uint32_t* pix, uint64_t npix,
                          uint16_t* radiance, uint8_t* albedo, float* gbuf,
                          orc_counters* counters, int threads) {
  const uint32_t w = (uint32_t)u->params1[0];
#pragma omp parallel num_threads(nthreads(threads))
  {
    orc_counters local{0, 0, 0, 0};
#pragma omp for schedule(dynamic, 64)
    for (int64_t k = 0; k < (int64_t)npix; ++k)
      pathtrace_pixel(g, u, fclamp, spp, pix[k] % w, pix[k] / w, radiance, albedo, gbuf,
                      &local);
    add_counters(counters, local);
  }
}

// temporal.wgsl:17-77
void orc_temporal(const vrt_uniforms* u, const uint16_t* radiance, const float* gbuf_cur,
                  const void* accum_prev, const float* gbuf_prev, uint32_t format,
                  void* accum_out, int threads) {
  const int w = (int)u->params1[0], h = (int)u->params1[1];
  const bool temporalOn = (u->params0[2] & VRT_FLAG_TEMPORAL) != 0u;
#pragma omp parallel for schedule(dynamic, 8) num_threads(nthreads(threads))
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) {
      const uint64_t i = (uint64_t)y * w + x;
      v3 cur = mk(f16_to_f32(radiance[i * 4 + 0]), f16_to_f32(radiance[i * 4 + 1]),
                  f16_to_f32(radiance[i * 4 + 2]));
      const float* g = gbuf_cur + i * 4;
      if (g[3] < 0.0f || !temporalOn) {
        store_tex(accum_out, format, i * 4 + 0, cur.x);
        store_tex(accum_out, format, i * 4 + 1, cur.y);
        store_tex(accum_out, format, i * 4 + 2, cur.z);
        store_tex(accum_out, format, i * 4 + 3, 1.0f);
        continue;
      }
      float uvx = ((float)x + 0.5f) / (float)w;
      float uvy = ((float)y + 0.5f) / (float)h;
      v3 cp = mk(u->cam_pos[0], u->cam_pos[1], u->cam_pos[2]);
      v3 worldPos = add(cp, muls(camera_ray(u, uvx, uvy), g[3]));
      float wp4[4] = {worldPos.x, worldPos.y, worldPos.z, 1.0f};
      float prevClip[4];
      mat_vec4(u->prev_view_proj, wp4, prevClip);
      v3 color = cur;
      float history = 1.0f;
      if (prevClip[3] > 0.0f) {
        float ndcx = prevClip[0] / prevClip[3];
        float ndcy = prevClip[1] / prevClip[3];
        float pux = ndcx * 0.5f + 0.5f;
        float puy = (-ndcy) * 0.5f + 0.5f;
        if (pux > 0.0f && puy > 0.0f && pux < 1.0f && puy < 1.0f) {
          int ppx = (int)(pux * (float)w);
          int ppy = (int)(puy * (float)h);
          if (ppx > w - 1) ppx = w - 1;
          if (ppy > h - 1) ppy = h - 1;
          const float* pg = gbuf_prev + ((uint64_t)ppy * w + ppx) * 4;
          v3 dwp = sub(worldPos, mk(u->prev_cam_pos[0], u->prev_cam_pos[1], u->prev_cam_pos[2]));
          float expectedT = std::sqrt(dot3(dwp, dwp));
          bool depthOk = pg[3] > 0.0f && std::fabs(pg[3] - expectedT) < 0.08f * expectedT + 0.02f;
          bool normalOk = dot3(mk(pg[0], pg[1], pg[2]), mk(g[0], g[1], g[2])) > 0.85f;
          if (depthOk && normalOk) {
            // textureSampleLevel with a linear, clamp-to-edge sampler, in FP32;
            // the texel coordinate carries 8 fractional bits, as in a texture
            // unit's fixed-point filter (sub_texel)
            float flx, fly, fx, fy;
            sub_texel(pux * (float)w - 0.5f, &flx, &fx);
            sub_texel(puy * (float)h - 0.5f, &fly, &fy);
            int x0 = (int)flx, y0 = (int)fly;
            int x1 = x0 + 1, y1 = y0 + 1;
            x0 = x0 < 0 ? 0 : (x0 > w - 1 ? w - 1 : x0);
            x1 = x1 < 0 ? 0 : (x1 > w - 1 ? w - 1 : x1);
            y0 = y0 < 0 ? 0 : (y0 > h - 1 ? h - 1 : y0);
            y1 = y1 < 0 ? 0 : (y1 > h - 1 ? h - 1 : y1);
            float prev[4];
            for (int c = 0; c < 4; ++c) {
              float t00 = load_tex(accum_prev, format, ((uint64_t)y0 * w + x0) * 4 + c);
              float t10 = load_tex(accum_prev, format, ((uint64_t)y0 * w + x1) * 4 + c);
              float t01 = load_tex(accum_prev, format, ((uint64_t)y1 * w + x0) * 4 + c);
              float t11 = load_tex(accum_prev, format, ((uint64_t)y1 * w + x1) * 4 + c);
              prev[c] = mixf(mixf(t00, t10, fx), mixf(t01, t11, fx), fy);
            }
            float histCap = u->params4[1];
            history = fmin_(prev[3], histCap) + 1.0f;
            float alpha = 1.0f / history;
            color = mix3(mk(prev[0], prev[1], prev[2]), cur, alpha);
          }
        }
      }
      store_tex(accum_out, format, i * 4 + 0, color.x);
      store_tex(accum_out, format, i * 4 + 1, color.y);
      store_tex(accum_out, format, i * 4 + 2, color.z);
      store_tex(accum_out, format, i * 4 + 3, history);
    }
}

// atrous.wgsl:20-83
void orc_atrous(const vrt_uniforms* u, const void* src, uint32_t src_format, const float* gbuf,
                int step, uint16_t* dst, int threads) {
  const int w = (int)u->params1[0], h = (int)u->params1[1];
  const float K[3] = {0.375f, 0.25f, 0.0625f};
#pragma omp parallel for schedule(dynamic, 8) num_threads(nthreads(threads))
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) {
      const uint64_t i = (uint64_t)y * w + x;
      float center[4];
      for (int c = 0; c < 4; ++c) center[c] = load_tex(src, src_format, i * 4 + c);
      const float* g = gbuf + i * 4;
      if (g[3] < 0.0f) {
        for (int c = 0; c < 4; ++c) dst[i * 4 + c] = f32_to_f16(center[c]);
        continue;
      }
      float history = fmax_(center[3], 1.0f);
      const v3 lw = mk(0.2126f, 0.7152f, 0.0722f);
      float lumC = dot3(mk(center[0], center[1], center[2]), lw);
      float relax = 2.0f / std::sqrt(history);
      float sigmaL = fmax_(0.08f, relax);
      v3 sum = muls(mk(center[0], center[1], center[2]), K[0] * K[0]);
      float wSum = K[0] * K[0];
      v3 gn = mk(g[0], g[1], g[2]);
      for (int dy = -2; dy <= 2; ++dy)
        for (int dx = -2; dx <= 2; ++dx) {
          if (dx == 0 && dy == 0) continue;
          int qx = x + dx * step, qy = y + dy * step;
          if (qx < 0 || qy < 0 || qx >= w || qy >= h) continue;
          const uint64_t j = (uint64_t)qy * w + qx;
          const float* gq = gbuf + j * 4;
          if (gq[3] < 0.0f) continue;
          v3 cq = mk(load_tex(src, src_format, j * 4 + 0), load_tex(src, src_format, j * 4 + 1),
                     load_tex(src, src_format, j * 4 + 2));
          float wKernel = K[dx < 0 ? -dx : dx] * K[dy < 0 ? -dy : dy];
          float wNormal = pow32(fmax_(dot3(gn, mk(gq[0], gq[1], gq[2])), 0.0f));
          // WGSL's divisions (2.5 ulps allowed) as products with the
          // correctly rounded reciprocals of the per-pixel divisors
          float wDepth = exp_w(-(std::fabs(g[3] - gq[3]) * (1.0f / (0.05f * g[3] + 0.01f))));
          float wLum = exp_w(-(std::fabs(lumC - dot3(cq, lw)) * (1.0f / sigmaL)));
          float wt = ((wKernel * wNormal) * wDepth) * wLum;
          sum = add(sum, muls(cq, wt));
          wSum += wt;
        }
      v3 o = divs(sum, fmax_(wSum, 1e-6f));
      dst[i * 4 + 0] = f32_to_f16(o.x);
      dst[i * 4 + 1] = f32_to_f16(o.y);
      dst[i * 4 + 2] = f32_to_f16(o.z);
      dst[i * 4 + 3] = f32_to_f16(center[3]);
    }
}

void orc_hdr(int w, int h, const void* illum, uint32_t illum_format, const uint8_t* albedo,
             float* hdr) {
  for (int64_t i = 0; i < (int64_t)w * h; ++i)
    for (int c = 0; c < 3; ++c) {
      double ir = (double)load_tex(illum, illum_format, (uint64_t)i * 4 + c);
      double a = (double)albedo[i * 4 + c] / 255.0;
      hdr[i * 3 + c] = (float)(ir * a * a);
    }
}

void orc_present(int w, int h, const void* illum, uint32_t illum_format, const uint8_t* albedo, float exposure,
                 uint8_t* rgba) {
  const float inv_gamma = (float)(1.0 / 2.2);  // WGSL folds 1.0 / 2.2 as an abstract float
  for (int64_t i = 0; i < (int64_t)w * h; ++i) {
    for (int c = 0; c < 3; ++c) {
      const float s = (float)albedo[i * 4 + c] / 255.0f;  // rgba8unorm sample
      const float x = (load_tex(illum, illum_format, (uint64_t)i * 4 + c) * (s * s)) * exposure;
      const float num = x * (2.51f * x + 0.03f);
      const float den = x * (2.43f * x + 0.59f) + 0.14f;
      const float t = fclamp_(num / den, 0.0f, 1.0f);
      rgba[i * 4 + c] = f32_to_unorm8(pow_own(t, inv_gamma));
    }
    rgba[i * 4 + 3] = 255;
  }
}

}  // extern "C"
