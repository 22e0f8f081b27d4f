// vrt_device.h -- device-side arithmetic for libvoxelrt_hip (gfx950).
//
// Every function here fixes its operation order so that results are
// reproducible bit for bit on any IEEE-754 FP32 implementation compiled
// without fused multiply-add (-ffp-contract=off):
//   min(a,b) = b<a ? b : a,  max(a,b) = a<b ? b : a (first operand on ties,
//   which also fixes the sign of zero), dot = (x*x' + y*y') + z*z',
//   normalize(v) = v / sqrt(dot(v,v)), mix(a,b,t) = a*(1-t) + b*t,
//   IEEE division and square root (hipcc's default
//   -fhip-fp32-correctly-rounded-divide-sqrt).
// sin/cos/exp2/log2 are implemented here from textbook methods (Cody-Waite
// reduction, Taylor series with 1/n! coefficients, atanh series for log)
// instead of the vendor math library, whose last bits are not specified.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// Host and device: the host instantiation lets CPU tests run the same DDA code.
#define VRT_D __host__ __device__ __forceinline__

namespace vrt {

struct f3 {
  float x, y, z;
};
VRT_D f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
VRT_D f3 operator+(f3 a, f3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
VRT_D f3 operator-(f3 a, f3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
VRT_D f3 operator*(f3 a, f3 b) { return mk3(a.x * b.x, a.y * b.y, a.z * b.z); }
VRT_D f3 operator*(f3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
VRT_D f3 operator/(f3 a, f3 b) { return mk3(a.x / b.x, a.y / b.y, a.z / b.z); }
VRT_D f3 operator/(f3 a, float s) { return mk3(a.x / s, a.y / s, a.z / s); }

// Lanes of the calling wave for which p holds (bit l = lane l).  The host
// instantiation (CPU tests) runs one lane at a time, which is lane 0.
VRT_D uint64_t wave_ballot(bool p) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __ballot(p);
#else
  return p ? 1ull : 0ull;
#endif
}

VRT_D float vmin(float a, float b) { return b < a ? b : a; }

VRT_D float vmax(float a, float b) { return a < b ? b : a; }
VRT_D float vclamp(float x, float lo, float hi) { return vmin(vmax(x, lo), hi); }
VRT_D float vabs(float x) { return __builtin_fabsf(x); }
VRT_D float vfloor(float x) { return __builtin_floorf(x); }
VRT_D float vsqrt(float x) { return __builtin_sqrtf(x); }
VRT_D float dot(f3 a, f3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
VRT_D f3 normalize(f3 v) { return v / vsqrt(dot(v, v)); }
VRT_D f3 cross(f3 a, f3 b) {
  return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
VRT_D float mix(float a, float b, float t) { return a * (1.0f - t) + b * t; }
VRT_D f3 mix(f3 a, f3 b, float t) { return mk3(mix(a.x, b.x, t), mix(a.y, b.y, t), mix(a.z, b.z, t)); }

// common.wgsl:7-11
constexpr float PI_F = 3.14159265358979f;
constexpr float TWO_PI_F = 2.0f * PI_F;
constexpr float INV_PI_F = 0.31830988618f;
constexpr float EMISSIVE_SCALE = 24.0f;

// ---- transcendentals -------------------------------------------------------
// Reduction x = k*(pi/2) + r with pi/2 split into three floats whose leading
// part has 8 significant bits (k*HI exact for |k| < 2^16).
VRT_D void sincos_r(float x, float* s, float* c) {
  const float TWO_OVER_PI = 0x1.45f306p-1f;
  const float HI = 0x1.92p+0f, MID = 0x1.fb6p-12f, LO = -0x1.777a5cp-25f;
  float k = vfloor(x * TWO_OVER_PI + 0.5f);
  float r = ((x - k * HI) - k * MID) - k * LO;
  float r2 = r * r;
  // sin r = r + r^3 (-1/3! + r^2 (1/5! + r^2 (-1/7! + r^2/9!)))
  float sp = (((float)(1.0 / 362880.0) * r2 + (float)(-1.0 / 5040.0)) * r2 +
              (float)(1.0 / 120.0)) * r2 + (float)(-1.0 / 6.0);
  float sn = r + (r * r2) * sp;
  // cos r = 1 + r^2 (-1/2 + r^2 (1/4! + r^2 (-1/6! + r^2 (1/8! - r^2/10!))))
  float cp = ((((float)(-1.0 / 3628800.0) * r2 + (float)(1.0 / 40320.0)) * r2 +
               (float)(-1.0 / 720.0)) * r2 + (float)(1.0 / 24.0)) * r2 - 0.5f;
  float cs = 1.0f + r2 * cp;
  int q = ((int)k) & 3;
  float ss = (q & 1) ? cs : sn;
  float cc = (q & 1) ? sn : cs;
  *s = (q & 2) ? -ss : ss;
  *c = ((q + 1) & 2) ? -cc : cc;
}

// 2^x = 2^k * e^(f ln2), k = round(x), Taylor series to degree 7 in
// fused multiply-add Horner form (each step one correctly rounded fma).
// Branch-free: the special cases (NaN -> NaN, x <= -126 -> 0, x >= 128 ->
// inf) are selected over the polynomial of an in-range stand-in, so lanes of
// a wave never diverge here; in-range arguments take exactly the operations
// of the scalar definition (the oracle's exp2_own).
VRT_D float exp2_r(float x) {
  const bool nan = x != x, lo = !(x > -126.0f), hi = x >= 128.0f;
  const float xs = (nan || lo || hi) ? 0.0f : x;
  const float LN2 = 0x1.62e43p-1f;
  const float k = vfloor(xs + 0.5f);
  const float f = xs - k;
  const float y = f * LN2;
  float p = __builtin_fmaf((float)(1.0 / 5040.0), y, (float)(1.0 / 720.0));
  p = __builtin_fmaf(p, y, (float)(1.0 / 120.0));
  p = __builtin_fmaf(p, y, (float)(1.0 / 24.0));
  p = __builtin_fmaf(p, y, (float)(1.0 / 6.0));
  p = __builtin_fmaf(p, y, 0.5f);
  p = __builtin_fmaf(p, y, 1.0f);
  p = __builtin_fmaf(p, y, 1.0f);
  int ki = (int)k;
  const bool top = ki > 127;  // 2^128 = 2 * 2^127
  p = top ? p * 2.0f : p;
  ki = top ? ki - 1 : ki;
  float r = p * __builtin_bit_cast(float, (uint32_t)(ki + 127) << 23);
  r = hi ? __builtin_inff() : r;
  r = lo ? 0.0f : r;
  return nan ? x : r;
}

// log2 x = e + ln(m)/ln2, ln m = 2 atanh(s), s = (m-1)/(m+1), m in
// (sqrt(1/2), sqrt(2)].
VRT_D float log2_r(float x) {
  if (x != x || x < 0.0f) return __builtin_nanf("");
  if (x == 0.0f) return -__builtin_inff();
  if (x == __builtin_inff()) return x;
  int e = 0;
  uint32_t b = __builtin_bit_cast(uint32_t, x);
  if (b < 0x00800000u) {
    x = x * 8388608.0f;
    b = __builtin_bit_cast(uint32_t, x);
    e = -23;
  }
  e += (int)((b >> 23) & 0xffu) - 127;
  float m = __builtin_bit_cast(float, (b & 0x007fffffu) | 0x3f800000u);
  if (m > 1.41421356f) {
    m = m * 0.5f;
    e += 1;
  }
  float s = (m - 1.0f) / (m + 1.0f);
  float s2 = s * s;
  float p = ((((float)(2.0 / 11.0) * s2 + (float)(2.0 / 9.0)) * s2 + (float)(2.0 / 7.0)) * s2 +
             (float)(2.0 / 5.0)) * s2 + (float)(2.0 / 3.0);
  p = p * s2 + 2.0f;
  return (float)e + (s * p) * 0x1.715476p+0f;
}

VRT_D float pow_r(float x, float y) {
  if (x == 0.0f) return y > 0.0f ? 0.0f : (y == 0.0f ? 1.0f : __builtin_inff());
  return exp2_r(y * log2_r(x));
}
VRT_D float exp_r(float x) { return exp2_r(x * 0x1.715476p+0f); }

// e^x for the a-trous weights, whose arguments are <= 0 or NaN:
// z = max(x log2(e), -160) (IEEE maximum: a NaN argument becomes -160 and
// the weight 0), k = rint(z) in [-160, 0], f = z - k in [-1/2, 1/2] (exact),
// e^x = 2^k 2^f with 2^f from a degree-5 polynomial fitted for relative
// error on [-1/2, 1/2] (constant term 1, so e^0 = 1 exactly) in fused
// multiply-add steps, scaled by ldexp: within 2.1 ulps of 2^f, plus the
// rounding of x log2(e) (|x| 2^-24 relative), inside the 3 + 2|x| ulps WGSL
// allows exp (11 instructions).
VRT_D float exp_w(float x) {
  const float z = __builtin_fmaxf(x * 0x1.715476p+0f, -160.0f);
  const float k = __builtin_rintf(z);
  const float f = z - k;
  float p = __builtin_fmaf(0x1.5bb9fep-10f, f, 0x1.3cea88p-7f);
  p = __builtin_fmaf(p, f, 0x1.c6b754p-5f);
  p = __builtin_fmaf(p, f, 0x1.ebf9bcp-3f);
  p = __builtin_fmaf(p, f, 0x1.62e42ap-1f);
  p = __builtin_fmaf(p, f, 1.0f);
  return __builtin_ldexpf(p, (int)k);
}
VRT_D float pow8(float x) {
  float x2 = x * x;
  float x4 = x2 * x2;
  return x4 * x4;
}
VRT_D float pow32(float x) {
  float a = x * x;
  a = a * a;
  a = a * a;
  a = a * a;
  return a * a;
}

// Bilinear filter position of texel coordinate tc (= uv * size - 0.5):
// floor and fraction after rounding tc to 8 fractional bits (nearest, ties
// to even), the sub-texel precision of a hardware texture filter.  Every
// step is exact: scaling by 256 and 1/256, rounding to an integer, and the
// difference of two integers below 2^24.
VRT_D void sub_texel(float tc, float& fl, float& fr) {
  const float q = __builtin_rintf(tc * 256.0f);
  fl = vfloor(q * (1.0f / 256.0f));
  fr = (q - fl * 256.0f) * (1.0f / 256.0f);
}

// ---- storage formats ----------------------------------------------------------
VRT_D uint16_t f32_to_f16(float x) { return __builtin_bit_cast(uint16_t, (_Float16)x); }
VRT_D float f16_to_f32(uint16_t h) { return (float)__builtin_bit_cast(_Float16, h); }
VRT_D uint32_t f32_to_unorm8(float x) {
  return (uint32_t)vfloor(vclamp(x, 0.0f, 1.0f) * 255.0f + 0.5f);
}

// ---- PCG, common.wgsl:69-84 -------------------------------------------------
VRT_D uint32_t pcg_hash(uint32_t v_in) {
  uint32_t v = v_in * 747796405u + 2891336453u;
  uint32_t w = ((v >> ((v >> 28u) + 4u)) ^ v) * 277803737u;
  return (w >> 22u) ^ w;
}
VRT_D uint32_t init_rng(uint32_t px, uint32_t py, uint32_t frame) {
  return pcg_hash(px + pcg_hash(py + pcg_hash(frame)));
}
VRT_D float rand01(uint32_t& s) {
  s = pcg_hash(s);
  return (float)s * (1.0f / 4294967296.0f);
}

// ---- material, voxel.wgsl:41-51 ---------------------------------------------
struct Material {
  f3 albedo;
  float emissive;
};
VRT_D Material unpack_material(uint32_t m) {
  const float inv255 = 1.0f / 255.0f;
  Material r;
  r.albedo = mk3((float)(m & 0xffu), (float)((m >> 8) & 0xffu), (float)((m >> 16) & 0xffu)) * inv255;
  r.albedo = r.albedo * r.albedo;
  r.emissive = (float)((m >> 24) & 0xffu) * inv255;
  return r;
}

}  // namespace vrt
