// vrt_trace.h -- two-level Amanatides-Woo DDA on the device (voxel.wgsl:65-157).
//
// Scene layout in HBM (built by vrt_scene_upload):
//   bricks[bgx*bgy*bgz]  one u32 per 8^3 brick, x fastest: 0 = empty, else
//                        1 + slot of the brick in the pool.  Non-zero exactly
//                        where the reference's bricks[] is 1 (scene.js:432).
//   pool[slot]           528 u32 = 2112 B (33 x 64 B lines): words 0..15 are
//                        the brick's 512-bit occupancy mask, bit
//                        lx + 8*ly + 64*lz (scene.js:429-430); words 16..527
//                        the brick's packed voxels in the same order.
// The coarse loop reads one brick word per step, the fine loop one mask word
// per step, and a hit one voxel word -- the access pattern of voxel.wgsl, with
// the mask and voxels of a brick in one contiguous 2 KB block so a fine walk
// stays within a few cache lines.  Empty bricks cost no pool storage.
#pragma once
#include "vrt_device.h"

namespace vrt {

constexpr int POOL_WORDS = 16 + 512;

struct SceneDev {
  const uint32_t* bricks;
  const uint32_t* pool;
  const uint32_t* bits;    // occupancy bit per brick: bit (i & 31) of word i >> 5
  const uint32_t* prefix;  // prefix[w] = set bits in bits[0 .. w-1] (= pool slot of the word's first brick)
  uint32_t bit_words;      // ceil(bricks / 32)
  // The same bitmap over a power-of-two padded brick grid, indexed by the
  // packed position bx | by << psh1 | bz << psh2 (padding bits are 0), and
  // its prefix counts: the layout of the single-loop DDA (vrt_dda_fast.h).
  const uint32_t* pbits;
  const uint32_t* pprefix;
  // Per padded bitmap word, the Chebyshev distance (in bricks, over y and z)
  // from the brick rows the word covers to the nearest row holding an
  // occupied brick; 0 when one of its rows is occupied, at most 255.  Every
  // brick of a row within distance d - 1 of such a row is empty, whatever its
  // x.  pwords bytes, padded to whole words.
  const uint8_t* pdist;
  uint32_t pwords;
  uint32_t psh1, psh2;
  int nx, ny, nz;
  int bgx, bgy, bgz;
  float voxel_size, inv_voxel;
  // Single-loop DDA in the traversal kernels: lanes of a wave that reach an
  // occupied brick wait until this many of them (or every lane still
  // walking) wait, then enter their bricks together (vrt_dda_fast.h).
  // 1 = enter at once.
  uint32_t entry_batch;
  // Order in which the blocks of a one-pixel-per-lane path-trace launch take
  // its tiles (map_tile in vrt_kernels.hip): 0 row-major, 1 one contiguous
  // band of tiles per XCD (blocks are dealt to the 8 XCDs round-robin, so
  // each XCD's L2 sees the bricks of one band), 2 row-major from the last
  // tile, 3 bands from their last tile.
  uint32_t tile_order;
  // The box of occupied bricks, in brick coordinates (inclusive; lo > hi on
  // every axis for a scene with none): outside it every brick is empty, so
  // a ray that leaves it cannot hit.  The path tracer's DDA ends rays there
  // (DdaFastT<.., BOX = true>).
  int box_lo[3], box_hi[3];
  // Per pool slot, the box of the brick's solid voxels: lo x | lo y << 3 |
  // lo z << 6 | hi x << 9 | hi y << 12 | hi z << 15 (brick-local, inclusive).
  // A ray entering the brick that cannot reach it walks no voxel of it
  // (DdaFastT<.., FBOX = true>).
  const uint32_t* vbox;
  // Always 0.0f.  Passed at run time so that (zero - x) stays a real
  // subtraction: the gfx950 instruction selector folds a literal (0.0f - x)
  // feeding a multiply into a negated operand, which turns +0 into -0.
  float zero;
};

struct TraceResult {
  float t;          // world metres, -1 on miss
  f3 n;             // hit normal (-rd on miss)
  uint32_t mat;     // packed material, 0 on miss
  uint32_t voxel;   // dense linear index, 0xFFFFFFFF on miss
  uint32_t bsteps;  // brick-loop iterations
  uint32_t vsteps;  // voxel-loop iterations
  uint32_t face;    // VoxelBench face id, 6 = zero normal, 255 = miss
};

// voxel.wgsl:65-71: s / max(|v|, 1e-8) with s = +-1, correctly rounded.  On
// the device, for divisors up to 2^64, as v_rcp_f32 refined by one Newton
// step (e = 1 - b*r, r + e*r, each one FMA): that equals the IEEE quotient
// 1/b for every float b in [1e-8, 2^64] (checked exhaustively on gfx950 by
// tests/rcp_exhaustive.hip), and -1/b is its negation.  Larger divisors and
// NaN take the division.
VRT_D float safe_inv(float v) {
  const float b = vmax(vabs(v), 1e-8f);
#if defined(__HIP_DEVICE_COMPILE__)
  if (b <= 0x1p64f) {
    const float r = __builtin_amdgcn_rcpf(b);
    const float q = __builtin_fmaf(__builtin_fmaf(-b, r, 1.0f), r, r);
    return v >= 0.0f ? q : -q;
  }
#endif
  const float s = v >= 0.0f ? 1.0f : -1.0f;
  return s / b;
}
VRT_D int sign_i(float v) { return v > 0.0f ? 1 : (v < 0.0f ? -1 : 0); }
VRT_D int floor_clamp(float x, int lo, int hi) {
  return (int)vclamp(vfloor(x), (float)lo, (float)hi);
}

// Where the DDA reads occupancy from.  GlobalMem: the brick word per coarse
// step and the mask word per fine step straight from HBM/L2 (the access
// pattern of voxel.wgsl).  LdsMem: the coarse test reads the scene's
// occupancy bitmap staged in LDS (one bit per brick); on entering an occupied
// brick its 64-byte mask is copied once into this lane's LDS slot, so every
// fine step reads LDS.  Same cells tested in the same order, so hits, normals
// and step counts are identical.
struct GlobalMem {
  VRT_D bool brick(const SceneDev& s, uint32_t bi, uint32_t& bw) const {
    bw = s.bricks[bi];
    return bw != 0u;
  }
  VRT_D void enter(const uint32_t*) const {}
  VRT_D uint32_t mask_word(const uint32_t* blk, uint32_t w) const { return blk[w]; }
};

struct LdsMem {
  const uint32_t* bits;  // LDS copy of SceneDev::bits
  uint32_t* mask;        // this lane's mask slot: word w at mask[w * stride]
  uint32_t stride;       // lanes in the block (word-major layout: no bank conflicts)
  VRT_D bool brick(const SceneDev& s, uint32_t bi, uint32_t& bw) const {
    if (((bits[bi >> 5] >> (bi & 31u)) & 1u) == 0u) return false;
    bw = s.bricks[bi];
    return true;
  }
  VRT_D void enter(const uint32_t* blk) const {
    const uint4* q = (const uint4*)blk;  // pool entries are 2112 B: 16-byte aligned
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint4 v = q[k];
      mask[(4 * k + 0) * stride] = v.x;
      mask[(4 * k + 1) * stride] = v.y;
      mask[(4 * k + 2) * stride] = v.z;
      mask[(4 * k + 3) * stride] = v.w;
    }
  }
  VRT_D uint32_t mask_word(const uint32_t*, uint32_t w) const { return mask[w * stride]; }
};

// A ray with a NaN component or an origin that is not finite has no defined
// walk (its cell indices would come from NaN, and an index from a NaN can
// leave the grid); every DDA ends it at setup as a miss.
VRT_D bool ray_undefined(const float ro[3], const float rd[3]) {
  bool bad = false;
#pragma unroll
  for (int c = 0; c < 3; ++c) bad |= !__builtin_isfinite(ro[c]) || rd[c] != rd[c];
  return bad;
}

// voxel.wgsl:73-153.  Non-cubic grids: clip against [0,n]^3 per axis; the
// brick-loop bound 3*BGRID+2 becomes bgx+bgy+bgz+2 (identical for cubes).
template <class Mem>
VRT_D void trace_m(const SceneDev& s, const Mem& mem, f3 roW, f3 rd, float maxTWorld, TraceResult& o) {
  o.t = -1.0f;
  o.mat = 0u;
  o.n = mk3(-rd.x, -rd.y, -rd.z);
  o.voxel = 0xFFFFFFFFu;
  o.face = 255u;
  o.bsteps = 0u;
  o.vsteps = 0u;

  const float ro[3] = {roW.x * s.inv_voxel, roW.y * s.inv_voxel, roW.z * s.inv_voxel};
  const float maxT = maxTWorld * s.inv_voxel;
  const float rdv[3] = {rd.x, rd.y, rd.z};
  if (ray_undefined(ro, rdv)) return;
  float inv[3], ge[3], adl[3];
  int sd[3];
  float tmin[3], tmax[3];
  const int G[3] = {s.nx, s.ny, s.nz};
  const int BG[3] = {s.bgx, s.bgy, s.bgz};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    inv[c] = safe_inv(rdv[c]);
    adl[c] = vabs(inv[c]);
    sd[c] = sign_i(rdv[c]);
    ge[c] = rdv[c] >= 0.0f ? 1.0f : 0.0f;
    float t0 = (s.zero - ro[c]) * inv[c];
    float t1 = ((float)G[c] - ro[c]) * inv[c];
    tmin[c] = vmin(t0, t1);
    tmax[c] = vmax(t0, t1);
  }
  const float tEnter = vmax(vmax(tmin[0], tmin[1]), vmax(tmin[2], 0.0f));
  const float tExit = vmin(vmin(tmax[0], tmax[1]), vmin(tmax[2], maxT));
  if (tEnter > tExit) return;

  int ax = (tmin[0] >= tmin[1] && tmin[0] >= tmin[2]) ? 0 : (tmin[1] >= tmin[2] ? 1 : 2);
  float t = tEnter;
  int bpos[3];
  float tDeltaB[3], tMaxB[3];
  {
    const float ts = t + 1e-4f;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      float sp = ro[c] + rdv[c] * ts;
      bpos[c] = floor_clamp(sp / 8.0f, 0, BG[c] - 1);
      tDeltaB[c] = adl[c] * 8.0f;
      tMaxB[c] = (((float)bpos[c] + ge[c]) * 8.0f - ro[c]) * inv[c];
    }
  }
  const int maxIter = s.bgx + s.bgy + s.bgz + 2;
  const uint32_t bstride_y = (uint32_t)s.bgx, bstride_z = (uint32_t)s.bgx * (uint32_t)s.bgy;
  uint32_t bsteps = 0, vsteps = 0;
  for (int iter = 0; iter < maxIter; ++iter) {
    if (t > tExit) break;
    const uint32_t bi = (uint32_t)bpos[0] + (uint32_t)bpos[1] * bstride_y + (uint32_t)bpos[2] * bstride_z;
    ++bsteps;
    uint32_t bw;
    if (mem.brick(s, bi, bw)) {
      const uint32_t* blk = s.pool + (uint64_t)(bw - 1u) * POOL_WORDS;
      mem.enter(blk);
      int bmin[3] = {bpos[0] * 8, bpos[1] * 8, bpos[2] * 8};
      int vpos[3];
      float tMaxV[3];
      const float ts = t + 1e-4f;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        float p = ro[c] + rdv[c] * ts;
        vpos[c] = floor_clamp(p, bmin[c], bmin[c] + 7);
        tMaxV[c] = (((float)vpos[c] + ge[c]) - ro[c]) * inv[c];
      }
      float tCur = t;
      int vax = ax;
      for (int j = 0; j < 24; ++j) {
        ++vsteps;
        const uint32_t bit = (uint32_t)((vpos[0] - bmin[0]) + (vpos[1] - bmin[1]) * 8 + (vpos[2] - bmin[2]) * 64);
        const uint32_t word = mem.mask_word(blk, bit >> 5);
        if ((word >> (bit & 31u)) & 1u) {
          const uint32_t m = blk[16 + bit];
          o.bsteps = bsteps;
          o.vsteps = vsteps;
          if (tCur > maxT) return;
          o.t = tCur * s.voxel_size;
          float nn[3];
#pragma unroll
          for (int c = 0; c < 3; ++c) nn[c] = (c == vax ? -1.0f : -0.0f) * (float)sd[c];
          o.n = mk3(nn[0], nn[1], nn[2]);
          o.mat = m;
          o.voxel = (uint32_t)vpos[0] + (uint32_t)vpos[1] * (uint32_t)s.nx +
                    (uint32_t)vpos[2] * (uint32_t)s.nx * (uint32_t)s.ny;
          const float nv = vax == 0 ? nn[0] : (vax == 1 ? nn[1] : nn[2]);
          o.face = nv > 0.0f ? (uint32_t)(2 * vax) : (nv < 0.0f ? (uint32_t)(2 * vax + 1) : 6u);
          return;
        }
        if (tMaxV[0] <= tMaxV[1] && tMaxV[0] <= tMaxV[2]) {
          tCur = tMaxV[0]; tMaxV[0] += adl[0]; vpos[0] += sd[0]; vax = 0;
        } else if (tMaxV[1] <= tMaxV[2]) {
          tCur = tMaxV[1]; tMaxV[1] += adl[1]; vpos[1] += sd[1]; vax = 1;
        } else {
          tCur = tMaxV[2]; tMaxV[2] += adl[2]; vpos[2] += sd[2]; vax = 2;
        }
        if ((unsigned)(vpos[0] - bmin[0]) > 7u || (unsigned)(vpos[1] - bmin[1]) > 7u ||
            (unsigned)(vpos[2] - bmin[2]) > 7u)
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
    if ((unsigned)bpos[0] >= (unsigned)s.bgx || (unsigned)bpos[1] >= (unsigned)s.bgy ||
        (unsigned)bpos[2] >= (unsigned)s.bgz)
      break;
  }
  o.bsteps = bsteps;
  o.vsteps = vsteps;
}

VRT_D void trace(const SceneDev& s, f3 roW, f3 rd, float maxTWorld, TraceResult& o) {
  trace_m(s, GlobalMem{}, roW, rd, maxTWorld, o);
}

}  // namespace vrt
