// vrt_dda_fast.h -- the two-level DDA of voxel.wgsl:73-153 as one tight loop
// whose every iteration reads one LDS word, at either level.
//
// Occupancy lives in LDS (one array per block, see LdsScene): a 16-word slot
// per lane into which the 512-bit mask of the brick being walked is copied on
// entry, the scene's brick bitmap over the power-of-two padded brick grid
// (SceneDev::pbits, one bit per brick) and per-word prefix counts of set bits.
// The coarse cell index is the packed position bx | by << psh1 | bz << psh2,
// which is also its bit index, so the brick coordinates are bit fields.  A
// brick's pool slot is the number of occupied bricks before it (slots follow
// brick order, and padding bits are 0), i.e. prefix[w] + popcount of the
// lower bits of its bitmap word -- no dependent HBM read of the brick word.
//
// One iteration = test the current cell + advance the current level by one
// cell.  The level never appears in that code: it only sets registers that
// the iteration reads (LDS base and word stride, step-counter increment,
// t bound, tMax deltas, index strides, exit decrements), so lanes at
// different levels run the same instructions together.  Rare events (entering
// an occupied brick, leaving it, hit, miss) branch.
//
// Exactness: every t value comes from the same FP32 operations in the same
// order as vrt::trace_m (and the oracle): tMaxB += |invRd|*8, tMaxV += |invRd|,
// fine tMax recomputed from the origin on entry, ties x, then y, then z.  The
// reference's integer bounds tests become counts of cells left along each
// axis (an axis with zero direction never leaves, as its position never
// changes) and its iteration caps are kept.  Step counts are the reference's
// loop iterations.
#pragma once
#include "vrt_trace.h"

namespace vrt {

// LDS of a 256-lane block: words [0, 4096) hold the lanes' mask slots
// (word w of lane l at w * 256 + l), then the bitmap at bits_off, then its
// prefix counts at prefix_off.
struct LdsScene {
  uint32_t* lds;
  uint32_t lane;        // this lane's slot base (0..255)
  uint32_t bits_off;    // 4096
  uint32_t prefix_off;  // bits_off + SceneDev::pwords
};

// Cells left along each axis before leaving [0, n) from p moving by sd,
// packed 8 bits per axis, and the per-axis decrements; an axis with sd == 0
// gets 255 and a zero decrement, so it never reports an exit.
VRT_D uint32_t pack_left(const int p[3], const int n[3], const int sd[3]) {
  uint32_t cnt = 0u;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const uint32_t left = sd[c] > 0 ? (uint32_t)(n[c] - 1 - p[c]) : (sd[c] < 0 ? (uint32_t)p[c] : 255u);
    cnt |= left << (8 * c);
  }
  return cnt;
}

VRT_D void trace_fast(const SceneDev& s, const LdsScene& L, f3 roW, f3 rd, float maxTWorld, TraceResult& o) {
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
    const float t0 = (s.zero - ro[c]) * inv[c];
    const float t1 = ((float)G[c] - ro[c]) * inv[c];
    tmin[c] = vmin(t0, t1);
    tmax[c] = vmax(t0, t1);
  }
  const float tEnter = vmax(vmax(tmin[0], tmin[1]), vmax(tmin[2], 0.0f));
  const float tExit = vmin(vmin(tmax[0], tmax[1]), vmin(tmax[2], maxT));
  if (tEnter > tExit) return;

  uint32_t ax = (tmin[0] >= tmin[1] && tmin[0] >= tmin[2]) ? 0u : (tmin[1] >= tmin[2] ? 1u : 2u);
  float t = tEnter;
  const uint32_t sh1 = s.psh1, sh2 = s.psh2;
  const uint32_t mx = (1u << sh1) - 1u, my = (1u << (sh2 - sh1)) - 1u;
  const uint32_t dec[3] = {sd[0] != 0 ? 1u : 0u, sd[1] != 0 ? 256u : 0u, sd[2] != 0 ? 65536u : 0u};

  // current level (starts coarse)
  float tm[3], d[3];
  int di[3];
  uint32_t idx, cnt;
  {
    const float ts = t + 1e-4f;
    int bp[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float sp = ro[c] + rdv[c] * ts;
      bp[c] = floor_clamp(sp / 8.0f, 0, BG[c] - 1);
      d[c] = adl[c] * 8.0f;
      tm[c] = (((float)bp[c] + ge[c]) * 8.0f - ro[c]) * inv[c];
    }
    di[0] = sd[0];
    di[1] = sd[1] * (int)(1u << sh1);
    di[2] = sd[2] * (int)(1u << sh2);
    idx = (uint32_t)bp[0] | ((uint32_t)bp[1] << sh1) | ((uint32_t)bp[2] << sh2);
    cnt = pack_left(bp, BG, sd);
  }
  uint32_t iter = (uint32_t)(s.bgx + s.bgy + s.bgz + 2);
  uint32_t base = L.bits_off, wshift = 0u, inc = 1u;  // coarse: bitmap word idx>>5, brick steps
  float tBound = tExit;                                // coarse: leave at t > tExit; fine: never
  // coarse state saved while walking a brick
  float tmB[3] = {0.0f, 0.0f, 0.0f};
  uint32_t idxB = 0u, cntB = 0u, iterB = 0u, blk = 0u;
  uint32_t steps = 0u;  // brick steps | voxel steps << 16

  for (;;) {
    // ---- test the current cell ----
    if (t > tBound) break;
    const uint32_t word = L.lds[base + ((idx >> 5) << wshift)];
    steps += inc;
    if ((word >> (idx & 31u)) & 1u) {
      if (inc != 1u) {  // fine level: hit (voxel.wgsl:122-128)
        o.bsteps = steps & 0xffffu;
        o.vsteps = steps >> 16;
        if (t > maxT) return;
        o.t = t * s.voxel_size;
        float nn[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) nn[c] = ((uint32_t)c == ax ? -1.0f : -0.0f) * (float)sd[c];
        o.n = mk3(nn[0], nn[1], nn[2]);
        o.mat = s.pool[(uint64_t)blk + 16u + idx];
        const uint32_t bx = idxB & mx, by = (idxB >> sh1) & my, bz = idxB >> sh2;
        const uint32_t vx = bx * 8u + (idx & 7u), vy = by * 8u + ((idx >> 3) & 7u), vz = bz * 8u + (idx >> 6);
        o.voxel = vx + vy * (uint32_t)s.nx + vz * (uint32_t)s.nx * (uint32_t)s.ny;
        const float nv = ax == 0u ? nn[0] : (ax == 1u ? nn[1] : nn[2]);
        o.face = nv > 0.0f ? 2u * ax : (nv < 0.0f ? 2u * ax + 1u : 6u);
        return;
      }
      // coarse level: enter the occupied brick (voxel.wgsl:111-118)
      const uint32_t w = idx >> 5;
      const uint32_t slot = L.lds[L.prefix_off + w] + (uint32_t)__builtin_popcount(word & ((1u << (idx & 31u)) - 1u));
      blk = slot * (uint32_t)POOL_WORDS;
      const uint4* q = (const uint4*)(s.pool + blk);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const uint4 v = q[k];
        L.lds[((4 * k + 0) << 8) + L.lane] = v.x;
        L.lds[((4 * k + 1) << 8) + L.lane] = v.y;
        L.lds[((4 * k + 2) << 8) + L.lane] = v.z;
        L.lds[((4 * k + 3) << 8) + L.lane] = v.w;
      }
      const uint32_t bx = idx & mx, by = (idx >> sh1) & my, bz = idx >> sh2;
      const int bmin[3] = {(int)bx * 8, (int)by * 8, (int)bz * 8};
      const float ts = t + 1e-4f;
      int lp[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        tmB[c] = tm[c];
        const float p = ro[c] + rdv[c] * ts;
        const int vp = floor_clamp(p, bmin[c], bmin[c] + 7);
        tm[c] = (((float)vp + ge[c]) - ro[c]) * inv[c];
        lp[c] = vp - bmin[c];
        d[c] = adl[c];
      }
      di[0] = sd[0];
      di[1] = sd[1] * 8;
      di[2] = sd[2] * 64;
      idxB = idx;
      cntB = cnt;
      iterB = iter;
      idx = (uint32_t)lp[0] + 8u * (uint32_t)lp[1] + 64u * (uint32_t)lp[2];
      {
        const int n8[3] = {8, 8, 8};
        cnt = pack_left(lp, n8, sd);
      }
      iter = 24u;
      base = L.lane;
      wshift = 8u;
      inc = 0x10000u;
      tBound = __builtin_inff();
      continue;  // test the entry voxel
    }
    // ---- advance the current level by one cell (ties x, then y, then z) ----
    bool out;
    {
      const bool c0 = tm[0] <= tm[1] && tm[0] <= tm[2];
      const bool c1 = !c0 && tm[1] <= tm[2];
      t = c0 ? tm[0] : (c1 ? tm[1] : tm[2]);
      ax = c0 ? 0u : (c1 ? 1u : 2u);
      const float n0 = tm[0] + d[0], n1 = tm[1] + d[1], n2 = tm[2] + d[2];
      tm[0] = c0 ? n0 : tm[0];
      tm[1] = c1 ? n1 : tm[1];
      tm[2] = (!c0 && !c1) ? n2 : tm[2];
      idx += (uint32_t)(c0 ? di[0] : (c1 ? di[1] : di[2]));
      const uint32_t sh = c0 ? 0u : (c1 ? 8u : 16u);
      out = ((cnt >> sh) & 255u) == 0u || iter == 1u;
      cnt -= c0 ? dec[0] : (c1 ? dec[1] : dec[2]);
      iter -= 1u;
    }
    if (out) {
      if (inc == 1u) break;  // coarse: left the grid / out of brick iterations -> miss
      // fine: leave the brick, back to the coarse DDA and its advance
      // (voxel.wgsl:142-150 follows the fine loop)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        tm[c] = tmB[c];
        d[c] = adl[c] * 8.0f;
      }
      di[0] = sd[0];
      di[1] = sd[1] * (int)(1u << sh1);
      di[2] = sd[2] * (int)(1u << sh2);
      idx = idxB;
      cnt = cntB;
      iter = iterB;
      base = L.bits_off;
      wshift = 0u;
      inc = 1u;
      tBound = tExit;
      const bool c0 = tm[0] <= tm[1] && tm[0] <= tm[2];
      const bool c1 = !c0 && tm[1] <= tm[2];
      t = c0 ? tm[0] : (c1 ? tm[1] : tm[2]);
      ax = c0 ? 0u : (c1 ? 1u : 2u);
      const float n0 = tm[0] + d[0], n1 = tm[1] + d[1], n2 = tm[2] + d[2];
      tm[0] = c0 ? n0 : tm[0];
      tm[1] = c1 ? n1 : tm[1];
      tm[2] = (!c0 && !c1) ? n2 : tm[2];
      idx += (uint32_t)(c0 ? di[0] : (c1 ? di[1] : di[2]));
      const uint32_t sh = c0 ? 0u : (c1 ? 8u : 16u);
      const bool gone = ((cnt >> sh) & 255u) == 0u || iter == 1u;
      cnt -= c0 ? dec[0] : (c1 ? dec[1] : dec[2]);
      iter -= 1u;
      if (gone) break;
    }
  }
  o.bsteps = steps & 0xffffu;
  o.vsteps = steps >> 16;
}

}  // namespace vrt
