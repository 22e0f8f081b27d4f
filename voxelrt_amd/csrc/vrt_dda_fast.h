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
// Entering a brick (copy of its mask into LDS, fine-level setup) is the
// longest of those branches, and in a wave of 64 lanes some lane enters a
// brick in most iterations, so the whole wave would pay for it each time with
// a few lanes active.  step() therefore stops at an occupied brick
// (DDA_ENTER) and run() lets the lanes that stopped enter together: its inner
// loop steps the other lanes until SceneDev::entry_batch lanes of the wave
// wait (a wave-uniform count, so the loop exit is a scalar branch) or no lane
// is left stepping.  That changes when a lane enters, never what it computes.
//
// Empty rows (the Skip instantiations).  At an empty coarse cell whose bitmap
// word has row distance D >= 2, every cell the ray can reach with at most
// D - 1 steps along y and along z (and any number along x, inside the grid)
// is empty.  skip_rows() then performs the coarse steps that stay inside that
// region in bulk: with T below the t of the first step that could leave it
// (and below tExit), the DDA takes, in some interleaving, exactly the steps
// whose tMax is < T.  Their order does not change where they lead: each axis'
// tMax runs through the same additions tMax += tDelta as one step at a time,
// and the position, the last step's t and axis (largest t; on ties the
// later axis, as the per-step order x, y, z gives), the cell counts, the
// iteration budget and the step count follow from the per-axis step counts.
// No cell among them is occupied, so no test inside the run could have ended
// the ray, and the state after the run is the step-by-step state: results,
// including step counts, are unchanged.
//
// Exactness: every t value comes from the same FP32 operations in the same
// order as vrt::trace_m (and the oracle): tMaxB += |invRd|*8, tMaxV += |invRd|,
// fine tMax recomputed from the origin on entry, ties x, then y, then z.  The
// reference's integer bounds tests become counts of cells left along each
// axis (an axis with zero direction never leaves, as its position never
// changes) and its iteration caps are kept (they bind for degenerate
// directions, e.g. a zero vector, whose steps never leave the cell) -- except
// in the path tracer's form (NOITER below), which ends those walks by the
// counts with the same result.  Step counts are the reference's loop
// iterations (the path tracer's forms count fewer for walks they end early).
#pragma once
#include "vrt_trace.h"

namespace vrt {

// LDS of a 256-lane block: words [0, 4096) hold the lanes' mask slots
// (word w of lane l at w * 256 + l), then the bitmap at bits_off, then its
// prefix counts at prefix_off, then the row distances (SceneDev::pdist, one
// byte per bitmap word) at dist_off.
struct LdsScene {
  uint32_t* lds;
  uint32_t lane;        // this lane's slot base (0..255)
  uint32_t bits_off;    // 4096
  uint32_t prefix_off;  // bits_off + SceneDev::pwords
  uint32_t dist_off;    // prefix_off + SceneDev::pwords
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

// Resumable state of one ray (all in registers once inlined).  begin() sets
// the ray up; step() runs one iteration and returns DDA_RUN, DDA_MISS or
// DDA_HIT; finish() writes the result record.  trace_fast() runs them in a
// loop; the persistent traversal kernel interleaves rays per lane.
enum : uint32_t { DDA_RUN = 0u, DDA_MISS = 1u, DDA_HIT = 2u, DDA_ENTER = 3u, DDA_LEAVE = 4u };

// Bitwise selects (c ? a : b, bit for bit).  Written as masks so that the
// compiler cannot turn a choice between two state-array elements into a load
// through a chosen address, which would keep the DDA state in scratch memory.
VRT_D float vsel(bool c, float a, float b) {
  const uint32_t m = 0u - (uint32_t)c;
  return __builtin_bit_cast(float, (m & __builtin_bit_cast(uint32_t, a)) | (~m & __builtin_bit_cast(uint32_t, b)));
}
VRT_D int isel(bool c, int a, int b) {
  const uint32_t m = 0u - (uint32_t)c;
  return (int)((m & (uint32_t)a) | (~m & (uint32_t)b));
}

// floor(x) clamped to [lo, lo + 7] (lo an integer-valued float), as a float:
// (float)floor_clamp(x, lo, lo + 7) for every finite x.  One median-of-three
// on the device instead of two compare-and-selects.
VRT_D float floor_clamp7f(float x, float lo) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_fmed3f(vfloor(x), lo, lo + 7.0f);
#else
  return vclamp(vfloor(x), lo, lo + 7.0f);
#endif
}

// Bit (i & 31) of w.
VRT_D bool bit_at(uint32_t w, uint32_t i) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_ubfe(w, i, 1u) != 0u;
#else
  return ((w >> (i & 31u)) & 1u) != 0u;
#endif
}

// GM = false: occupancy in LDS (LdsScene), the entered brick's mask copied into
// the lane's LDS slot.  GM = true: the same loop on global memory -- the padded
// bitmap, its prefix counts and the row distances read from the scene's
// buffers (L1/L2-resident), the entered brick's mask read in place from its
// pool entry (no copy) -- for scenes whose bitmaps do not fit in LDS.  Both
// levels then read words of one array at a base pointer (bitmap, or the pool
// entry) with a 4-byte stride, so the level still only switches registers.
// KM: the brick-entry cell counts through two per-ray constants (kfine,
// mfine) instead of per-axis selects: fewer instructions on the entry path,
// two more live registers -- the step-level kernels' choice (path tracer and
// traversal, whose occupancy LDS, not registers, bounds); the one-ray-per-lane
// kernels keep pack_left.
// BOX: the coarse level's cells-left counts run to the box of occupied
// bricks (SceneDev::box_lo / box_hi) instead of the grid: every brick outside
// it is empty, and along each axis the walk is monotone, so a ray that leaves
// the box (or starts outside it moving away) cannot hit; it ends there as a
// miss.  Hits, their t and both step counts are those of the grid walk; a
// miss records fewer brick steps (the path tracer's choice, whose step counts
// are statistics; the traversal's miss records keep the grid walk's counts).
// FBOX: the same at the fine level, against the entered brick's box of solid
// voxels (SceneDev::vbox): the fine walk ends where the ray leaves that box,
// and enter() returns false for a ray that cannot reach it (the caller then
// leaves the brick at once: resume()).  The coarse walk resumes from its own
// saved state either way, so every hit (voxel, face, t) and the brick steps
// are unchanged; the voxel steps spent in bricks passed without a hit are
// fewer.
// NOITER (with BOX and FBOX: the path tracer's form): no iteration budget.
// Every step spends one cell of the count of the axis it selects, and an
// axis the ray does not move along starts with 2 instead of 255.  Such an
// axis has tDelta 1e8 (8e8 at the coarse level: the clamped reciprocal of a
// zero component), so after its first selection it can only be selected
// again at t >= 1e8 voxel units; the path tracer's walks end at maxT = 1e4 m
// = 1.6e5 voxels (the coarse level's t test; a hit beyond maxT is reported
// as a miss), so every cell that NOITER leaves untested lies where no hit
// counts.  Hits, misses and their t are the budgeted walk's for every ray --
// a zero direction, whose walk only the budget would end (it tests the same
// cell again and again), included -- and only the step counts of walks cut
// short are fewer.  (Rays with a NaN component end at setup in every form:
// ray_undefined.)  Requires maxT * inv_voxel < 1e8.
template <bool GM, bool KM = false, bool BOX = false, bool FBOX = false, bool NOITER = false>
struct DdaFastT {
  static_assert(!NOITER || (BOX && FBOX), "NOITER sets the counts of BOX and FBOX");
  static constexpr uint32_t STILL = NOITER ? 2u : 255u;  // count of an axis the ray does not move along
  float ro[3], rdv[3], inv[3];
  int sd[3];
  float tExit, maxT;
  // current level
  float tm[3], d[3];
  int di[3];
  uint32_t idx, cnt, iter;
  uint32_t base4, wsh, inc;  // LDS byte base, word stride shift (2: words, 10: slot rows), step increment
  const char* gbase;         // GM: the current level's word array
  float tBound, t;             // coarse: leave at t > tExit; fine: never
  uint32_t axs;  // axis of the last step x 8 (its field offset in cnt)
  // KM: pack_left over a brick's 8 cells as kfine ^ (cells & mfine): per
  // axis 7 - p = 7 ^ p (sd > 0), p = 0 ^ p (sd < 0), 255 = 255 ^ 0 (sd == 0)
  uint32_t kfine, mfine;
  // coarse state saved while walking a brick
  float tmB[3];
  uint32_t idxB, cntB, iterB, blk;
  uint32_t steps;  // brick steps | voxel steps << 16

  // voxel.wgsl:79-104.  Returns true when the ray misses the grid at setup
  // (o is then the miss record).
  VRT_D bool begin(const SceneDev& s, const LdsScene& L, f3 roW, f3 rd, float maxTWorld, TraceResult& o) {
    o.t = -1.0f;
    o.mat = 0u;
    o.n = mk3(-rd.x, -rd.y, -rd.z);
    o.voxel = 0xFFFFFFFFu;
    o.face = 255u;
    o.bsteps = 0u;
    o.vsteps = 0u;
    ro[0] = roW.x * s.inv_voxel;
    ro[1] = roW.y * s.inv_voxel;
    ro[2] = roW.z * s.inv_voxel;
    maxT = maxTWorld * s.inv_voxel;
    rdv[0] = rd.x;
    rdv[1] = rd.y;
    rdv[2] = rd.z;
    if (ray_undefined(ro, rdv)) return true;
    float tmin[3], tmax[3];
    const int G[3] = {s.nx, s.ny, s.nz};
    const int BG[3] = {s.bgx, s.bgy, s.bgz};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      inv[c] = safe_inv(rdv[c]);
      sd[c] = sign_i(rdv[c]);
      const float t0 = (s.zero - ro[c]) * inv[c];
      const float t1 = ((float)G[c] - ro[c]) * inv[c];
      tmin[c] = vmin(t0, t1);
      tmax[c] = vmax(t0, t1);
    }
    const float tEnter = vmax(vmax(tmin[0], tmin[1]), vmax(tmin[2], 0.0f));
    tExit = vmin(vmin(tmax[0], tmax[1]), vmin(tmax[2], maxT));
    if (tEnter > tExit) return true;
    axs = (tmin[0] >= tmin[1] && tmin[0] >= tmin[2]) ? 0u : (tmin[1] >= tmin[2] ? 8u : 16u);
    t = tEnter;
    const float ts = t + 1e-4f;
    int bp[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float sp = ro[c] + rdv[c] * ts;
      bp[c] = floor_clamp(sp / 8.0f, 0, BG[c] - 1);
      d[c] = vabs(inv[c]) * 8.0f;
      const float ge = rdv[c] >= 0.0f ? 1.0f : 0.0f;
      tm[c] = (((float)bp[c] + ge) * 8.0f - ro[c]) * inv[c];
    }
    di[0] = sd[0];
    di[1] = sd[1] * (int)(1u << s.psh1);
    di[2] = sd[2] * (int)(1u << s.psh2);
    idx = (uint32_t)bp[0] | ((uint32_t)bp[1] << s.psh1) | ((uint32_t)bp[2] << s.psh2);
    if constexpr (BOX) {
      // cells left to the box's far side along each moving axis; negative
      // (outside and moving away, or outside on an axis it does not move
      // along) means the ray never reaches an occupied brick
      cnt = 0u;
      bool out = false;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const int lo = s.box_lo[c], hi = s.box_hi[c];
        const int left = sd[c] > 0 ? hi - bp[c] : (sd[c] < 0 ? bp[c] - lo : (bp[c] < lo || bp[c] > hi ? -1 : (int)STILL));
        out |= left < 0;
        cnt |= (uint32_t)(left & 255) << (8 * c);
      }
      if (out) return true;
    } else {
      cnt = pack_left(bp, BG, sd);
    }
    kfine = mfine = 0u;
    if constexpr (KM) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        kfine |= (sd[c] > 0 ? 7u : (sd[c] < 0 ? 0u : 255u)) << (8 * c);
        mfine |= (sd[c] != 0 ? 7u : 0u) << (8 * c);
      }
    }
    iter = (uint32_t)(s.bgx + s.bgy + s.bgz + 2);
    if constexpr (GM) {
      gbase = (const char*)s.pbits;
      base4 = 0u;
    } else {
      gbase = nullptr;
      base4 = 4u * L.bits_off;
    }
    wsh = 2u;
    inc = 1u;
    tBound = tExit;
    tmB[0] = tmB[1] = tmB[2] = 0.0f;
    idxB = cntB = iterB = blk = 0u;
    steps = 0u;
    return false;
  }

  // Advance the current level by one cell (ties x, then y, then z); true
  // when that leaves the level's box or exhausts its iteration budget.
  VRT_D bool advance() {
    const float m0 = tm[0], m1 = tm[1], m2 = tm[2];
    const int i0 = di[0], i1 = di[1], i2 = di[2];
    const bool c0 = m0 <= m1 && m0 <= m2;
    const bool c1 = !c0 && m1 <= m2;
    t = vsel(c0, m0, vsel(c1, m1, m2));
    const float n0 = m0 + d[0], n1 = m1 + d[1], n2 = m2 + d[2];
    tm[0] = vsel(c0, n0, m0);
    tm[1] = vsel(c1, n1, m1);
    tm[2] = vsel(!c0 && !c1, n2, m2);
    const int dsel = isel(c0, i0, isel(c1, i1, i2));  // 0 exactly when that axis does not move
    idx += (uint32_t)dsel;
    const uint32_t sh = c0 ? 0u : (c1 ? 8u : 16u);
    axs = sh;
    bool out = ((cnt >> sh) & 255u) == 0u;
    if constexpr (NOITER) {
      cnt -= 1u << sh;
    } else {
      out |= iter == 1u;
      cnt -= (dsel != 0 ? 1u : 0u) << sh;
      iter -= 1u;
    }
    return out;
  }

  // Enter the occupied brick at idx (voxel.wgsl:111-118): its pool slot is
  // prefix[w] + the set bits below it in bitmap word w; its mask goes into
  // the lane's LDS slot; the fine level starts at the voxel containing
  // ro + rd * (t + 1e-4), clamped into the brick, with tMax from the origin.
  VRT_D bool enter(const SceneDev& s, const LdsScene& L) {
    const uint32_t w = idx >> 5;
    uint32_t slot;
    if constexpr (GM) {
      const uint32_t word = s.pbits[w];
      slot = s.pprefix[w] + (uint32_t)__builtin_popcount(word & ((1u << (idx & 31u)) - 1u));
      blk = slot * (uint32_t)POOL_WORDS;
      gbase = (const char*)(s.pool + blk);  // mask words 0..15 of the entry
    } else {
      const uint32_t word = L.lds[L.bits_off + w];
      slot = L.lds[L.prefix_off + w] + (uint32_t)__builtin_popcount(word & ((1u << (idx & 31u)) - 1u));
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
    }
    const uint32_t mx = (1u << s.psh1) - 1u, my = (1u << (s.psh2 - s.psh1)) - 1u;
    const uint32_t bx = idx & mx, by = (idx >> s.psh1) & my, bz = idx >> s.psh2;
    const int bmin[3] = {(int)bx * 8, (int)by * 8, (int)bz * 8};
    const float ts = t + 1e-4f;
    int lp[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      tmB[c] = tm[c];
      const float p = ro[c] + rdv[c] * ts;  // finite: ro, rd finite, ts <= tExit
      const float vp = floor_clamp7f(p, (float)bmin[c]);  // integer-valued
      const float ge = rdv[c] >= 0.0f ? 1.0f : 0.0f;
      tm[c] = ((vp + ge) - ro[c]) * inv[c];
      lp[c] = (int)vp - bmin[c];
      d[c] = vabs(inv[c]);
    }
    di[0] = sd[0];
    di[1] = sd[1] * 8;
    di[2] = sd[2] * 64;
    idxB = idx;
    cntB = cnt;
    if constexpr (!NOITER) iterB = iter;
    idx = (uint32_t)lp[0] + 8u * (uint32_t)lp[1] + 64u * (uint32_t)lp[2];
    bool reach = true;
    if constexpr (FBOX) {
      // the three axes at once, a byte each (values 0..7, so 8 + a - b never
      // borrows across bytes): A = 8 + hi - lp, B = 8 + lp - lo; an axis
      // moving up needs A >= 8 (left = A - 8), one moving down B >= 8, one
      // that does not move lo <= lp <= hi (both >= 8; left 255).  The axis
      // signs come from the ray's kfine / mfine (7 / 0 / 255, 7 / 0).
      const uint32_t vb = s.vbox[slot];
      const uint32_t lpk = (uint32_t)lp[0] | ((uint32_t)lp[1] << 8) | ((uint32_t)lp[2] << 16);
      const uint32_t lok = (vb & 7u) | ((vb << 5) & 0x700u) | ((vb << 10) & 0x70000u);
      const uint32_t hik = ((vb >> 9) & 7u) | ((vb >> 4) & 0x700u) | ((vb << 1) & 0x70000u);
      const uint32_t A = hik + 0x080808u - lpk, B = lpk + 0x080808u - lok;
      const uint32_t up = kfine & mfine, down = mfine & ~kfine;      // 7 per byte
      const uint32_t mu = up | (up << 1), md = down | (down << 1);   // 15 per byte
      const uint32_t S = (A & mu) | (B & md);
      const uint32_t mov = (mu | md) & 0x080808u, still = ~(mu | md) & 0x080808u;
      reach = (S & mov) == mov && (A & B & still) == still;
      cnt = (S - mov) | ((still >> 3) * STILL);
    } else if constexpr (KM) {
      cnt = kfine ^ (((uint32_t)lp[0] | ((uint32_t)lp[1] << 8) | ((uint32_t)lp[2] << 16)) & mfine);
    } else {
      const int n8[3] = {8, 8, 8};
      cnt = pack_left(lp, n8, sd);
    }
    if constexpr (!NOITER) iter = 24u;
    if constexpr (!GM) {
      base4 = 4u * L.lane;
      wsh = 10u;
    }
    inc = 0x10000u;
    tBound = __builtin_inff();
    return reach;
  }

  // Leave the brick: back to the coarse registers saved on entry.
  VRT_D void leave(const SceneDev& s, const LdsScene& L) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      tm[c] = tmB[c];
      d[c] = vabs(inv[c]) * 8.0f;
    }
    di[0] = sd[0];
    di[1] = sd[1] * (int)(1u << s.psh1);
    di[2] = sd[2] * (int)(1u << s.psh2);
    idx = idxB;
    cnt = cntB;
    if constexpr (!NOITER) iter = iterB;
    if constexpr (GM) {
      gbase = (const char*)s.pbits;
    } else {
      base4 = 4u * L.bits_off;
      wsh = 2u;
    }
    inc = 1u;
    tBound = tExit;
  }

  // Coarse level, at an empty cell with row distance D >= 1: take in bulk the
  // steps whose tMax is below T (see the file comment).  Returns false when
  // no step qualifies (the caller advances one step as usual).
  VRT_D bool skip_rows(uint32_t D) {
    const uint32_t c0 = cnt & 255u, c1 = (cnt >> 8) & 255u, c2 = (cnt >> 16) & 255u;
    const uint32_t r1 = D - 1u < c1 ? D - 1u : c1, r2 = D - 1u < c2 ? D - 1u : c2;
    if (iter <= c0 + r1 + r2) return false;  // keep the iteration budget out of the run
    // T: below the t of step r+1 along each moving axis (half a tDelta of
    // margin over the rounding of r additions), and not beyond tExit
    float T = tBound;
    if (sd[0] != 0) T = vmin(T, tm[0] + ((float)c0 - 0.5f) * d[0]);
    if (sd[1] != 0) T = vmin(T, tm[1] + ((float)r1 - 0.5f) * d[1]);
    if (sd[2] != 0) T = vmin(T, tm[2] + ((float)r2 - 0.5f) * d[2]);
    uint32_t k[3];
    float last[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      uint32_t n = 0u;
      float m = tm[c], p = 0.0f;
      while (m < T) {
        p = m;
        m = m + d[c];
        ++n;
      }
      tm[c] = m;
      k[c] = n;
      last[c] = p;
    }
    const uint32_t K = k[0] + k[1] + k[2];
    if (K == 0u) return false;
    // the last step: largest t, ties to the later axis
    float tl = -1.0f;
    uint32_t al = 0u;
#pragma unroll
    for (int c = 0; c < 3; ++c)
      if (k[c] != 0u && last[c] >= tl) {
        tl = last[c];
        al = (uint32_t)c;
      }
    t = tl;
    axs = 8u * al;
    iter -= K;
    idx += (uint32_t)((int)k[0] * di[0] + (int)k[1] * di[1] + (int)k[2] * di[2]);
    cnt -= (di[0] != 0 ? k[0] : 0u) | ((di[1] != 0 ? k[1] : 0u) << 8) | ((di[2] != 0 ? k[2] : 0u) << 16);
    steps += K - 1u;  // the cells passed over are tested (empty); the last one is tested next
    return true;
  }

  // One iteration: test the current cell, then advance.  Returns DDA_ENTER
  // (without advancing) at an occupied brick of the coarse level; the caller
  // calls enter() and steps on.  Skip: take runs of empty rows in bulk.
  template <bool Skip = false>
  VRT_D uint32_t step(const SceneDev& s, const LdsScene& L) {
    if (t > tBound) return DDA_MISS;
    uint32_t word;
    if constexpr (GM)
      word = *(const uint32_t*)(gbase + ((idx >> 5) << 2));
    else
      word = *(const uint32_t*)((const char*)L.lds + (((idx >> 5) << wsh) + base4));
    steps += inc;
    if (bit_at(word, idx)) return inc != 1u ? DDA_HIT : DDA_ENTER;  // voxel.wgsl:122 / 111
    if constexpr (Skip) {
      if (inc == 1u) {
        const uint32_t D = GM ? (uint32_t)s.pdist[idx >> 5]
                             : (L.lds[L.dist_off + (idx >> 7)] >> (8u * ((idx >> 5) & 3u))) & 255u;
        if (D >= 2u && skip_rows(D)) return DDA_RUN;  // runs that can also move in y and z
      }
    }
    if (!advance()) return DDA_RUN;
    if (inc == 1u) return DDA_MISS;  // coarse: left the grid / out of brick iterations
    // fine: leave the brick, back to the coarse DDA and its advance
    // (voxel.wgsl:142-150 follows the fine loop)
    leave(s, L);
    return advance() ? DDA_MISS : DDA_RUN;
  }

  // step() with both level switches left to the caller: returns DDA_ENTER at
  // an occupied brick of the coarse level (call enter()), DDA_LEAVE when the
  // fine level has left its brick (call resume(), which returns the coarse
  // level to the caller and takes its advance).  A caller that runs enter()
  // and resume() for many lanes at once keeps those long branches out of the
  // per-step code.  The same operations as step(), in the same order.
  VRT_D uint32_t step_ev(const SceneDev& s, const LdsScene& L) {
    if (t > tBound) return DDA_MISS;
    uint32_t word;
    if constexpr (GM)
      word = *(const uint32_t*)(gbase + ((idx >> 5) << 2));
    else
      word = *(const uint32_t*)((const char*)L.lds + (((idx >> 5) << wsh) + base4));
    steps += inc;
    if (bit_at(word, idx)) return inc != 1u ? DDA_HIT : DDA_ENTER;
    if (!advance()) return DDA_RUN;
    return inc == 1u ? DDA_MISS : DDA_LEAVE;
  }
  // After DDA_LEAVE: back to the coarse level and its advance (voxel.wgsl:142-150).
  VRT_D uint32_t resume(const SceneDev& s, const LdsScene& L) {
    leave(s, L);
    return advance() ? DDA_MISS : DDA_RUN;
  }
  // The first half of the coarse level's next step_ev, for a caller that can
  // enter at once: DDA_MISS when t is past the bound, DDA_ENTER (the step
  // counted) at an occupied brick, else DDA_RUN with nothing done -- the
  // next step_ev then takes the whole step.  Either way the same operations
  // as step_ev.
  VRT_D uint32_t peek_enter(const SceneDev& s, const LdsScene& L) {
    if (t > tBound) return DDA_MISS;
    uint32_t word;
    if constexpr (GM)
      word = *(const uint32_t*)(gbase + ((idx >> 5) << 2));
    else
      word = *(const uint32_t*)((const char*)L.lds + (((idx >> 5) << wsh) + base4));
    if (!bit_at(word, idx)) return DDA_RUN;
    steps += inc;
    return DDA_ENTER;
  }

  // Walk to the end (DDA_HIT or DDA_MISS).  Batched: enter bricks in groups
  // of up to s.entry_batch lanes, and every lane of the wave that walks a ray
  // calls it; otherwise (and with entry_batch <= 1) enter at once.
  template <bool Batched, bool Skip>
  VRT_D uint32_t run(const SceneDev& s, const LdsScene& L) {
    uint32_t r = DDA_RUN;
    if (!Batched || s.entry_batch <= 1u) {  // enter at once: no wave-level bookkeeping
      do {
        r = step<Skip>(s, L);
        if (r == DDA_ENTER) {
          enter(s, L);
          r = DDA_RUN;
        }
      } while (r == DDA_RUN);
      return r;
    }
    for (;;) {
      uint32_t waiting = 0u;  // lanes of the wave stopped at a brick (uniform)
      while (r == DDA_RUN) {
        r = step<Skip>(s, L);
        waiting += (uint32_t)__builtin_popcountll(wave_ballot(r == DDA_ENTER));
        if (waiting >= s.entry_batch) break;
      }
      if (r == DDA_ENTER) {
        enter(s, L);
        r = DDA_RUN;
      }
      if (r != DDA_RUN) return r;
    }
  }

  // The result record of a finished ray (status DDA_MISS or DDA_HIT).
  VRT_D void finish(const SceneDev& s, uint32_t status, TraceResult& o) const {
    o.t = -1.0f;
    o.mat = 0u;
    o.n = mk3(-rdv[0], -rdv[1], -rdv[2]);
    o.voxel = 0xFFFFFFFFu;
    o.face = 255u;
    o.bsteps = steps & 0xffffu;
    o.vsteps = steps >> 16;
    if (status != DDA_HIT || t > maxT) return;  // a hit beyond maxT reports a miss (voxel.wgsl:124)
    const uint32_t ax = axs >> 3;
    o.t = t * s.voxel_size;  // voxel.wgsl:122-128
    float nn[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) nn[c] = ((uint32_t)c == ax ? -1.0f : -0.0f) * (float)sd[c];
    o.n = mk3(nn[0], nn[1], nn[2]);
    o.mat = s.pool[(uint64_t)blk + 16u + idx];
    const uint32_t mx = (1u << s.psh1) - 1u, my = (1u << (s.psh2 - s.psh1)) - 1u;
    const uint32_t bx = idxB & mx, by = (idxB >> s.psh1) & my, bz = idxB >> s.psh2;
    const uint32_t vx = bx * 8u + (idx & 7u), vy = by * 8u + ((idx >> 3) & 7u), vz = bz * 8u + (idx >> 6);
    o.voxel = vx + vy * (uint32_t)s.nx + vz * (uint32_t)s.nx * (uint32_t)s.ny;
    const float nv = ax == 0u ? nn[0] : (ax == 1u ? nn[1] : nn[2]);
    o.face = nv > 0.0f ? 2u * ax : (nv < 0.0f ? 2u * ax + 1u : 6u);
  }
};

using DdaFast = DdaFastT<false>;

template <bool Batched = false, bool Skip = false, bool GM = false, bool KM = false>
VRT_D void trace_fast(const SceneDev& s, const LdsScene& L, f3 roW, f3 rd, float maxTWorld, TraceResult& o) {
  DdaFastT<GM, KM> st;
  if (st.begin(s, L, roW, rd, maxTWorld, o)) return;
  st.finish(s, st.template run<Batched, Skip>(s, L), o);
}

}  // namespace vrt
