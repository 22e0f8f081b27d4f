// vrt_dda_state.h -- the two-level DDA of voxel.wgsl:73-153 as a resumable
// per-lane state machine.  One call of step() performs exactly one loop
// iteration of the reference: a brick-loop iteration (one brick word) or a
// voxel-loop iteration (one mask word), with the same operations in the same
// order as vrt::trace() -- results (voxel, face, t, step counts) are
// identical.  Persistent kernels use it to interleave rays: a lane whose ray
// finishes takes the next ray from a queue while the rest of its wave keeps
// stepping, so a wave is not held by its longest ray.
#pragma once
#include "vrt_trace.h"

namespace vrt {

struct DdaState {
  // ray constants (voxel units)
  float ro[3], rd[3], inv[3], adl[3], ge[3];
  int sd[3];
  float tExit, maxT;
  // coarse (brick) level
  int bpos[3];
  float tMaxB[3];
  float t;
  int ax, iter, maxIter;
  // fine (voxel) level, local coordinates 0..7 inside the current brick
  bool in_brick;
  uint32_t blk;  // word offset of the brick's pool entry
  int lpos[3];
  float tMaxV[3];
  float tCur;
  int vax, j;
  uint32_t bsteps, vsteps;

  // Setup (voxel.wgsl:79-104).  Returns true when the ray already missed
  // (no overlap with the grid): the result is then in `o`.
  VRT_D bool begin(const SceneDev& s, f3 roW, f3 rdIn, float maxTWorld, TraceResult& o) {
    o.t = -1.0f;
    o.mat = 0u;
    o.n = mk3(-rdIn.x, -rdIn.y, -rdIn.z);
    o.voxel = 0xFFFFFFFFu;
    o.face = 255u;
    o.bsteps = 0u;
    o.vsteps = 0u;
    bsteps = vsteps = 0u;
    ro[0] = roW.x * s.inv_voxel;
    ro[1] = roW.y * s.inv_voxel;
    ro[2] = roW.z * s.inv_voxel;
    rd[0] = rdIn.x;
    rd[1] = rdIn.y;
    rd[2] = rdIn.z;
    maxT = maxTWorld * s.inv_voxel;
    const int G[3] = {s.nx, s.ny, s.nz};
    const int BG[3] = {s.bgx, s.bgy, s.bgz};
    float tmin[3], tmax[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      inv[c] = safe_inv(rd[c]);
      adl[c] = vabs(inv[c]);
      sd[c] = sign_i(rd[c]);
      ge[c] = rd[c] >= 0.0f ? 1.0f : 0.0f;
      const float t0 = (s.zero - ro[c]) * inv[c];
      const float t1 = ((float)G[c] - ro[c]) * inv[c];
      tmin[c] = vmin(t0, t1);
      tmax[c] = vmax(t0, t1);
    }
    const float tEnter = vmax(vmax(tmin[0], tmin[1]), vmax(tmin[2], 0.0f));
    tExit = vmin(vmin(tmax[0], tmax[1]), vmin(tmax[2], maxT));
    if (tEnter > tExit) return true;
    ax = (tmin[0] >= tmin[1] && tmin[0] >= tmin[2]) ? 0 : (tmin[1] >= tmin[2] ? 1 : 2);
    t = tEnter;
    const float ts = t + 1e-4f;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float sp = ro[c] + rd[c] * ts;
      bpos[c] = floor_clamp(sp / 8.0f, 0, BG[c] - 1);
      tMaxB[c] = (((float)bpos[c] + ge[c]) * 8.0f - ro[c]) * inv[c];
    }
    iter = 0;
    maxIter = s.bgx + s.bgy + s.bgz + 2;
    in_brick = false;
    return false;
  }

  // Brick step (voxel.wgsl:142-150); returns true when the ray left the grid
  // or used up the iteration bound (a miss).
  // One Amanatides-Woo advance on (tm, pos), ties broken x, then y, then z
  // (voxel.wgsl:131-137 / 143-149), written with selects: the chosen axis
  // gets tm += d and pos += sd, *tcur the old tm, *axis its index.
  VRT_D void advance(float tm[3], int pos[3], const float d[3], float* tcur, int* axis) const {
    const bool cx = tm[0] <= tm[1] && tm[0] <= tm[2];
    const bool cy = !cx && tm[1] <= tm[2];
    const bool cz = !cx && !cy;
    *tcur = cx ? tm[0] : (cy ? tm[1] : tm[2]);
    *axis = cx ? 0 : (cy ? 1 : 2);
    const float n0 = tm[0] + d[0], n1 = tm[1] + d[1], n2 = tm[2] + d[2];
    tm[0] = cx ? n0 : tm[0];
    tm[1] = cy ? n1 : tm[1];
    tm[2] = cz ? n2 : tm[2];
    pos[0] += cx ? sd[0] : 0;
    pos[1] += cy ? sd[1] : 0;
    pos[2] += cz ? sd[2] : 0;
  }

  VRT_D bool next_brick(const SceneDev& s) {
    const float dB[3] = {adl[0] * 8.0f, adl[1] * 8.0f, adl[2] * 8.0f};
    advance(tMaxB, bpos, dB, &t, &ax);
    ++iter;
    return (unsigned)bpos[0] >= (unsigned)s.bgx || (unsigned)bpos[1] >= (unsigned)s.bgy ||
           (unsigned)bpos[2] >= (unsigned)s.bgz || iter >= maxIter;
  }

  // One iteration.  Returns true when the ray finished; `o` then holds the
  // result (hit or miss) including the step counts.
  VRT_D bool step(const SceneDev& s, TraceResult& o) {
    if (!in_brick) {
      if (t > tExit) return finish_miss(o);
      const uint32_t bi = (uint32_t)bpos[0] + (uint32_t)bpos[1] * (uint32_t)s.bgx +
                          (uint32_t)bpos[2] * (uint32_t)s.bgx * (uint32_t)s.bgy;
      ++bsteps;
      const uint32_t bw = s.bricks[bi];
      if (bw != 0u) {
        blk = (bw - 1u) * (uint32_t)POOL_WORDS;
        const float ts = t + 1e-4f;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const int bmin = bpos[c] * 8;
          const float p = ro[c] + rd[c] * ts;
          const int vp = floor_clamp(p, bmin, bmin + 7);
          lpos[c] = vp - bmin;
          tMaxV[c] = (((float)vp + ge[c]) - ro[c]) * inv[c];
        }
        tCur = t;
        vax = ax;
        j = 0;
        in_brick = true;
        return false;
      }
      if (next_brick(s)) return finish_miss(o);
      return false;
    }
    ++vsteps;
    const uint32_t bit = (uint32_t)(lpos[0] + lpos[1] * 8 + lpos[2] * 64);
    const uint32_t* blkp = s.pool + blk;
    const uint32_t word = blkp[bit >> 5];
    if ((word >> (bit & 31u)) & 1u) {
      const uint32_t m = blkp[16 + bit];
      if (tCur > maxT) return finish_miss(o);
      o.t = tCur * s.voxel_size;
      float nn[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) nn[c] = (c == vax ? -1.0f : -0.0f) * (float)sd[c];
      o.n = mk3(nn[0], nn[1], nn[2]);
      o.mat = m;
      const uint32_t vx = (uint32_t)(bpos[0] * 8 + lpos[0]), vy = (uint32_t)(bpos[1] * 8 + lpos[1]),
                     vz = (uint32_t)(bpos[2] * 8 + lpos[2]);
      o.voxel = vx + vy * (uint32_t)s.nx + vz * (uint32_t)s.nx * (uint32_t)s.ny;
      const float nv = vax == 0 ? nn[0] : (vax == 1 ? nn[1] : nn[2]);
      o.face = nv > 0.0f ? (uint32_t)(2 * vax) : (nv < 0.0f ? (uint32_t)(2 * vax + 1) : 6u);
      o.bsteps = bsteps;
      o.vsteps = vsteps;
      return true;
    }
    advance(tMaxV, lpos, adl, &tCur, &vax);
    ++j;
    if ((unsigned)lpos[0] > 7u || (unsigned)lpos[1] > 7u || (unsigned)lpos[2] > 7u || j >= 24) {
      in_brick = false;
      if (next_brick(s)) return finish_miss(o);
    }
    return false;
  }

  VRT_D bool finish_miss(TraceResult& o) {
    o.t = -1.0f;
    o.mat = 0u;
    o.n = mk3(-rd[0], -rd[1], -rd[2]);
    o.voxel = 0xFFFFFFFFu;
    o.face = 255u;
    o.bsteps = bsteps;
    o.vsteps = vsteps;
    return true;
  }
};

}  // namespace vrt
