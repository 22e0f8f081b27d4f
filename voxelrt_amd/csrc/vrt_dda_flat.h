// vrt_dda_flat.h -- the two-level DDA of voxel.wgsl:73-153 as a flat,
// resumable state with ONE instruction stream for both levels.
//
// A step tests the cell under the active DDA (a brick word at the coarse
// level, a mask bit at the fine level) and advances it.  The level only
// selects strides, base pointer, cell test, per-axis delta and bounds, so a
// wave whose lanes sit at different levels still issues each instruction
// once for all of them.  Branches remain only for the rare events: entering
// an occupied brick, leaving a brick, the hit, the miss.
//
// Exactness: every tMax/t value is produced by the same FP32 operations in
// the same order as vrt::trace() (and the oracle): the coarse advance adds
// |invRd|*8, the fine advance |invRd|, ties go x, then y, then z; brick entry
// recomputes the fine tMax from the origin.  Step counts follow the
// reference's loop iterations: one brick word per coarse test, one mask word
// per fine test.
#pragma once
#include "vrt_trace.h"

namespace vrt {

struct DdaFlat {
  // ray constants (voxel units)
  float ro[3], rd[3], inv[3];
  int sd[3];
  float tExit, maxT;
  // active level: 0 = bricks, 1 = voxels inside brick `blk`
  float tm[3];
  int pos[3];
  float tcur;
  int axis;
  uint32_t level, vj, iter, blk;
  // the coarse DDA while the fine one runs
  float tmB[3];
  int posB[3];
  float tB;
  int axB;
  uint32_t bsteps, vsteps;

  VRT_D void fill_miss(TraceResult& o) const {
    o.t = -1.0f;
    o.mat = 0u;
    o.n = mk3(-rd[0], -rd[1], -rd[2]);
    o.voxel = 0xFFFFFFFFu;
    o.face = 255u;
    o.bsteps = bsteps;
    o.vsteps = vsteps;
  }

  // voxel.wgsl:79-104.  Returns true when the ray misses the grid at setup.
  VRT_D bool begin(const SceneDev& s, f3 roW, f3 rdIn, float maxTWorld, TraceResult& o) {
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
      sd[c] = sign_i(rd[c]);
      const float t0 = (s.zero - ro[c]) * inv[c];
      const float t1 = ((float)G[c] - ro[c]) * inv[c];
      tmin[c] = vmin(t0, t1);
      tmax[c] = vmax(t0, t1);
    }
    const float tEnter = vmax(vmax(tmin[0], tmin[1]), vmax(tmin[2], 0.0f));
    tExit = vmin(vmin(tmax[0], tmax[1]), vmin(tmax[2], maxT));
    if (tEnter > tExit) {
      fill_miss(o);
      return true;
    }
    axis = (tmin[0] >= tmin[1] && tmin[0] >= tmin[2]) ? 0 : (tmin[1] >= tmin[2] ? 1 : 2);
    tcur = tEnter;
    const float ts = tcur + 1e-4f;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float sp = ro[c] + rd[c] * ts;
      pos[c] = floor_clamp(sp / 8.0f, 0, BG[c] - 1);
      const float ge = rd[c] >= 0.0f ? 1.0f : 0.0f;
      tm[c] = (((float)pos[c] + ge) * 8.0f - ro[c]) * inv[c];
    }
    level = 0u;
    vj = 0u;
    iter = 0u;
    return false;
  }

  // Amanatides-Woo advance of the active level (ties x, then y, then z).
  VRT_D void advance(float scale) {
    const bool cx = tm[0] <= tm[1] && tm[0] <= tm[2];
    const bool cy = !cx && tm[1] <= tm[2];
    const bool cz = !cx && !cy;
    tcur = cx ? tm[0] : (cy ? tm[1] : tm[2]);
    axis = cx ? 0 : (cy ? 1 : 2);
    const float n0 = tm[0] + vabs(inv[0]) * scale;
    const float n1 = tm[1] + vabs(inv[1]) * scale;
    const float n2 = tm[2] + vabs(inv[2]) * scale;
    tm[0] = cx ? n0 : tm[0];
    tm[1] = cy ? n1 : tm[1];
    tm[2] = cz ? n2 : tm[2];
    pos[0] += cx ? sd[0] : 0;
    pos[1] += cy ? sd[1] : 0;
    pos[2] += cz ? sd[2] : 0;
  }

  // Coarse advance + grid bound + iteration bound; true = the ray missed.
  VRT_D bool brick_advance(const SceneDev& s) {
    advance(8.0f);  // |invRd| * 8 == tDeltaB (voxel.wgsl:103)
    ++iter;
    return (unsigned)pos[0] >= (unsigned)s.bgx || (unsigned)pos[1] >= (unsigned)s.bgy ||
           (unsigned)pos[2] >= (unsigned)s.bgz || iter >= (uint32_t)(s.bgx + s.bgy + s.bgz + 2);
  }

  // One loop iteration of the reference.  Returns true when the ray finished
  // (result in o).
  VRT_D bool step(const SceneDev& s, TraceResult& o) {
    const bool fine = level != 0u;
    if (!fine && tcur > tExit) {
      fill_miss(o);
      return true;
    }
    // cell index: brick word (x + y*bgx + z*bgx*bgy) or mask bit (x + 8y + 64z)
    const uint32_t sy = fine ? 8u : (uint32_t)s.bgx;
    const uint32_t sz = fine ? 64u : (uint32_t)s.bgx * (uint32_t)s.bgy;
    const uint32_t idx = (uint32_t)pos[0] + (uint32_t)pos[1] * sy + (uint32_t)pos[2] * sz;
    const uint32_t* base = fine ? s.pool + blk : s.bricks;
    const uint32_t word = base[fine ? (idx >> 5) : idx];
    bsteps += fine ? 0u : 1u;
    vsteps += fine ? 1u : 0u;
    const bool occ = fine ? ((word >> (idx & 31u)) & 1u) != 0u : word != 0u;
    if (occ) {
      if (fine) {  // hit (voxel.wgsl:122-128)
        const uint32_t m = s.pool[blk + 16u + idx];
        if (tcur > maxT) {
          fill_miss(o);
          return true;
        }
        o.t = tcur * s.voxel_size;
        float nn[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) nn[c] = (c == axis ? -1.0f : -0.0f) * (float)sd[c];
        o.n = mk3(nn[0], nn[1], nn[2]);
        o.mat = m;
        const uint32_t vx = (uint32_t)(posB[0] * 8 + pos[0]), vy = (uint32_t)(posB[1] * 8 + pos[1]),
                       vz = (uint32_t)(posB[2] * 8 + pos[2]);
        o.voxel = vx + vy * (uint32_t)s.nx + vz * (uint32_t)s.nx * (uint32_t)s.ny;
        const float nv = axis == 0 ? nn[0] : (axis == 1 ? nn[1] : nn[2]);
        o.face = nv > 0.0f ? (uint32_t)(2 * axis) : (nv < 0.0f ? (uint32_t)(2 * axis + 1) : 6u);
        o.bsteps = bsteps;
        o.vsteps = vsteps;
        return true;
      }
      // enter the occupied brick (voxel.wgsl:111-118): save the coarse DDA
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        tmB[c] = tm[c];
        posB[c] = pos[c];
      }
      tB = tcur;
      axB = axis;
      blk = (word - 1u) * (uint32_t)POOL_WORDS;
      const float ts = tcur + 1e-4f;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const int bmin = posB[c] * 8;
        const float p = ro[c] + rd[c] * ts;
        const int vp = floor_clamp(p, bmin, bmin + 7);
        const float ge = rd[c] >= 0.0f ? 1.0f : 0.0f;
        pos[c] = vp - bmin;
        tm[c] = (((float)vp + ge) - ro[c]) * inv[c];
      }
      level = 1u;
      vj = 0u;
      return false;
    }
    if (fine) {
      ++vj;
      bool leave = vj >= 24u;
      if (!leave) {
        advance(1.0f);
        leave = (unsigned)pos[0] > 7u || (unsigned)pos[1] > 7u || (unsigned)pos[2] > 7u;
      }
      if (!leave) return false;
      // back to the coarse DDA (voxel.wgsl:142-150 follows the fine loop)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        tm[c] = tmB[c];
        pos[c] = posB[c];
      }
      tcur = tB;
      axis = axB;
      level = 0u;
    }
    if (brick_advance(s)) {
      fill_miss(o);
      return true;
    }
    return false;
  }
};

// Whole-ray trace on the flat state: same results as vrt::trace(), but a wave
// whose lanes are at different levels does not serialise the two loops.
VRT_D void trace_flat(const SceneDev& s, f3 roW, f3 rd, float maxTWorld, TraceResult& o) {
  DdaFlat st;
  if (st.begin(s, roW, rd, maxTWorld, o)) return;
  while (!st.step(s, o)) {
  }
}

}  // namespace vrt
