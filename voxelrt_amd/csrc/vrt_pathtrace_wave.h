// vrt_pathtrace_wave.h -- the base path tracer (pathtrace.wgsl:248-330) with
// step-level refill on persistent one-wave blocks.  Included by
// vrt_kernels.hip after the sampling helpers it uses.
//
// Why: a wave that traces one ray per lane to the end runs as long as its
// longest ray, and ray lengths in a voxel world spread widely (a shadow ray
// to the sky walks tens of bricks, a bounce into a wall a few voxels), so a
// quarter of the lanes were active per DDA instruction.  And inside the walk,
// the single-loop DDA's two long branches -- entering an occupied brick
// (~80 instructions) and leaving it for the coarse level (~45) -- are taken
// by some lane of a 64-lane wave in nearly every iteration, so the wave
// issued them every iteration for a few lanes.
//
// Here the step loop runs only the common iteration (vrt_dda_fast.h
// step_ev): a lane that reaches an occupied brick, leaves one, or finishes
// its ray stops and waits.  When `min_wait` lanes wait at a brick boundary,
// `shade_min` lanes have finished rays, or no lane is left stepping, the wave
// runs one pass in which the waiting lanes switch level together, and (when
// shade_min have finished, or nothing else can run) every finished lane
// shades its result and starts its next ray -- shadow, bounce, the next
// sample, or the first sample of a new pixel.  Then all lanes step on.  The
// long branches run once for many lanes.
//
// Exactness: each path performs the operations and RNG draws of
// pathtrace.wgsl in the same order, so every pixel equals the other path
// tracer forms bit for bit.  Two rearrangements keep that: (1) at a vertex
// the bounce direction and the Russian-roulette draw are taken before the
// shadow ray is traced (they come after the sun sample in the stream and do
// not depend on the shadow result), so the shadow ray only decides whether
// the precomputed NEE term is added; (2) a ray's world origin is read back
// from the DDA's voxel-unit origin (scaling by the power-of-two voxel size is
// exact both ways), so the path keeps no copy of it.
//
// Register budget: the DDA state of DdaFastT<GM, true> (~40 VGPRs) plus a
// packed path state
// (WavePath, 21) stay live across the loop; the shading code's temporaries
// exist only inside the service pass.  Each expensive helper (sky, sun
// sample, cosine sample, camera ray, DDA setup) has one call site.
#pragma once

namespace vrt {

// DDA steps a lane takes between two counts of the wave's waiting and
// finished lanes (4 measured fastest at 1080p, against 1, 2, 3, 6 and 8).
constexpr int STEPS_PER_CHECK = 4;

// Lane status: a DDA status while a ray is in flight or just finished
// (DDA_RUN / DDA_MISS / DDA_HIT), or one of these.
enum : uint32_t { LS_NEED = 8u, LS_IDLE = 9u };

// Path flags (WavePath::pf).
enum : uint32_t {
  PF_PRIMARY = 0u, PF_SHADOW = 1u, PF_BOUNCE = 2u,  // bits 0-1: the ray in flight
  PF_ALIVE = 4u,                                     // a bounce ray follows the shadow ray
  PF_BOUNCE_SHIFT = 8u,                              // bits 8-31: vertices after the primary
};

struct WavePath {
  uint32_t pix, smp, rng, pf;
  f3 tp, rad, alb1, bdir, neeC, acc;
};

// 8x8 patch b of a launch over `np` patches, in the order of
// SceneDev::tile_order's bit 1 (last patch first: the costly bottom rows).
VRT_D uint32_t wave_patch(uint32_t b, uint32_t np, uint32_t order) { return (order & 2u) ? np - 1u - b : b; }

// NT = 64: one wave per block, the DDA on global memory (DdaFastT<true>);
// NT = 256: four independent waves per block sharing the block's LDS copy of
// the scene's padded bitmap and prefix counts, each lane with an LDS slot for
// the mask of the brick it walks (DdaFastT<false>, every step an LDS read).
template <int WV, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WV))) void k_pathtrace_wave(
    SceneDev s, Uniforms u, float fclamp, uint32_t spp, ushort4* __restrict__ radiance, uchar4* __restrict__ albedo,
    float4* __restrict__ gbuf, unsigned long long* __restrict__ counters, uint32_t* __restrict__ queue,
    uint32_t min_wait, uint32_t shade_min) {
  constexpr bool GM = NT == 64;
  extern __shared__ uint32_t lds_dyn[];
  LdsScene L0{nullptr, 0u, 0u, 0u, 0u};  // the global-memory DDA reads no LDS
  if constexpr (!GM) L0 = lds_setup<TV_FAST>(s, lds_dyn).f;  // every thread of the block, one barrier
  const uint32_t w = (uint32_t)u.params1[0], h = (uint32_t)u.params1[1];
  const uint32_t pw = (w + 7u) >> 3, np = pw * ((h + 7u) >> 3);
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t lt_mask = lane == 0u ? 0ull : (~0ull >> (64u - lane));
  const uint32_t nb = u.params0[1];
  const bool indirect = (u.params0[2] & VRT_FLAG_INDIRECT) != 0u;

  DdaFastT<GM, true> st;
  WavePath P;
  P.pix = 0u;
  P.smp = 0u;
  P.rng = 0u;
  P.pf = 0u;
  P.tp = P.rad = P.alb1 = P.bdir = P.neeC = P.acc = mk3(0.0f, 0.0f, 0.0f);
  uint32_t status = LS_NEED;
  RayCount rc{0u, 0u, 0u, 0u};
  // the wave's current 8x8 patch and the next of its 64 pixels (uniform)
  uint32_t patch = 0u, cur = 64u;
  SplitQueue sq(queue, np, blockIdx.x);
  bool drained = false;

  for (;;) {
    // ---------------- level switches (vrt_dda_fast.h step_ev) -------------
    if (status == DDA_ENTER) {
      st.enter(s, L0);
      status = DDA_RUN;
    }
    if (status == DDA_LEAVE) status = st.resume(s, L0);
    // ---------------- service: results of finished rays -------------------
    // when shade_min lanes have finished rays, or no lane is left stepping
    const uint32_t nfin = (uint32_t)__popcll(__ballot(status == DDA_HIT || status == DDA_MISS));
    const bool shade = nfin >= shade_min || __ballot(status == DDA_RUN) == 0ull;
    const bool fin = shade && (status == DDA_HIT || status == DDA_MISS);
    const uint32_t kind = P.pf & 3u;
    bool hit = false;
    if (fin) {
      hit = status == DDA_HIT && !(st.t > st.maxT);  // a hit beyond maxT is a miss (voxel.wgsl:124)
      rc.rays += 1u;
      rc.bsteps += st.steps & 0xffffu;
      rc.vsteps += st.steps >> 16;
      rc.hits += hit ? 1u : 0u;
    }
    bool sample_done = false, new_primary = false, sky_sample = false, begin = false;
    f3 ro = mk3(0.0f, 0.0f, 0.0f), rd = mk3(0.0f, 0.0f, 1.0f);
    const f3 rdv = mk3(st.rdv[0], st.rdv[1], st.rdv[2]);
    const f3 rov = mk3(st.ro[0] * s.voxel_size, st.ro[1] * s.voxel_size, st.ro[2] * s.voxel_size);
    if (fin && kind == PF_SHADOW) {  // pathtrace.wgsl:304-306
      if (!hit) P.rad = P.rad + P.neeC;
      if (P.pf & PF_ALIVE) {
        P.pf = PF_BOUNCE | (P.pf & ~3u & ~PF_ALIVE);
        ro = rov;  // the shadow ray started at surf, where the bounce ray starts
        rd = P.bdir;
        begin = true;
      } else {
        sample_done = true;
      }
    }
    if (fin && kind != PF_SHADOW && !hit) {  // miss: sky (pathtrace.wgsl:267-276, 316-319)
      const f3 sky = sky_color(u, rdv, kind == PF_PRIMARY);
      if (kind == PF_PRIMARY) {
        P.rad = indirect ? mk3(0.0f, 0.0f, 0.0f) : sky;
        sky_sample = true;
        if (P.smp == 0u) {
          albedo[P.pix] = make_uchar4(255, 255, 255, 255);
          gbuf[P.pix] = make_float4(-rdv.x, -rdv.y, -rdv.z, -1.0f);
        }
      } else {
        P.rad = P.rad + P.tp * sky;
      }
      sample_done = true;
    }
    if (fin && kind != PF_SHADOW && hit) {  // path vertex (pathtrace.wgsl:279-315)
      const float tw = st.t * s.voxel_size;
      const uint32_t ax = st.axs >> 3;
      const f3 n = mk3((ax == 0u ? -1.0f : -0.0f) * (float)st.sd[0], (ax == 1u ? -1.0f : -0.0f) * (float)st.sd[1],
                       (ax == 2u ? -1.0f : -0.0f) * (float)st.sd[2]);
      const Material mat = unpack_material(s.pool[(uint64_t)st.blk + 16u + st.idx]);
      const f3 pos = rov + rdv * tw;
      uint32_t bounce = P.pf >> PF_BOUNCE_SHIFT;
      if (kind == PF_PRIMARY) {
        P.alb1 = mat.albedo;
        if (P.smp == 0u) {
          albedo[P.pix] = make_uchar4((uint8_t)f32_to_unorm8(vsqrt(mat.albedo.x)),
                                      (uint8_t)f32_to_unorm8(vsqrt(mat.albedo.y)),
                                      (uint8_t)f32_to_unorm8(vsqrt(mat.albedo.z)), 255);
          gbuf[P.pix] = make_float4(n.x, n.y, n.z, tw);
        }
        P.rad = mk3(0.0f, 0.0f, 0.0f);
        P.tp = mk3(1.0f, 1.0f, 1.0f);
        bounce = 0u;
      } else {
        bounce += 1u;
      }
      const bool direct = !(indirect && bounce == 0u);
      if (mat.emissive > 0.0f && direct) P.rad = P.rad + (P.tp * mat.albedo) * (mat.emissive * EMISSIVE_SCALE);
      const f3 surf = pos + n * 1e-3f;
      const f3 l = sample_sun_dir(u, P.rng);
      const float ndl = dot(n, l);
      const bool nee = ndl > 0.0f && l.y > 0.0f && direct;
      const f3 sunRad = mk3(u.sun_radiance[0], u.sun_radiance[1], u.sun_radiance[2]);
      P.neeC = (((P.tp * mat.albedo) * INV_PI_F) * ndl) * sunRad;
      bool alive = false;
      if (bounce != nb) {
        P.bdir = cosine_hemisphere(n, P.rng);
        P.tp = P.tp * mat.albedo;
        alive = true;
        if (bounce >= 2u) {
          const float p = vclamp(vmax(P.tp.x, vmax(P.tp.y, P.tp.z)), 0.05f, 0.95f);
          if (rand01(P.rng) > p) alive = false;
          else P.tp = P.tp / p;
        }
      }
      P.pf = (bounce << PF_BOUNCE_SHIFT) | (alive ? PF_ALIVE : 0u);
      if (nee) {
        P.pf |= PF_SHADOW;
        ro = surf;
        rd = l;
        begin = true;
      } else if (alive) {
        P.pf = (P.pf & ~PF_ALIVE) | PF_BOUNCE;
        ro = surf;
        rd = P.bdir;
        begin = true;
      } else {
        sample_done = true;
      }
    }
    if (sample_done) {  // pathtrace.wgsl:326-328
      f3 val = P.rad;
      if (!sky_sample) {
        if (!(fclamp <= 0.0f)) val = mk3(vmin(val.x, fclamp), vmin(val.y, fclamp), vmin(val.z, fclamp));
        val = val / mk3(vmax(P.alb1.x, 1e-3f), vmax(P.alb1.y, 1e-3f), vmax(P.alb1.z, 1e-3f));
      }
      P.acc = P.acc + val;
      P.smp += 1u;
      if (P.smp < spp) {
        new_primary = true;
      } else {
        const f3 o = P.acc / (float)spp;
        radiance[P.pix] = make_ushort4(f32_to_f16(o.x), f32_to_f16(o.y), f32_to_f16(o.z), f32_to_f16(1.0f));
        status = LS_NEED;
      }
    }
    // ---------------- service: pixels for lanes that need one ------------
    uint64_t need = __ballot(status == LS_NEED);
    while (need != 0ull && !drained) {
      if (cur >= 64u) {  // the wave takes the next patch
        const uint32_t b = sq.next(lane);
        if (b >= np) {
          drained = true;
          break;
        }
        patch = wave_patch(b, np, s.tile_order);
        cur = 0u;
      }
      const uint32_t cnt = (uint32_t)__popcll(need), room = 64u - cur;
      const uint32_t take = cnt < room ? cnt : room;
      if (status == LS_NEED) {
        const uint32_t rank = (uint32_t)__popcll(need & lt_mask);
        if (rank < take) {
          const uint32_t q = cur + rank;
          const uint32_t px = (patch % pw) * 8u + (q & 7u), py = (patch / pw) * 8u + (q >> 3);
          if (px < w && py < h) {  // else outside the image: ask again
            P.pix = py * w + px;
            P.smp = 0u;
            P.acc = mk3(0.0f, 0.0f, 0.0f);
            status = DDA_MISS;  // placeholder until the ray begins
            new_primary = true;
          }
        }
      }
      cur += take;
      need = __ballot(status == LS_NEED);
    }
    if (status == LS_NEED && drained) status = LS_IDLE;
    // ---------------- service: camera rays (pathtrace.wgsl:255-262) -------
    if (new_primary) {
      const uint32_t px = P.pix % w, py = P.pix / w;
      P.rng = init_rng(px, py, u.params0[0] * spp + P.smp);
      const float jx = rand01(P.rng) - 0.5f;
      const float jy = rand01(P.rng) - 0.5f;
      const float uvx = (((float)px + 0.5f) + jx) / (float)w;
      const float uvy = (((float)py + 0.5f) + jy) / (float)h;
      ro = mk3(u.cam_pos[0], u.cam_pos[1], u.cam_pos[2]);
      rd = camera_ray(u, uvx, uvy);
      P.pf = PF_PRIMARY;
      begin = true;
    }
    // ---------------- service: ray setup (voxel.wgsl:79-104) -------------
    if (begin) {
      TraceResult unused;
      if (st.begin(s, L0, ro, rd, 1e4f, unused)) {
        status = DDA_MISS;  // misses the grid: serviced next round
        st.steps = 0u;
      } else {
        status = DDA_RUN;
      }
    }
    if (__ballot(status != LS_IDLE) == 0ull) break;
    // ---------------- stepping --------------------------------------------
    // until min_wait lanes wait at a brick boundary, shade_min lanes have a
    // finished ray, or no lane steps; the counts are taken every
    // STEPS_PER_CHECK steps (a lane that stops within them waits for the
    // rest: the ballots and compares cost about as much as a step)
    for (;;) {
#pragma unroll
      for (int k = 0; k < STEPS_PER_CHECK; ++k)
        if (status == DDA_RUN) status = st.step_ev(s, L0);
      if (__ballot(status == DDA_RUN) == 0ull) break;
      if ((uint32_t)__popcll(__ballot(status == DDA_ENTER || status == DDA_LEAVE)) >= min_wait) break;
      if ((uint32_t)__popcll(__ballot(status == DDA_HIT || status == DDA_MISS)) >= shade_min) break;
    }
  }
  count4(counters, rc.rays, rc.bsteps, rc.vsteps, rc.hits);
}

}  // namespace vrt
