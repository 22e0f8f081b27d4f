// vrt_pathtrace_persistent.h -- the base path tracer (pathtrace.wgsl:248-330)
// as a persistent, per-lane state machine.  Included by vrt_kernels.hip after
// the sampling helpers it uses.
//
// Each lane owns one pixel (all of its samples) and one ray in flight.  The
// loop of a wave:
//   1. lanes without a pixel take pixel indices from a chunked global queue;
//   2. lanes whose ray finished (or that just got a pixel) run the shading
//      transition -- primary hit / shadow result / bounce hit, the sun-cone
//      NEE and cosine-bounce sampling, Russian roulette, the clamp and the
//      demodulation -- up to the point where they begin their next ray or
//      finish the pixel.  Transitions run only once at least `min_refill`
//      lanes wait for one (or nothing is in flight), so the long shading code
//      is issued for many lanes at once;
//   3. lanes with a ray in flight advance their DDA `steps` iterations.
// Every path consumes its PCG stream in exactly the order of the reference
// loop, so results equal the one-pixel-per-lane kernel bit for bit.  The DDA
// is a Runner of vrt_kernels.hip: FlatRunner (single-stream, occupancy read
// from HBM/L2) or FastRunner (single loop on the LDS-staged occupancy; its
// brick entries wait for SceneDev::entry_batch lanes of the wave).
#pragma once

namespace vrt {

enum : uint32_t {
  PH_NONE = 0,     // lane needs a pixel
  PH_PRIMARY = 1,  // primary ray in flight
  PH_SHADOW = 2,   // sun shadow ray in flight
  PH_BOUNCE = 3,   // bounce ray in flight
  PH_DONE = 4,     // queue exhausted
};

struct PathLane {
  uint32_t pix, smp, rng, phase, bounce;
  f3 radiance, tp, surf, n, alb, alb1, acc;
  float emis, ndl;
  bool sky_sample;
  bool fin;  // ray finished, transition pending
};

// Camera ray of sample L.smp of pixel L.pix (pathtrace.wgsl:255-262).
VRT_D f3 primary_dir(const Uniforms& u, uint32_t spp, PathLane& L, uint32_t w, uint32_t h) {
  const uint32_t px = L.pix % w, py = L.pix / w;
  L.rng = init_rng(px, py, u.params0[0] * spp + L.smp);
  const float jx = rand01(L.rng) - 0.5f;
  const float jy = rand01(L.rng) - 0.5f;
  const float uvx = (((float)px + 0.5f) + jx) / (float)w;
  const float uvy = (((float)py + 0.5f) + jy) / (float)h;
  return camera_ray(u, uvx, uvy);
}

// Run the path from a finished ray (or a freshly assigned pixel) to the next
// ray begin, or to the end of the pixel (phase -> PH_NONE).  Written with one
// ray-begin site and one sky evaluation so each is emitted once.
template <class Runner>
VRT_D void path_transition(const SceneDev& s, const LaneMem& M, const Uniforms& u, float fclamp, uint32_t spp,
                           PathLane& L, Runner& st, TraceResult& o, bool have_result, ushort4* __restrict__ radiance_out,
                           uchar4* __restrict__ albedo_out, float4* __restrict__ gbuf_out, RayCount& rc,
                           uint32_t w, uint32_t h) {
  const uint32_t nb = u.params0[1];
  const bool indirect = (u.params0[2] & VRT_FLAG_INDIRECT) != 0u;
  enum { S_RESULT, S_VERTEX, S_AFTER_SHADOW, S_SAMPLE_DONE, S_START, S_PRIMARY, S_BEGIN } next =
      have_result ? S_RESULT : S_START;
  f3 ray_o = mk3(0.0f, 0.0f, 0.0f), ray_d = mk3(0.0f, 0.0f, 1.0f);
  // Terminates: every state either returns (a ray is in flight, or the pixel
  // is finished) or moves to a later state of the same sample, or to the next
  // of at most spp samples.
  for (;;) {
    if (next == S_BEGIN) {
      if (!st.begin(s, M, ray_o, ray_d, 1e4f, o)) return;
      next = S_RESULT;  // missed the grid at setup
    }
    if (next == S_START) {
      L.smp = 0u;
      L.acc = mk3(0.0f, 0.0f, 0.0f);
      next = S_PRIMARY;
    }
    if (next == S_PRIMARY) {
      L.phase = PH_PRIMARY;
      ray_o = mk3(u.cam_pos[0], u.cam_pos[1], u.cam_pos[2]);
      ray_d = primary_dir(u, spp, L, w, h);
      next = S_BEGIN;
      continue;
    }
    if (next == S_RESULT) {
      rc.add(o);
      const f3 rd = st.dir();
      if (o.t < 0.0f && L.phase != PH_SHADOW) {
        // primary miss: sky with the sun disc; bounce miss: sky without
        const f3 sky = sky_color(u, rd, L.phase == PH_PRIMARY);
        if (L.phase == PH_PRIMARY) {
          L.radiance = indirect ? mk3(0.0f, 0.0f, 0.0f) : sky;
          L.sky_sample = true;
          if (L.smp == 0u) {
            albedo_out[L.pix] = make_uchar4(255, 255, 255, 255);
            gbuf_out[L.pix] = make_float4(-rd.x, -rd.y, -rd.z, -1.0f);
          }
        } else {
          L.radiance = L.radiance + L.tp * sky;
        }
        next = S_SAMPLE_DONE;
        continue;
      }
      if (L.phase == PH_SHADOW) {
        if (!(o.t >= 0.0f)) {
          const f3 sunRad = mk3(u.sun_radiance[0], u.sun_radiance[1], u.sun_radiance[2]);
          L.radiance = L.radiance + (((L.tp * L.alb) * INV_PI_F) * L.ndl) * sunRad;
        }
        next = S_AFTER_SHADOW;
        continue;
      }
      const Material m = unpack_material(o.mat);
      if (L.phase == PH_PRIMARY) {
        L.alb1 = m.albedo;
        if (L.smp == 0u) {
          albedo_out[L.pix] = make_uchar4((uint8_t)f32_to_unorm8(vsqrt(m.albedo.x)),
                                          (uint8_t)f32_to_unorm8(vsqrt(m.albedo.y)),
                                          (uint8_t)f32_to_unorm8(vsqrt(m.albedo.z)), 255);
          gbuf_out[L.pix] = make_float4(o.n.x, o.n.y, o.n.z, o.t);
        }
        L.radiance = mk3(0.0f, 0.0f, 0.0f);
        L.tp = mk3(1.0f, 1.0f, 1.0f);
        L.surf = mk3(u.cam_pos[0], u.cam_pos[1], u.cam_pos[2]);  // pos = ro + rd * t
        L.bounce = 0u;
        L.sky_sample = false;
      } else {
        L.bounce += 1u;  // bounce hit: pos = surf + rd * t
      }
      L.surf = L.surf + rd * o.t;
      L.n = o.n;
      L.alb = m.albedo;
      L.emis = m.emissive;
      next = S_VERTEX;
      continue;
    }
    if (next == S_VERTEX) {  // pathtrace.wgsl:298-306, L.surf holds pos
      const bool direct = !(indirect && L.bounce == 0u);
      if (L.emis > 0.0f && direct) L.radiance = L.radiance + (L.tp * L.alb) * (L.emis * EMISSIVE_SCALE);
      L.surf = L.surf + L.n * 1e-3f;
      const f3 l = sample_sun_dir(u, L.rng);
      L.ndl = dot(L.n, l);
      if (L.ndl > 0.0f && l.y > 0.0f && direct) {
        L.phase = PH_SHADOW;
        ray_o = L.surf;
        ray_d = l;
        next = S_BEGIN;
        continue;
      }
      next = S_AFTER_SHADOW;
      continue;
    }
    if (next == S_AFTER_SHADOW) {  // pathtrace.wgsl:307-316
      if (L.bounce == nb) {
        next = S_SAMPLE_DONE;
        continue;
      }
      const f3 rd = cosine_hemisphere(L.n, L.rng);
      L.tp = L.tp * L.alb;
      if (L.bounce >= 2u) {
        const float p = vclamp(vmax(L.tp.x, vmax(L.tp.y, L.tp.z)), 0.05f, 0.95f);
        if (rand01(L.rng) > p) {
          next = S_SAMPLE_DONE;
          continue;
        }
        L.tp = L.tp / p;
      }
      L.phase = PH_BOUNCE;
      ray_o = L.surf;
      ray_d = rd;
      next = S_BEGIN;
      continue;
    }
    // S_SAMPLE_DONE: pathtrace.wgsl:326-328 (a sky sample is stored as is)
    f3 val = L.radiance;
    if (!L.sky_sample) {
      if (!(fclamp <= 0.0f)) val = mk3(vmin(val.x, fclamp), vmin(val.y, fclamp), vmin(val.z, fclamp));
      val = val / mk3(vmax(L.alb1.x, 1e-3f), vmax(L.alb1.y, 1e-3f), vmax(L.alb1.z, 1e-3f));
    }
    L.acc = L.acc + val;
    L.smp += 1u;
    if (L.smp < spp) {
      next = S_PRIMARY;
      continue;
    }
    const f3 out = L.acc / (float)spp;
    radiance_out[L.pix] = make_ushort4(f32_to_f16(out.x), f32_to_f16(out.y), f32_to_f16(out.z), f32_to_f16(1.0f));
    L.phase = PH_NONE;
    return;
  }
}

// WV: minimum waves per SIMD the register allocation must allow (1 = the
// compiler's choice), as for k_pathtrace.
template <class Runner, int TV, int WV>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WV))) void k_pathtrace_persistent(
    SceneDev s, Uniforms u, float fclamp, uint32_t spp, ushort4* __restrict__ radiance, uchar4* __restrict__ albedo,
    float4* __restrict__ gbuf, unsigned long long* __restrict__ counters, uint32_t* __restrict__ queue, int steps,
    int min_refill, uint32_t chunk) {
  extern __shared__ uint32_t lds_dyn[];
  const LaneMem M = lds_setup<TV>(s, lds_dyn);
  const uint32_t w = (uint32_t)u.params1[0], h = (uint32_t)u.params1[1];
  // queue ids walk 8x8 pixel tiles row-major, 64 ids per tile, so a wave's
  // chunk covers compact tiles (coherent primary rays)
  const uint32_t tiles_x = (w + 7u) >> 3, tiles_y = (h + 7u) >> 3;
  const uint32_t nids = tiles_x * tiles_y * 64u;
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t lt_mask = lane == 0u ? 0ull : (~0ull >> (64u - lane));
  PathLane L;
  L.phase = PH_NONE;
  L.fin = false;
  Runner st;
  TraceResult o;
  RayCount rc{0u, 0u, 0u, 0u};
  uint32_t cbase = 0u, cleft = 0u;
  for (;;) {
    const bool in_flight = L.phase == PH_PRIMARY || L.phase == PH_SHADOW || L.phase == PH_BOUNCE;
    const uint64_t busy = __ballot(in_flight && !L.fin);
    const uint64_t need = __ballot(L.phase == PH_NONE);
    const uint64_t waiting = __ballot(L.fin) | need;
    if (waiting != 0ull && (busy == 0ull || __popcll(waiting) >= min_refill)) {
      // hand out pixels to lanes that need one
      bool fresh = false;
      if (need != 0ull) {
        if (cleft == 0u) {
          uint32_t b = 0u;
          if (lane == 0u) b = atomicAdd(queue, chunk);
          cbase = __shfl(b, 0, 64);
          cleft = chunk;
        }
        const uint32_t cnt = (uint32_t)__popcll(need);
        const uint32_t take = cnt < cleft ? cnt : cleft;
        const uint32_t rank = (uint32_t)__popcll(need & lt_mask);
        const uint32_t base = cbase;
        cbase += take;
        cleft -= take;
        if (L.phase == PH_NONE && rank < take) {
          const uint32_t my = base + rank;
          if (my < nids) {
            const uint32_t t = my >> 6, q = my & 63u;
            const uint32_t px = (t % tiles_x) * 8u + (q & 7u);
            const uint32_t py = (t / tiles_x) * 8u + (q >> 3);
            if (px < w && py < h) {  // else: id outside the image, ask again
              L.pix = py * w + px;
              fresh = true;
            }
          } else {
            L.phase = PH_DONE;
          }
        }
      }
      if (L.fin || fresh) {
        path_transition(s, M, u, fclamp, spp, L, st, o, L.fin, radiance, albedo, gbuf, rc, w, h);
        L.fin = false;
      }
    }
    if (__ballot(L.phase != PH_DONE) == 0ull) break;
    for (int k = 0; k < steps; ++k) {
      const bool walking = !L.fin && (L.phase == PH_PRIMARY || L.phase == PH_SHADOW || L.phase == PH_BOUNCE);
      if (walking) {
        if (st.step(s, M, o)) L.fin = true;
      }
      st.settle(s, M, walking);
    }
  }
  count4(counters, rc.rays, rc.bsteps, rc.vsteps, rc.hits);
}

}  // namespace vrt
