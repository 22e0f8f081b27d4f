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
// finished lanes (6 measured fastest at 1080p, against 3, 4, 5 and 8, once
// rays ended at the occupied boxes; 4 before, when walks were longer).
constexpr int STEPS_PER_CHECK = 6;

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

// What a wave renders: one view (WaveView1) or the views of a batch table
// (WaveViewBatch).  Both give the view's scene, uniforms, sizes and outputs,
// and the next 8x8 patch: PATCH_OK (in `patch`, view-local), PATCH_NONE (no
// work left) or PATCH_SWITCH (the next work is another view's: the wave
// finishes every path in flight, then switch() moves it there).
enum : uint32_t { PATCH_OK = 0u, PATCH_NONE = 1u, PATCH_SWITCH = 2u };

struct WaveView1 {
  static constexpr bool MULTI = false;
  const SceneDev& sc;
  const Uniforms& un;
  float fclamp;
  uint32_t spp;
  ushort4* radiance;
  uchar4* albedo;
  float4* gbuf;
  uint32_t w, h, pw, np, nb;
  bool indirect;
  SplitQueue sq;
  VRT_D WaveView1(const SceneDev& s_, const Uniforms& u_, float fc, uint32_t spp_, ushort4* r, uchar4* a, float4* g,
                  uint32_t* queue)
      : sc(s_), un(u_), fclamp(fc), spp(spp_), radiance(r), albedo(a), gbuf(g),
        w((uint32_t)u_.params1[0]), h((uint32_t)u_.params1[1]), pw((w + 7u) >> 3), np(pw * ((h + 7u) >> 3)),
        nb(u_.params0[1]), indirect((u_.params0[2] & VRT_FLAG_INDIRECT) != 0u), sq(queue, np, blockIdx.x) {}
  VRT_D const SceneDev& s() const { return sc; }
  VRT_D const Uniforms& u() const { return un; }
  VRT_D uint32_t next(uint32_t lane, uint32_t& patch) {
    const uint32_t b = sq.next(lane);
    if (b >= np) return PATCH_NONE;
    patch = wave_patch(b, np, sc.tile_order);
    return PATCH_OK;
  }
  VRT_D uint32_t switch_view(uint32_t&) { return PATCH_NONE; }
  VRT_D void flush(RayCount&, uint32_t) {}
};

// The batched form's view source.  Work items are chunks of BatchHead::wchunk
// patches of one view (ViewDesc::wchunk0: a view's first chunk), drawn from a
// SplitQueue over the table's BatchHead::wchunks; a wave stays on one view
// until its next chunk belongs to another.

VRT_D uint32_t find_view_w(const ViewDesc* __restrict__ views, uint32_t nviews, uint32_t c) {
  uint32_t lo = 0, hi = nviews;  // views sorted by wchunk0, wchunk0[0] = 0
  while (hi - lo > 1u) {
    const uint32_t mid = (lo + hi) >> 1;
    if (views[mid].wchunk0 <= c) lo = mid;
    else hi = mid;
  }
  return lo;
}

struct WaveViewBatch {
  static constexpr bool MULTI = true;
  const ViewDesc* __restrict__ views;
  uint32_t nviews, nchunks, chunk;
  uint32_t vi = 0u, pend = 0u, pend_first = 0u;  // current view; the next one and its first patch
  uint32_t next_p = 0u, left = 0u;               // next patch of the current chunk, patches left in it
  float fclamp = 0.0f;
  uint32_t spp = 1u;
  ushort4* radiance = nullptr;
  uchar4* albedo = nullptr;
  float4* gbuf = nullptr;
  uint32_t w = 0u, h = 0u, pw = 0u, np = 0u, nb = 0u;
  bool indirect = false;
  SplitQueue sq;
  VRT_D WaveViewBatch(const BatchHead* head, const ViewDesc* v, uint32_t* queue)
      : views(v), nviews(head->nviews), nchunks(head->wchunks), chunk(head->wchunk),
        sq(queue, head->wchunks, blockIdx.x) {}
  VRT_D const ViewDesc& d() const { return views[vi]; }
  VRT_D const SceneDev& s() const { return d().s; }
  VRT_D const Uniforms& u() const { return d().u; }
  VRT_D void load(uint32_t k, uint32_t first) {
    vi = k;
    const ViewDesc& dv = views[k];
    fclamp = dv.fclamp;
    spp = dv.spp;
    radiance = (ushort4*)dv.radiance;
    albedo = (uchar4*)dv.albedo;
    gbuf = (float4*)dv.gbuf_cur;
    w = dv.w;
    h = dv.h;
    pw = (w + 7u) >> 3;
    np = pw * ((h + 7u) >> 3);
    nb = dv.u.params0[1];
    indirect = (dv.u.params0[2] & VRT_FLAG_INDIRECT) != 0u;
    next_p = first;
    left = np - first < chunk ? np - first : chunk;
  }
  VRT_D uint32_t next(uint32_t lane, uint32_t& patch) {
    if (left == 0u) {
      const uint32_t c = sq.next(lane);
      if (c >= nchunks) return PATCH_NONE;
      const uint32_t k = __builtin_amdgcn_readfirstlane(find_view_w(views, nviews, c));
      const uint32_t first = (c - views[k].wchunk0) * chunk;
      if (k != vi || np == 0u) {
        pend = k;
        pend_first = first;
        return PATCH_SWITCH;
      }
      next_p = first;
      left = np - first < chunk ? np - first : chunk;
    }
    patch = wave_patch(next_p, np, s().tile_order);
    ++next_p;
    --left;
    return PATCH_OK;
  }
  // Called by the whole wave once no path of the current view is in flight.
  VRT_D uint32_t switch_view(uint32_t& patch) {
    load(pend, pend_first);
    patch = wave_patch(next_p, np, s().tile_order);
    ++next_p;
    --left;
    return PATCH_OK;
  }
  // The current view's share of the counts so far (per-view counters).
  VRT_D void flush(RayCount& rc, uint32_t) {
    if (np != 0u && d().vcount) count4(d().vcount, rc.rays, rc.bsteps, rc.vsteps, rc.hits);
  }
};

// The loop of the step-level refill path tracer over a view source V.
template <bool GM, class V>
VRT_D void wave_paths(V& v, const LdsScene& L0, uint32_t min_wait, uint32_t shade_min,
                      unsigned long long* __restrict__ counters) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t lt_mask = lane == 0u ? 0ull : (~0ull >> (64u - lane));
  // rays end at the box of occupied bricks, fine walks at the entered
  // brick's box of solid voxels; no iteration budget
  DdaFastT<GM, true, true, true, true> st;
  WavePath P;
  P.pix = 0u;
  P.smp = 0u;
  P.rng = 0u;
  P.pf = 0u;
  P.tp = P.rad = P.alb1 = P.bdir = P.neeC = P.acc = mk3(0.0f, 0.0f, 0.0f);
  uint32_t status = LS_NEED;
  RayCount rc{0u, 0u, 0u, 0u};
  // the wave's current 8x8 patch and the next of its 64 pixels (uniform);
  // pending: the next work is another view's (batched form)
  uint32_t patch = 0u, cur = 64u;
  bool drained = false, pending = false;
  RayCount total{0u, 0u, 0u, 0u};

  for (;;) {
    // ---------------- service: results of finished rays -------------------
    // when shade_min lanes have finished rays, or no lane is left stepping
    // (lanes stopped at a level switch step on after this service)
    const uint32_t nfin = (uint32_t)__popcll(__ballot(status == DDA_HIT || status == DDA_MISS));
    const bool shade =
        nfin >= shade_min || __ballot(status == DDA_RUN || status == DDA_ENTER || status == DDA_LEAVE) == 0ull;
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
    const f3 rov = mk3(st.ro[0] * v.s().voxel_size, st.ro[1] * v.s().voxel_size, st.ro[2] * v.s().voxel_size);
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
      const f3 sky = sky_color(v.u(), rdv, kind == PF_PRIMARY);
      if (kind == PF_PRIMARY) {
        P.rad = v.indirect ? mk3(0.0f, 0.0f, 0.0f) : sky;
        sky_sample = true;
        if (P.smp == 0u) {
          v.albedo[P.pix] = make_uchar4(255, 255, 255, 255);
          v.gbuf[P.pix] = make_float4(-rdv.x, -rdv.y, -rdv.z, -1.0f);
        }
      } else {
        P.rad = P.rad + P.tp * sky;
      }
      sample_done = true;
    }
    if (fin && kind != PF_SHADOW && hit) {  // path vertex (pathtrace.wgsl:279-315)
      const float tw = st.t * v.s().voxel_size;
      const uint32_t ax = st.axs >> 3;
      const f3 n = mk3((ax == 0u ? -1.0f : -0.0f) * (float)st.sd[0], (ax == 1u ? -1.0f : -0.0f) * (float)st.sd[1],
                       (ax == 2u ? -1.0f : -0.0f) * (float)st.sd[2]);
      const Material mat = unpack_material(v.s().pool[(uint64_t)st.blk + 16u + st.idx]);
      const f3 pos = rov + rdv * tw;
      uint32_t bounce = P.pf >> PF_BOUNCE_SHIFT;
      if (kind == PF_PRIMARY) {
        P.alb1 = mat.albedo;
        if (P.smp == 0u) {
          v.albedo[P.pix] = make_uchar4((uint8_t)f32_to_unorm8(vsqrt(mat.albedo.x)),
                                      (uint8_t)f32_to_unorm8(vsqrt(mat.albedo.y)),
                                      (uint8_t)f32_to_unorm8(vsqrt(mat.albedo.z)), 255);
          v.gbuf[P.pix] = make_float4(n.x, n.y, n.z, tw);
        }
        P.rad = mk3(0.0f, 0.0f, 0.0f);
        P.tp = mk3(1.0f, 1.0f, 1.0f);
        bounce = 0u;
      } else {
        bounce += 1u;
      }
      const bool direct = !(v.indirect && bounce == 0u);
      if (mat.emissive > 0.0f && direct) P.rad = P.rad + (P.tp * mat.albedo) * (mat.emissive * EMISSIVE_SCALE);
      const f3 surf = pos + n * 1e-3f;
      const f3 l = sample_sun_dir(v.u(), P.rng);
      const float ndl = dot(n, l);
      const bool nee = ndl > 0.0f && l.y > 0.0f && direct;
      const f3 sunRad = mk3(v.u().sun_radiance[0], v.u().sun_radiance[1], v.u().sun_radiance[2]);
      P.neeC = (((P.tp * mat.albedo) * INV_PI_F) * ndl) * sunRad;
      bool alive = false;
      if (bounce != v.nb) {
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
        if (!(v.fclamp <= 0.0f)) val = mk3(vmin(val.x, v.fclamp), vmin(val.y, v.fclamp), vmin(val.z, v.fclamp));
        val = val / mk3(vmax(P.alb1.x, 1e-3f), vmax(P.alb1.y, 1e-3f), vmax(P.alb1.z, 1e-3f));
      }
      P.acc = P.acc + val;
      P.smp += 1u;
      if (P.smp < v.spp) {
        new_primary = true;
      } else {
        const f3 o = P.acc / (float)v.spp;
        v.radiance[P.pix] = make_ushort4(f32_to_f16(o.x), f32_to_f16(o.y), f32_to_f16(o.z), f32_to_f16(1.0f));
        status = LS_NEED;
      }
    }
    // ---------------- service: pixels for lanes that need one ------------
    for (;;) {
    uint64_t need = __ballot(status == LS_NEED);
    while (need != 0ull && !drained && !pending) {
      if (cur >= 64u) {  // the wave takes the next patch
        const uint32_t r = v.next(lane, patch);
        if (r == PATCH_NONE) {
          drained = true;
          break;
        }
        if (V::MULTI && r == PATCH_SWITCH) {
          pending = true;
          break;
        }
        cur = 0u;
      }
      const uint32_t cnt = (uint32_t)__popcll(need), room = 64u - cur;
      const uint32_t take = cnt < room ? cnt : room;
      if (status == LS_NEED) {
        const uint32_t rank = (uint32_t)__popcll(need & lt_mask);
        if (rank < take) {
          const uint32_t q = cur + rank;
          const uint32_t px = (patch % v.pw) * 8u + (q & 7u), py = (patch / v.pw) * 8u + (q >> 3);
          if (px < v.w && py < v.h) {  // else outside the image: ask again
            P.pix = py * v.w + px;
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
    // batched form: once no path of the current view is in flight, the wave
    // moves to the next view and hands out its pixels
    if constexpr (V::MULTI) {
      if (pending && __ballot(status != LS_NEED && status != LS_IDLE) == 0ull) {
        v.flush(rc, lane);
        total.rays += rc.rays;
        total.bsteps += rc.bsteps;
        total.vsteps += rc.vsteps;
        total.hits += rc.hits;
        rc = RayCount{0u, 0u, 0u, 0u};
        v.switch_view(patch);
        cur = 0u;
        pending = false;
        continue;
      }
    }
    break;
    }
    if (status == LS_NEED && drained) status = LS_IDLE;
    // ---------------- service: camera rays (pathtrace.wgsl:255-262) -------
    if (new_primary) {
      const uint32_t px = P.pix % v.w, py = P.pix / v.w;
      P.rng = init_rng(px, py, v.u().params0[0] * v.spp + P.smp);
      const float jx = rand01(P.rng) - 0.5f;
      const float jy = rand01(P.rng) - 0.5f;
      const float uvx = (((float)px + 0.5f) + jx) / (float)v.w;
      const float uvy = (((float)py + 0.5f) + jy) / (float)v.h;
      ro = mk3(v.u().cam_pos[0], v.u().cam_pos[1], v.u().cam_pos[2]);
      rd = camera_ray(v.u(), uvx, uvy);
      P.pf = PF_PRIMARY;
      begin = true;
    }
    // ---------------- service: ray setup (voxel.wgsl:79-104) -------------
    if (begin) {
      TraceResult unused;
      if (st.begin(v.s(), L0, ro, rd, 1e4f, unused)) {
        status = DDA_MISS;  // misses the grid: serviced next round
        st.steps = 0u;
      } else {
        status = DDA_RUN;
      }
    }
    if (__ballot(status != LS_IDLE) == 0ull) break;
    // ---------------- level switches (vrt_dda_fast.h step_ev) -------------
    // a lane whose ray has just begun, or is back at the coarse level, and
    // whose next brick is occupied enters it in this pass (peek_enter), with
    // the lanes that stopped at one: secondary rays start beside the surface
    // they leave, mostly in an occupied brick
    bool peek = begin && status == DDA_RUN;
    if (status == DDA_LEAVE) {
      status = st.resume(v.s(), L0);
      peek = status == DDA_RUN;
    }
    if (peek) status = st.peek_enter(v.s(), L0);
    if (status == DDA_ENTER)  // a brick whose solid voxels the ray cannot reach is left at once
      status = st.enter(v.s(), L0) ? DDA_RUN : st.resume(v.s(), L0);
    // ---------------- stepping --------------------------------------------
    // until min_wait lanes wait at a brick boundary, shade_min lanes have a
    // finished ray, or no lane steps; the counts are taken every
    // STEPS_PER_CHECK steps (a lane that stops within them waits for the
    // rest: the ballots and compares cost about as much as a step)
    for (;;) {
#pragma unroll
      for (int k = 0; k < STEPS_PER_CHECK; ++k)
        if (status == DDA_RUN) status = st.step_ev(v.s(), L0);
      if (__ballot(status == DDA_RUN) == 0ull) break;
      if ((uint32_t)__popcll(__ballot(status == DDA_ENTER || status == DDA_LEAVE)) >= min_wait) break;
      if ((uint32_t)__popcll(__ballot(status == DDA_HIT || status == DDA_MISS)) >= shade_min) break;
    }
  }
  v.flush(rc, lane);
  count4(counters, total.rays + rc.rays, total.bsteps + rc.bsteps, total.vsteps + rc.vsteps, total.hits + rc.hits);
}

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
  WaveView1 v(s, u, fclamp, spp, radiance, albedo, gbuf, queue);
  wave_paths<GM>(v, L0, min_wait, shade_min, counters);
}

// The batched form: one-wave blocks on the DDA in global memory (views of
// different scenes share no LDS copy), chunks of the table's views from the
// SplitQueue in head->wq; the same per-pixel code as the single-view kernel.
template <int WV>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(WV))) void k_pathtrace_wave_batch(
    BatchHead* __restrict__ head, const ViewDesc* __restrict__ views, unsigned long long* __restrict__ counters,
    uint32_t min_wait, uint32_t shade_min) {
  const LdsScene L0{nullptr, 0u, 0u, 0u, 0u};
  WaveViewBatch v(head, views, head->wq);
  v.load(0u, 0u);
  v.left = 0u;  // the first chunk comes from the queue
  v.np = 0u;    // no view yet: the first draw switches
  wave_paths<true>(v, L0, min_wait, shade_min, counters);
}

}  // namespace vrt
