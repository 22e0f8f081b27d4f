// vrt_trace_wave.h -- the traversal kernel (voxel.wgsl:73-153 over a ray
// list) with step-level refill on persistent waves: the path tracer's wave
// scheme (vrt_pathtrace_wave.h) applied to the ray-list traversal.  Included
// by vrt_kernels.hip after store_hit.
//
// Each lane steps its ray with step_ev (vrt_dda_fast.h) and stops at a brick
// boundary (enter / leave) or at the end of its ray.  When `min_wait` lanes
// wait at a boundary, `fin_min` have finished, or no lane is left stepping,
// the wave runs one pass: the waiting lanes switch level together, and (on
// fin_min finished lanes, or when nothing else can run) every finished lane
// writes its record and takes the next ray of the wave's chunk -- chunks of
// TRACE_CHUNK rays come from a SplitQueue, and their rays are handed to the
// lanes in lane order.  The long branches (ray setup ~190 instructions, brick entry
// ~80, brick exit ~45) so run once for many lanes instead of for the few that
// reach them in a given iteration.  Records are those of trace_fast bit for
// bit: the same DDA operations in the same order per ray.
#pragma once

namespace vrt {

// DDA steps between two counts of the wave's waiting and finished lanes (6
// against 4, 5 and 8 at p = 0.05 / 0.25 / 0.6: -1 / +1 / -3 % against 4).
constexpr int TRACE_STEPS_PER_CHECK = 6;
// Rays a wave takes from the queue at a time (512 measured fastest against
// 128, 256 and 1024 on 2^24-ray lists).
constexpr uint32_t TRACE_CHUNK = 512u;

// NT = 64: one wave per block, the DDA on global memory; NT = 256: four
// waves per block sharing the block's LDS copy of the scene's bitmap.  The
// queue's QWORDS words must be zero at launch.
template <int WV, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WV))) void k_trace_wave(
    SceneDev s, const float* __restrict__ rays, uint32_t n, float tmax, uint4* __restrict__ hits,
    unsigned long long* __restrict__ counters, uint32_t* __restrict__ queue, uint32_t min_wait, uint32_t fin_min) {
  constexpr bool GM = NT == 64;
  extern __shared__ uint32_t lds_dyn[];
  LdsScene L0{nullptr, 0u, 0u, 0u, 0u};
  if constexpr (!GM) L0 = lds_setup<TV_FAST>(s, lds_dyn).f;  // every thread of the block, one barrier
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t lt_mask = lane == 0u ? 0ull : (~0ull >> (64u - lane));
  DdaFastT<GM, true> st;
  uint32_t status = LS_NEED, ray = 0u;
  uint32_t pf_id = RAY_NONE;  // the lane's prefetched ray, RAY_NONE when its slot is empty
  float pf[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 1.0f};
  // the wave's chunk: rays [base, base + clen), the next at base + cur (uniform)
  uint32_t base = 0u, clen = 0u, cur = 0u;
  SplitQueue sq(queue, (uint32_t)(((uint64_t)n + TRACE_CHUNK - 1u) / TRACE_CHUNK), blockIdx.x);
  bool drained = false;
  uint32_t nr = 0u, nb = 0u, nv = 0u, nh = 0u;
  for (;;) {
    // level switches
    if (status == DDA_ENTER) {
      st.enter(s, L0);
      status = DDA_RUN;
    }
    if (status == DDA_LEAVE) status = st.resume(s, L0);
    // finished rays: write their records when fin_min have finished, or
    // when no lane is left stepping
    const uint32_t nfin = (uint32_t)__popcll(__ballot(status == DDA_HIT || status == DDA_MISS));
    const bool service = nfin >= fin_min || __ballot(status == DDA_RUN) == 0ull;
    if (service && (status == DDA_HIT || status == DDA_MISS)) {
      TraceResult o;
      st.finish(s, status, o);
      store_hit(hits, ray, o);
      nb += o.bsteps;
      nv += o.vsteps;
      nh += o.t >= 0.0f ? 1u : 0u;
      nr += 1u;
      status = LS_NEED;
    }
    // next rays.  Each lane keeps the next ray it will trace prefetched in
    // registers (pf, id pf_id): a lane that needs a ray takes its prefetched
    // one at once, and the emptied slots are given the next ids of the
    // wave's chunk (in lane order) and loaded while the rays are walked, so
    // the load latency is hidden behind a whole ray instead of stalling the
    // pass.  Only the first pass waits for its loads.
    bool begin = false;
    f3 ro = mk3(0.0f, 0.0f, 0.0f), rd = mk3(0.0f, 0.0f, 1.0f);
    if (status == LS_NEED && pf_id != RAY_NONE) {
      ray = pf_id;
      ro = mk3(pf[0], pf[1], pf[2]);
      rd = mk3(pf[3], pf[4], pf[5]);
      pf_id = RAY_NONE;
      status = DDA_MISS;  // placeholder until the ray begins
      begin = true;
    }
    uint64_t want = __ballot(pf_id == RAY_NONE);
    bool fetch = false;
    while (want != 0ull && !drained) {
      if (cur >= clen) {
        const uint32_t c = sq.next(lane);
        if (c >= sq.nitems) {
          drained = true;
          break;
        }
        base = c * TRACE_CHUNK;
        clen = n - base < TRACE_CHUNK ? n - base : TRACE_CHUNK;
        cur = 0u;
      }
      const uint32_t avail = clen - cur;
      const uint32_t cnt = (uint32_t)__popcll(want);
      const uint32_t take = cnt < avail ? cnt : avail;
      if (pf_id == RAY_NONE) {
        const uint32_t rank = (uint32_t)__popcll(want & lt_mask);
        if (rank < take) {
          pf_id = base + cur + rank;
          fetch = true;
        }
      }
      cur += take;
      want = __ballot(pf_id == RAY_NONE);
    }
    if (fetch) {
      const float* r = rays + (uint64_t)pf_id * 6;
#pragma unroll
      for (int c = 0; c < 6; ++c) pf[c] = r[c];
    }
    if (status == LS_NEED) {  // no prefetched ray: the first pass, or the queue is drained
      if (pf_id != RAY_NONE) {
        ray = pf_id;
        ro = mk3(pf[0], pf[1], pf[2]);
        rd = mk3(pf[3], pf[4], pf[5]);
        pf_id = RAY_NONE;
        status = DDA_MISS;
        begin = true;
      } else if (drained) {
        status = LS_IDLE;
      }
    }
    if (begin) {
      TraceResult o;
      if (st.begin(s, L0, ro, rd, tmax, o)) {
        status = DDA_MISS;  // misses the grid: its record is written at the next service
        st.steps = 0u;
      } else {
        status = DDA_RUN;
      }
    }
    if (__ballot(status != LS_IDLE) == 0ull) break;
    // stepping until min_wait lanes wait at a boundary, fin_min have
    // finished, or no lane steps
    for (;;) {
#pragma unroll
      for (int k = 0; k < TRACE_STEPS_PER_CHECK; ++k)
        if (status == DDA_RUN) status = st.step_ev(s, L0);
      if (__ballot(status == DDA_RUN) == 0ull) break;
      if ((uint32_t)__popcll(__ballot(status == DDA_ENTER || status == DDA_LEAVE)) >= min_wait) break;
      if ((uint32_t)__popcll(__ballot(status == DDA_HIT || status == DDA_MISS)) >= fin_min) break;
    }
  }
  count4(counters, nr, nb, nv, nh);
}

}  // namespace vrt
