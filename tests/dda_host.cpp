// Host build of the library's device DDA code (voxelrt_amd/csrc), for CPU
// tests: the scene layout (brick words, pool, bitmap, prefix counts) is built
// here from a dense grid the way vrt_scene_upload builds it on the device, and
// each trace variant runs on the CPU over a ray list.  Test infrastructure
// only; compiled by tests/test_dda_host.py.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../voxelrt_amd/csrc/vrt_dda_fast.h"
#include "../voxelrt_amd/csrc/vrt_trace.h"

namespace {

// One lane of the step-level refill path tracer: step_ev() until the ray
// ends, with the level switches it reports taken one at a time.
template <bool PEEK = false, class St>
void step_ev_walk(const vrt::SceneDev& s, const vrt::LdsScene& ls, St& st, vrt::f3 ro, vrt::f3 rd, float tmax,
                  vrt::TraceResult& o, uint64_t& nit) {
  if (st.begin(s, ls, ro, rd, tmax, o)) return;
  // enter() reports a brick whose solid voxels the ray cannot reach (FBOX):
  // the lane then leaves it at once
  auto enter = [&]() { return st.enter(s, ls) ? vrt::DDA_RUN : vrt::DDA_LEAVE; };
  uint32_t r = PEEK ? st.peek_enter(s, ls) : vrt::DDA_RUN;
  if (r == vrt::DDA_ENTER) r = enter();
  for (;;) {
    if (r == vrt::DDA_MISS) break;
    if (r == vrt::DDA_ENTER) {
      r = enter();
    } else if (r == vrt::DDA_LEAVE) {
      r = st.resume(s, ls);
      if (PEEK && r == vrt::DDA_RUN) {
        r = st.peek_enter(s, ls);
        if (r == vrt::DDA_ENTER) r = enter();
      }
    }
    if (r == vrt::DDA_RUN) {
      r = st.step_ev(s, ls);
      ++nit;
    }
    if (r == vrt::DDA_HIT || r == vrt::DDA_MISS) break;
  }
  st.finish(s, r, o);
}

struct HostScene {
  std::vector<uint32_t> bricks, pool, bits, prefix, pbits, pprefix, vbox;
  std::vector<uint8_t> pdist;
  vrt::SceneDev dev{};
};

void build(const uint32_t* vox, int nx, int ny, int nz, HostScene& h) {
  const int bgx = nx / 8, bgy = ny / 8, bgz = nz / 8;
  const uint64_t nb = (uint64_t)bgx * bgy * bgz;
  h.bricks.assign(nb, 0u);
  h.pool.clear();
  uint32_t slot = 0;
  for (int bz = 0; bz < bgz; ++bz)
    for (int by = 0; by < bgy; ++by)
      for (int bx = 0; bx < bgx; ++bx) {
        uint32_t mask[16] = {0};
        uint32_t words[512];
        bool any = false;
        for (int lz = 0; lz < 8; ++lz)
          for (int ly = 0; ly < 8; ++ly)
            for (int lx = 0; lx < 8; ++lx) {
              const uint64_t vi = (uint64_t)(bx * 8 + lx) + (uint64_t)(by * 8 + ly) * nx +
                                  (uint64_t)(bz * 8 + lz) * (uint64_t)nx * ny;
              const uint32_t bit = (uint32_t)(lx + 8 * ly + 64 * lz);
              words[bit] = vox[vi];
              if (vox[vi]) {
                mask[bit >> 5] |= 1u << (bit & 31u);
                any = true;
              }
            }
        const uint64_t bi = (uint64_t)bx + (uint64_t)by * bgx + (uint64_t)bz * bgx * bgy;
        if (!any) continue;
        h.bricks[bi] = ++slot;
        h.pool.insert(h.pool.end(), mask, mask + 16);
        h.pool.insert(h.pool.end(), words, words + 512);
        {  // the brick's box of solid voxels (packed as SceneDev::vbox)
          int lo[3] = {7, 7, 7}, hi[3] = {0, 0, 0};
          for (int lz = 0; lz < 8; ++lz)
            for (int ly = 0; ly < 8; ++ly)
              for (int lx = 0; lx < 8; ++lx)
                if (mask[(lx + 8 * ly + 64 * lz) >> 5] >> ((lx + 8 * ly + 64 * lz) & 31) & 1u) {
                  const int l[3] = {lx, ly, lz};
                  for (int c = 0; c < 3; ++c) {
                    lo[c] = std::min(lo[c], l[c]);
                    hi[c] = std::max(hi[c], l[c]);
                  }
                }
          h.vbox.push_back((uint32_t)(lo[0] | lo[1] << 3 | lo[2] << 6 | hi[0] << 9 | hi[1] << 12 | hi[2] << 15));
        }
      }
  if (h.pool.empty()) h.pool.assign(vrt::POOL_WORDS, 0u);
  if (h.vbox.empty()) h.vbox.assign(1, 0u);
  const uint64_t nw = (nb + 31) / 32;
  h.bits.assign(nw, 0u);
  for (uint64_t i = 0; i < nb; ++i)
    if (h.bricks[i]) h.bits[i >> 5] |= 1u << (i & 31u);
  h.prefix.assign(nw, 0u);
  for (uint64_t w = 1; w < nw; ++w) h.prefix[w] = h.prefix[w - 1] + (uint32_t)__builtin_popcount(h.bits[w - 1]);
  // padded layout: bit index bx | by << sh1 | bz << sh2
  auto lg = [](int v) {
    uint32_t k = 0;
    while ((1 << k) < v) ++k;
    return k;
  };
  const uint32_t sh1 = lg(bgx), sh2 = sh1 + lg(bgy), sh3 = sh2 + lg(bgz);
  const uint64_t pw = ((1ull << sh3) + 31) / 32;
  h.pbits.assign(pw, 0u);
  for (int bz = 0; bz < bgz; ++bz)
    for (int by = 0; by < bgy; ++by)
      for (int bx = 0; bx < bgx; ++bx)
        if (h.bricks[(uint64_t)bx + (uint64_t)by * bgx + (uint64_t)bz * bgx * bgy]) {
          const uint64_t i = (uint64_t)bx | ((uint64_t)by << sh1) | ((uint64_t)bz << sh2);
          h.pbits[i >> 5] |= 1u << (i & 31u);
        }
  h.pprefix.assign(pw, 0u);
  for (uint64_t w = 1; w < pw; ++w) h.pprefix[w] = h.pprefix[w - 1] + (uint32_t)__builtin_popcount(h.pbits[w - 1]);
  // row distances: Chebyshev distance over (by, bz) from each row of bricks to
  // the nearest row holding an occupied brick (brute force), then per padded
  // bitmap word the minimum over the grid rows it covers
  std::vector<int> rocc((size_t)bgy * bgz, 0), rdist((size_t)bgy * bgz, 255);
  for (int bz = 0; bz < bgz; ++bz)
    for (int by = 0; by < bgy; ++by)
      for (int bx = 0; bx < bgx; ++bx)
        if (h.bricks[(uint64_t)bx + (uint64_t)by * bgx + (uint64_t)bz * bgx * bgy]) rocc[by + bz * bgy] = 1;
  for (int bz = 0; bz < bgz; ++bz)
    for (int by = 0; by < bgy; ++by)
      for (int z = 0; z < bgz; ++z)
        for (int y = 0; y < bgy; ++y)
          if (rocc[y + z * bgy]) {
            const int dd = std::max(std::abs(y - by), std::abs(z - bz));
            if (dd < rdist[by + bz * bgy]) rdist[by + bz * bgy] = dd;
          }
  h.pdist.assign((pw + 3) & ~3ull, 255u);
  for (uint64_t w = 0; w < pw; ++w)
    for (uint64_t i = w * 32; i < w * 32 + 32; ++i) {
      const uint64_t y = (i >> sh1) & ((1ull << (sh2 - sh1)) - 1), z = i >> sh2;
      if (y < (uint64_t)bgy && z < (uint64_t)bgz)
        h.pdist[w] = (uint8_t)std::min<int>(h.pdist[w], rdist[y + z * bgy]);
    }
  vrt::SceneDev& s = h.dev;
  s.pbits = h.pbits.data();
  s.pprefix = h.pprefix.data();
  s.pdist = h.pdist.data();
  s.pwords = (uint32_t)pw;
  s.psh1 = sh1;
  s.psh2 = sh2;
  s.bricks = h.bricks.data();
  s.pool = h.pool.data();
  s.vbox = h.vbox.data();
  s.bits = h.bits.data();
  s.prefix = h.prefix.data();
  s.bit_words = (uint32_t)nw;
  s.nx = nx;
  s.ny = ny;
  s.nz = nz;
  s.bgx = bgx;
  s.bgy = bgy;
  s.bgz = bgz;
  s.voxel_size = 0.0625f;
  s.inv_voxel = 1.0f / 0.0625f;
  s.entry_batch = 1u;
  // the box of occupied bricks (brute force over the brick words)
  for (int c = 0; c < 3; ++c) {
    s.box_lo[c] = c == 0 ? bgx : (c == 1 ? bgy : bgz);
    s.box_hi[c] = -1;
  }
  for (int bz = 0; bz < bgz; ++bz)
    for (int by = 0; by < bgy; ++by)
      for (int bx = 0; bx < bgx; ++bx)
        if (h.bricks[(uint64_t)bx + (uint64_t)by * bgx + (uint64_t)bz * bgx * bgy]) {
          const int b[3] = {bx, by, bz};
          for (int c = 0; c < 3; ++c) {
            s.box_lo[c] = std::min(s.box_lo[c], b[c]);
            s.box_hi[c] = std::max(s.box_hi[c], b[c]);
          }
        }
  s.zero = 0.0f;
}

}  // namespace

// variant: 0 nested (HBM), 1 flat, 2 nested with LDS occupancy, 3 single loop
// with both levels in LDS, 4 the same with the empty-row skip, 5 the batched
// form of 4 (trace_fast<true>; one lane here), 6 the single loop on global
// memory (DdaFastT<true>), 7 its batched form, 8 and 9 the path tracer's
// forms of 6 and 3 (brick-entry cell counts from kfine / mfine, KM = true),
// 10 and 11 the step-level refill kernels' stepping (step_ev, level switches one
// at a time), 12 and 13 the same with the entry check after begin and resume,
// 14 and 15 those ending rays at the box of occupied bricks (BOX = true),
// 16 and 17 the path tracer's: also fine walks at the bricks' boxes of solid
// voxels, with no iteration budget (FBOX, NOITER).
// hits: n records of 4 u32
// (t bits, voxel, brick_steps | voxel_steps << 16, face), as vrt_trace_rays
// writes them.  *iters: loop iterations (step() calls) of variants 3 and 4.
extern "C" int dda_host_trace(const uint32_t* vox, int nx, int ny, int nz, const float* rays, uint64_t n,
                              float tmax, int variant, uint32_t* hits, uint64_t* iters) {
  if (nx % 8 || ny % 8 || nz % 8 || nx <= 0 || ny <= 0 || nz <= 0) return -1;
  HostScene h;
  build(vox, nx, ny, nz, h);
  // one block's LDS as the kernels lay it out: 256 mask slots, bitmap, prefix,
  // row distances
  const uint32_t pw = h.dev.pwords;
  std::vector<uint32_t> lds(16 * 256 + 2 * (size_t)pw + (pw + 3) / 4, 0u);
  std::memcpy(lds.data() + 16 * 256, h.pbits.data(), pw * 4);
  std::memcpy(lds.data() + 16 * 256 + pw, h.pprefix.data(), pw * 4);
  std::memcpy(lds.data() + 16 * 256 + 2 * pw, h.pdist.data(), h.pdist.size());
  uint64_t nit = 0;
  const uint32_t lane = 37;  // any slot
  vrt::LdsScene ls{lds.data(), lane, 16u * 256u, 16u * 256u + pw, 16u * 256u + 2u * pw};
  for (uint64_t i = 0; i < n; ++i) {
    const float* r = rays + 6 * i;
    const vrt::f3 ro = vrt::mk3(r[0], r[1], r[2]), rd = vrt::mk3(r[3], r[4], r[5]);
    vrt::TraceResult o;
    switch (variant) {
      case 0: vrt::trace(h.dev, ro, rd, tmax, o); break;
      case 3:
      case 4: {
        vrt::DdaFast st;
        if (st.begin(h.dev, ls, ro, rd, tmax, o)) break;
        uint32_t r;
        for (;;) {
          r = variant == 4 ? st.step<true>(h.dev, ls) : st.step(h.dev, ls);
          ++nit;
          if (r == vrt::DDA_ENTER) {
            st.enter(h.dev, ls);
            continue;
          }
          if (r != vrt::DDA_RUN) break;
        }
        st.finish(h.dev, r, o);
        break;
      }
      case 5: vrt::trace_fast<true, true>(h.dev, ls, ro, rd, tmax, o); break;
      case 6: vrt::trace_fast<false, false, true>(h.dev, ls, ro, rd, tmax, o); break;
      case 7: vrt::trace_fast<true, false, true>(h.dev, ls, ro, rd, tmax, o); break;
      case 8: vrt::trace_fast<false, false, true, true>(h.dev, ls, ro, rd, tmax, o); break;
      case 9: vrt::trace_fast<false, false, false, true>(h.dev, ls, ro, rd, tmax, o); break;
      case 10:  // the step-level refill path tracer's stepping: step_ev + enter / resume
      case 11: {
        if (variant == 10) {
          vrt::DdaFastT<false, true> st;
          step_ev_walk(h.dev, ls, st, ro, rd, tmax, o, nit);
        } else {
          vrt::DdaFastT<true, true> st;
          step_ev_walk(h.dev, ls, st, ro, rd, tmax, o, nit);
        }
        break;
      }
      case 12:  // 10 and 11 with the kernels' re-entry check after resume
      case 13: {
        if (variant == 12) {
          vrt::DdaFastT<false, true> st;
          step_ev_walk<true>(h.dev, ls, st, ro, rd, tmax, o, nit);
        } else {
          vrt::DdaFastT<true, true> st;
          step_ev_walk<true>(h.dev, ls, st, ro, rd, tmax, o, nit);
        }
        break;
      }
      case 16:  // 14 and 15 also ending fine walks at the entered brick's box of solid voxels
      case 17: {
        if (variant == 16) {
          vrt::DdaFastT<false, true, true, true, true> st;
          step_ev_walk<true>(h.dev, ls, st, ro, rd, tmax, o, nit);
        } else {
          vrt::DdaFastT<true, true, true, true, true> st;
          step_ev_walk<true>(h.dev, ls, st, ro, rd, tmax, o, nit);
        }
        break;
      }
      case 14:  // the path tracer's stepping: 12 and 13 ending rays at the occupied box
      case 15: {
        if (variant == 14) {
          vrt::DdaFastT<false, true, true> st;
          step_ev_walk<true>(h.dev, ls, st, ro, rd, tmax, o, nit);
        } else {
          vrt::DdaFastT<true, true, true> st;
          step_ev_walk<true>(h.dev, ls, st, ro, rd, tmax, o, nit);
        }
        break;
      }
      default: return -1;
    }
    uint32_t tb;
    std::memcpy(&tb, &o.t, 4);
    hits[4 * i + 0] = tb;
    hits[4 * i + 1] = o.voxel;
    hits[4 * i + 2] = (o.bsteps & 0xffffu) | (o.vsteps << 16);
    hits[4 * i + 3] = o.face;
  }
  if (iters) *iters = nit;
  return 0;
}

// vrt::exp_w (the a-trous weights' exponential, built from the device
// header) on the float bit patterns start, start + stride, ... (mod 2^32).
extern "C" void exp_w_table(uint64_t start, uint64_t stride, uint64_t n, float* out) {
  for (uint64_t i = 0; i < n; ++i) {
    const uint32_t b = (uint32_t)(start + i * stride);
    float x;
    std::memcpy(&x, &b, 4);
    out[i] = vrt::exp_w(x);
  }
}

