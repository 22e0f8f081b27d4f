// vrt_kernels.hip -- gfx950 kernels of libvoxelrt_hip.
//
//   k_brick_masks   scene upload: 512-bit masks + occupancy per 8^3 brick
//                   (scene.js:411-433), one 64-lane wave per brick, one ballot
//                   per 8x8 layer
//   k_brick_words   occupied-brick slots -> brick words (0 / 1+slot)
//   k_pack_pool     copy each occupied brick's mask + voxels into its slot
//   k_trace_rays    voxel.wgsl:73-153 on a ray list (traversal microbench)
//   k_pathtrace     pathtrace.wgsl:248-330, base branch
//   k_temporal      temporal.wgsl:17-77
//   k_atrous        atrous.wgsl:26-83 (LDS tile with a 2*step halo)
//   k_hdr           main.js:628-659 HDR capture
// No MFMA anywhere: nothing here is a dense contraction.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vrt_device.h"
#include "vrt_kernels.h"
#include "vrt_trace.h"
#include "vrt_dda_fast.h"
#include "vrt_batch.h"

namespace vrt {

// ---------------------------------------------------------------------------
// counters: wave-reduce, one 64-bit atomic per wave
// ---------------------------------------------------------------------------
VRT_D uint32_t wave_sum(uint32_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
VRT_D void count4(unsigned long long* c, uint32_t a, uint32_t b, uint32_t d, uint32_t e) {
  a = wave_sum(a);
  b = wave_sum(b);
  d = wave_sum(d);
  e = wave_sum(e);
  if ((threadIdx.x & 63u) == 0u && c != nullptr) {
    atomicAdd(c + 0, (unsigned long long)a);
    atomicAdd(c + 1, (unsigned long long)b);
    atomicAdd(c + 2, (unsigned long long)d);
    atomicAdd(c + 3, (unsigned long long)e);
  }
}

// ---------------------------------------------------------------------------
// scene build
// ---------------------------------------------------------------------------
// One wave per brick of a slab of whole brick layers (bz in [bz0, bz0+nbz)).
// vox points at voxel (0, 0, bz0*8).  masks/flags are indexed by the global
// brick index.
__global__ __launch_bounds__(256) void k_brick_masks(const uint32_t* __restrict__ vox, int nx,
                                                     int ny, int bgx, int bgy, int bz0, int nbz,
                                                     uint32_t* __restrict__ masks,
                                                     uint32_t* __restrict__ flags) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t wave = (uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const uint64_t nslab = (uint64_t)bgx * bgy * nbz;
  if (wave >= nslab) return;
  const int bx = (int)(wave % bgx);
  const int by = (int)((wave / bgx) % bgy);
  const int bzl = (int)(wave / ((uint64_t)bgx * bgy));
  const int lx = (int)(lane & 7u), ly = (int)(lane >> 3);
  const uint64_t gbi = (uint64_t)bx + (uint64_t)by * bgx + (uint64_t)(bz0 + bzl) * bgx * bgy;
  uint32_t any = 0;
  uint32_t mylo = 0, myhi = 0;
#pragma unroll
  for (int lz = 0; lz < 8; ++lz) {
    const uint64_t vi = (uint64_t)(bx * 8 + lx) + (uint64_t)(by * 8 + ly) * nx +
                        (uint64_t)(bzl * 8 + lz) * (uint64_t)nx * ny;
    const uint64_t b = __ballot(vox[vi] != 0u);
    // words 2*lz (bits 0..31 of the layer) and 2*lz+1 (bits 32..63)
    if (lane == (uint32_t)(2 * lz)) mylo = (uint32_t)b;
    if (lane == (uint32_t)(2 * lz + 1)) myhi = (uint32_t)(b >> 32);
    any |= (b != 0ull) ? 1u : 0u;
  }
  if (lane < 16u) masks[gbi * 16 + lane] = (lane & 1u) ? myhi : mylo;
  if (lane == 0u) flags[gbi] = any;
}

// Edit of the bricks a box of voxels touched (bricks [b0, b0 + ext) per
// axis), one 64-lane wave per brick: the brick's mask from the dense grid as
// k_brick_masks builds it; where its occupancy (empty / not) is what the
// brick word says, an occupied brick's pool entry (mask + voxels) is
// rewritten in place at its slot, as k_pack_pool writes it; where the
// occupancy changed, *changed is set and nothing of that brick is written
// (the caller then rebuilds the scene: slots follow brick order).
__global__ __launch_bounds__(256) void k_edit_bricks(const uint32_t* __restrict__ vox, int nx, int ny, int bgx,
                                                     int bgy, int bx0, int by0, int bz0, int ex, int ey, int ez,
                                                     const uint32_t* __restrict__ bricks, uint32_t* __restrict__ pool,
                                                     uint32_t* __restrict__ changed) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t wave = (uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (wave >= (uint64_t)ex * ey * ez) return;
  const int bx = bx0 + (int)(wave % (uint64_t)ex);
  const int by = by0 + (int)((wave / (uint64_t)ex) % (uint64_t)ey);
  const int bz = bz0 + (int)(wave / ((uint64_t)ex * ey));
  const uint64_t gbi = (uint64_t)bx + (uint64_t)by * bgx + (uint64_t)bz * bgx * bgy;
  const int lx = (int)(lane & 7u), ly = (int)(lane >> 3);
  uint32_t v[8];
  uint32_t any = 0, mylo = 0, myhi = 0;
#pragma unroll
  for (int lz = 0; lz < 8; ++lz) {
    const uint64_t vi = (uint64_t)(bx * 8 + lx) + (uint64_t)(by * 8 + ly) * nx +
                        (uint64_t)(bz * 8 + lz) * (uint64_t)nx * ny;
    v[lz] = vox[vi];
    const uint64_t b = __ballot(v[lz] != 0u);
    if (lane == (uint32_t)(2 * lz)) mylo = (uint32_t)b;
    if (lane == (uint32_t)(2 * lz + 1)) myhi = (uint32_t)(b >> 32);
    any |= (b != 0ull) ? 1u : 0u;
  }
  const uint32_t w = bricks[gbi];
  if ((w != 0u) != (any != 0u)) {
    if (lane == 0u) atomicOr(changed, 1u);
    return;
  }
  if (w == 0u) return;
  uint32_t* dst = pool + (uint64_t)(w - 1u) * POOL_WORDS;
  if (lane < 16u) dst[lane] = (lane & 1u) ? myhi : mylo;
#pragma unroll
  for (int lz = 0; lz < 8; ++lz) dst[16 + lz * 64 + (int)lane] = v[lz];
}

// bricks[i] = flags[i] ? 1 + slot[i] : 0, slot = exclusive prefix of flags
__global__ __launch_bounds__(256) void k_brick_words(const uint32_t* __restrict__ flags,
                                                     const uint32_t* __restrict__ slots,
                                                     uint64_t nb, uint32_t* __restrict__ bricks) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  bricks[i] = flags[i] ? slots[i] + 1u : 0u;
}

// bits: one occupancy bit per brick, 64 bricks per wave (one ballot).  The
// buffer holds whole 64-bit groups, so the tail bits of the last group are 0.
__global__ __launch_bounds__(256) void k_brick_bits(const uint32_t* __restrict__ bricks, uint64_t nb,
                                                    uint32_t* __restrict__ bits) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if ((i & ~63ull) >= nb) return;  // whole wave past the last group
  const uint64_t b = __ballot(i < nb && bricks[i < nb ? i : 0] != 0u);
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t g = i >> 6;  // 64-bit group of this wave
  if (lane == 0u) bits[2 * g] = (uint32_t)b;
  if (lane == 1u) bits[2 * g + 1] = (uint32_t)(b >> 32);
}

// Padded bitmap: bit i of the power-of-two brick grid, i = bx | by << sh1 |
// bz << sh2; bricks outside the real grid read as empty.  64 bits per wave.
__global__ __launch_bounds__(256) void k_brick_pbits(const uint32_t* __restrict__ bricks, int bgx, int bgy, int bgz,
                                                     uint32_t sh1, uint32_t sh2, uint64_t nbits,
                                                     uint32_t* __restrict__ bits) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if ((i & ~63ull) >= nbits) return;  // whole wave past the last group
  const uint32_t x = (uint32_t)(i & ((1ull << sh1) - 1)), y = (uint32_t)((i >> sh1) & ((1ull << (sh2 - sh1)) - 1));
  const uint64_t z = i >> sh2;
  const bool in = i < nbits && x < (uint32_t)bgx && y < (uint32_t)bgy && z < (uint64_t)bgz;
  const uint64_t bi = in ? (uint64_t)x + (uint64_t)y * bgx + z * (uint64_t)bgx * bgy : 0;
  const uint64_t b = __ballot(in && bricks[bi] != 0u);
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t g = i >> 6;
  if (lane == 0u) bits[2 * g] = (uint32_t)b;
  if (lane == 1u && 64 * g + 32 < nbits) bits[2 * g + 1] = (uint32_t)(b >> 32);
}

__global__ __launch_bounds__(256) void k_popc(const uint32_t* __restrict__ words, uint64_t n,
                                              uint32_t* __restrict__ counts) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) counts[i] = (uint32_t)__popc(words[i]);
}

// Row distances (SceneDev::pdist).  A row is the line of bricks with fixed
// (by, bz), index r = by + bz * bgy.  occ[r]: some brick of row r is occupied.
__global__ __launch_bounds__(256) void k_row_occ(const uint32_t* __restrict__ bricks, int bgx, int bgy, int bgz,
                                                 uint8_t* __restrict__ occ) {
  const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= (uint64_t)bgy * (uint64_t)bgz) return;
  const uint32_t* row = bricks + r * (uint64_t)bgx;  // bricks bx + by*bgx + bz*bgx*bgy = bx + r*bgx
  uint32_t o = 0u;
  for (int x = 0; x < bgx; ++x) o |= row[x];
  occ[r] = o != 0u ? 1 : 0;
}

// span[r]: the first and last occupied bx of row r (first | last << 16), or
// 0xFFFF when the row is empty: the host reduces them to the box of occupied
// bricks (SceneDev::box_lo / box_hi).
__global__ __launch_bounds__(256) void k_row_span(const uint32_t* __restrict__ bricks, int bgx, int bgy, int bgz,
                                                  uint32_t* __restrict__ span) {
  const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= (uint64_t)bgy * (uint64_t)bgz) return;
  const uint32_t* row = bricks + r * (uint64_t)bgx;
  uint32_t first = 0xFFFFu, last = 0u;
  for (int x = 0; x < bgx; ++x)
    if (row[x] != 0u) {
      first = first == 0xFFFFu ? (uint32_t)x : first;
      last = (uint32_t)x;
    }
  span[r] = first == 0xFFFFu ? 0xFFFFu : (first | last << 16);
}

// vbox[slot]: the box of the solid voxels of the brick in pool slot `slot`,
// from its 16 mask words (bit x + 8 y + 64 z; word k holds z = k >> 1, rows
// y = 4 (k & 1) .. + 3 in its four bytes): lo x | lo y << 3 | lo z << 6 |
// hi x << 9 | hi y << 12 | hi z << 15.
__global__ __launch_bounds__(256) void k_brick_vbox(const uint32_t* __restrict__ pool, uint64_t nslots,
                                                    uint32_t* __restrict__ vbox) {
  const uint64_t slot = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= nslots) return;
  const uint4* m = (const uint4*)(pool + slot * (uint64_t)POOL_WORDS);
  uint32_t xs = 0u, ys = 0u, zs = 0u;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const uint4 v = m[q];
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int k = 4 * q + j;
      const uint32_t word = w[j];
      xs |= (word | word >> 8 | word >> 16 | word >> 24) & 0xFFu;
#pragma unroll
      for (int b = 0; b < 4; ++b) ys |= ((word >> (8 * b)) & 0xFFu) != 0u ? 1u << (4 * (k & 1) + b) : 0u;
      zs |= word != 0u ? 1u << (k >> 1) : 0u;
    }
  }
  if (xs == 0u) {  // no solid voxel (not an occupied brick's entry)
    vbox[slot] = 0u;
    return;
  }
  const uint32_t lo[3] = {(uint32_t)__builtin_ctz(xs), (uint32_t)__builtin_ctz(ys), (uint32_t)__builtin_ctz(zs)};
  const uint32_t hi[3] = {31u - (uint32_t)__builtin_clz(xs), 31u - (uint32_t)__builtin_clz(ys),
                          31u - (uint32_t)__builtin_clz(zs)};
  vbox[slot] = lo[0] | lo[1] << 3 | lo[2] << 6 | hi[0] << 9 | hi[1] << 12 | hi[2] << 15;
}

// dy[r]: distance along y, within the row's z, to the nearest occupied row
// (255 when there is none or it is farther).
__global__ __launch_bounds__(256) void k_row_dy(const uint8_t* __restrict__ occ, int bgy, int bgz,
                                                uint8_t* __restrict__ dy) {
  const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= (uint64_t)bgy * (uint64_t)bgz) return;
  const int by = (int)(r % (uint64_t)bgy);
  const uint8_t* col = occ + (r - (uint64_t)by);
  int best = 255;
  for (int y = 0; y < bgy; ++y)
    if (col[y]) best = min(best, y > by ? y - by : by - y);
  dy[r] = (uint8_t)best;
}

// dist[r] = min over z' of max(|bz - z'|, dy[by, z']): the Chebyshev
// distance over (y, z) to the nearest occupied row, capped at 255.
__global__ __launch_bounds__(256) void k_row_dist(const uint8_t* __restrict__ dy, int bgy, int bgz,
                                                  uint8_t* __restrict__ dist) {
  const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= (uint64_t)bgy * (uint64_t)bgz) return;
  const int by = (int)(r % (uint64_t)bgy), bz = (int)(r / (uint64_t)bgy);
  int best = 255;
  for (int z = 0; z < bgz; ++z) best = min(best, max(z > bz ? z - bz : bz - z, (int)dy[(uint64_t)by + (uint64_t)z * bgy]));
  dist[r] = (uint8_t)best;
}

// pdist[w] = min of dist over the grid's rows that padded bitmap word w
// covers (255 for a word of padding only); nw is pwords rounded up to 4.
__global__ __launch_bounds__(256) void k_word_dist(const uint8_t* __restrict__ dist, int bgy, int bgz,
                                                   uint32_t sh1, uint32_t sh2, uint64_t pwords, uint64_t nw,
                                                   uint8_t* __restrict__ pdist) {
  const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nw) return;
  int best = 255;
  if (w < pwords) {
    const uint64_t r0 = (w * 32u) >> sh1, r1 = (w * 32u + 31u) >> sh1;
    const uint64_t ym = (1ull << (sh2 - sh1)) - 1ull;
    for (uint64_t r = r0; r <= r1; ++r) {
      const uint64_t y = r & ym, z = r >> (sh2 - sh1);
      if (y < (uint64_t)bgy && z < (uint64_t)bgz) best = min(best, (int)dist[y + z * (uint64_t)bgy]);
    }
  }
  pdist[w] = (uint8_t)best;
}

// One wave per brick of the slab; occupied bricks copy mask + voxels.
__global__ __launch_bounds__(256) void k_pack_pool(const uint32_t* __restrict__ vox, int nx, int ny,
                                                   int bgx, int bgy, int bz0, int nbz,
                                                   const uint32_t* __restrict__ masks,
                                                   const uint32_t* __restrict__ bricks,
                                                   uint32_t* __restrict__ pool) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t wave = (uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const uint64_t nslab = (uint64_t)bgx * bgy * nbz;
  if (wave >= nslab) return;
  const int bx = (int)(wave % bgx);
  const int by = (int)((wave / bgx) % bgy);
  const int bzl = (int)(wave / ((uint64_t)bgx * bgy));
  const uint64_t gbi = (uint64_t)bx + (uint64_t)by * bgx + (uint64_t)(bz0 + bzl) * bgx * bgy;
  const uint32_t w = bricks[gbi];
  if (w == 0u) return;
  uint32_t* dst = pool + (uint64_t)(w - 1u) * POOL_WORDS;
  if (lane < 16u) dst[lane] = masks[gbi * 16 + lane];
  const int lx = (int)(lane & 7u), ly = (int)(lane >> 3);
#pragma unroll
  for (int lz = 0; lz < 8; ++lz) {
    const uint64_t vi = (uint64_t)(bx * 8 + lx) + (uint64_t)(by * 8 + ly) * nx +
                        (uint64_t)(bzl * 8 + lz) * (uint64_t)nx * ny;
    dst[16 + lz * 64 + (int)lane] = vox[vi];
  }
}

// Trace variants of the one-ray-per-lane kernels (all give identical results;
// ids in vrt_kernels.h):
//   TV_NESTED  nested brick/voxel loops reading HBM/L2 (vrt_trace.h)
//   TV_FAST    single-loop DDA with both levels in LDS (vrt_dda_fast.h):
//              bitmap + prefix counts + per-lane mask slot; launched with
//              lds_bytes() of dynamic LDS
//   TV_FASTSKIP TV_FAST with the bulk skip of empty brick rows (traversal)
//   TV_FASTG   the single-loop DDA on global memory (DdaFastT<true>)
// What a lane needs to trace in a given variant.
struct LaneMem {
  LdsScene f;  // TV_FAST
};

template <int TV>
VRT_D void trace_sel(const SceneDev& s, const LaneMem& L, f3 ro, f3 rd, float tmax, TraceResult& h) {
  if constexpr (TV == TV_FASTSKIP)
    trace_fast<false, true>(s, L.f, ro, rd, tmax, h);
  else if constexpr (TV == TV_FASTG)
    trace_fast<false, false, true, true>(s, L.f, ro, rd, tmax, h);
  else if constexpr (tv_fast(TV))
    trace_fast<false, false, false, true>(s, L.f, ro, rd, tmax, h);
  else
    trace(s, ro, rd, tmax, h);
}

// Dynamic LDS of a TV_FAST block of 256 threads: 16 mask words per lane
// (word-major: word w of lane l at w * 256 + l), then the scene's padded
// occupancy bitmap, its prefix counts and its row distances.  Every thread of
// the block must call this.
template <int TV>
VRT_D LaneMem lds_setup(const SceneDev& s, uint32_t* lds) {
  LaneMem L{{nullptr, 0u, 0u, 0u, 0u}};
  if constexpr (tv_fast(TV)) {
    uint32_t* bits = lds + 16u * 256u;
    for (uint32_t i = threadIdx.x; i < s.pwords; i += blockDim.x) {
      bits[i] = s.pbits[i];
      bits[s.pwords + i] = s.pprefix[i];
    }
    const uint32_t dw = (s.pwords + 3u) / 4u;
    for (uint32_t i = threadIdx.x; i < dw; i += blockDim.x) bits[2u * s.pwords + i] = ((const uint32_t*)s.pdist)[i];
    __syncthreads();
    L.f.lds = lds;
    L.f.lane = threadIdx.x;
    L.f.bits_off = 16u * 256u;
    L.f.prefix_off = 16u * 256u + s.pwords;
    L.f.dist_off = 16u * 256u + 2u * s.pwords;
  }
  return L;
}

// ---------------------------------------------------------------------------
// traversal microbench: one ray per lane, grid-stride
// ---------------------------------------------------------------------------
// WV: minimum waves per SIMD the register allocation must allow (1 = the
// compiler's choice; 6 trades a few registers for more waves in flight,
// which pays on sparse scenes, where rays take long empty walks).
template <int TV, int WV = 1>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WV))) void k_trace_rays(SceneDev s, const float* __restrict__ rays,
                                                    uint64_t n, float tmax,
                                                    uint4* __restrict__ hits,
                                                    unsigned long long* __restrict__ counters) {
  extern __shared__ uint32_t lds_dyn[];
  const LaneMem L = lds_setup<TV>(s, lds_dyn);
  uint32_t nb = 0, nv = 0, nh = 0, nr = 0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  // every lane of a wave runs the same number of outer iterations so the
  // counter reduction below sees all 64 lanes
  const uint64_t base0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (uint64_t i = base0; i < n; i += stride) {
    const float* r = rays + i * 6;
    TraceResult h;
    if constexpr (tv_fast(TV) || TV == TV_FASTG)  // wave-batched brick entries (SceneDev::entry_batch)
      trace_fast<true, TV == TV_FASTSKIP, TV == TV_FASTG>(s, L.f, mk3(r[0], r[1], r[2]), mk3(r[3], r[4], r[5]),
                                                           tmax, h);
    else
      trace_sel<TV>(s, L, mk3(r[0], r[1], r[2]), mk3(r[3], r[4], r[5]), tmax, h);
    uint4 rec;
    rec.x = __builtin_bit_cast(uint32_t, h.t);
    rec.y = h.voxel;
    rec.z = (h.bsteps & 0xffffu) | (h.vsteps << 16);
    rec.w = h.face;
    hits[i] = rec;
    nb += h.bsteps;
    nv += h.vsteps;
    nh += h.t >= 0.0f ? 1u : 0u;
    nr += 1u;
  }
  count4(counters, nr, nb, nv, nh);
}

// ---------------------------------------------------------------------------
// traversal microbench, persistent form: each lane pulls rays from a global
// queue (one atomic per wave refill) and advances a resumable DDA state one
// loop iteration at a time (vrt_dda_fast.h).  `steps` iterations run between
// refill checks; a refill happens once at least `min_refill` lanes are idle
// (or the whole wave is).  Rays are taken from the global counter in chunks of
// QUEUE_CHUNK per wave, so the counter sees one atomic per 256 rays instead of
// one per refill.  The queue counter must be zero at launch.
// ---------------------------------------------------------------------------
constexpr uint32_t RAY_NONE = 0xFFFFFFFFu;  // lane wants a ray
constexpr uint32_t RAY_DONE = 0xFFFFFFFEu;  // queue exhausted for this lane
constexpr uint32_t QUEUE_CHUNK = 256u;      // rays a wave takes per queue atomic

// A work queue spread over QSPLIT counters QSTRIDE words apart (one 128-byte
// line each): item i is drawn from counter i % QSPLIT as the k-th draw there
// (i = k * QSPLIT + counter).  Atomics on one address are serialised in the
// memory system (about 12 ns each on MI355X), which capped a single-counter
// queue near 80 M draws/s; each wave draws from its home counter (its block
// index mod QSPLIT) and, when that is exhausted, from each of the others
// once in turn (a draw past the end tells the wave that counter is done).
// Items still come out in about increasing order, so a queue ordered
// costly-first stays so.  The QSPLIT counters start at zero; each counts at
// most nitems / QSPLIT + 1 + (waves of the launch) draws.
constexpr uint32_t QSPLIT = 8u, QSTRIDE = 32u, QWORDS = QSPLIT * QSTRIDE;
static_assert(QWORDS == QUEUE_WORDS, "queue size");
struct SplitQueue {
  uint32_t* q;
  uint32_t nitems, home, tried;
  VRT_D SplitQueue(uint32_t* queue, uint32_t n, uint32_t block) : q(queue), nitems(n), home(block % QSPLIT), tried(0u) {}
  // The next item for the calling wave (the same value in every lane), or
  // nitems when every counter is exhausted.  Called by the whole wave.
  VRT_D uint32_t next(uint32_t lane) {
    while (tried < QSPLIT) {
      const uint32_t c = (home + tried) % QSPLIT;
      uint32_t k = 0u;
      if (lane == 0u) k = atomicAdd(q + c * QSTRIDE, 1u);
      k = __builtin_amdgcn_readfirstlane(__shfl(k, 0, 64));
      const uint64_t item = (uint64_t)k * QSPLIT + c;
      if (item < nitems) return (uint32_t)item;
      ++tried;
    }
    return nitems;
  }
};

VRT_D void store_hit(uint4* hits, uint32_t i, const TraceResult& h) {
  uint4 rec;
  rec.x = __builtin_bit_cast(uint32_t, h.t);
  rec.y = h.voxel;
  rec.z = (h.bsteps & 0xffffu) | (h.vsteps << 16);
  rec.w = h.face;
  hits[i] = rec;
}

// The resumable DDA a persistent lane runs (the single-loop LDS form).
// A lane that reaches an occupied brick waits (pend) until entry_batch lanes
// of its wave wait, or every lane walking a ray does; settle(), called by the
// whole wave after each step, lets them enter together.
template <bool Skip>
struct FastRunner {
  DdaFast st;
  bool pend = false;
  VRT_D bool begin(const SceneDev& s, const LaneMem& L, f3 ro, f3 rd, float tmax, TraceResult& o) {
    return st.begin(s, L.f, ro, rd, tmax, o);
  }
  VRT_D bool step(const SceneDev& s, const LaneMem& L, TraceResult& o) {
    if (pend) return false;
    const uint32_t r = st.step<Skip>(s, L.f);
    if (r == DDA_RUN) return false;
    if (r == DDA_ENTER) {
      pend = true;
      return false;
    }
    st.finish(s, r, o);
    return true;
  }
  VRT_D void settle(const SceneDev& s, const LaneMem& L, bool walking) {
    const uint64_t w = wave_ballot(pend);
    if (w == 0ull) return;
    const uint32_t nw = (uint32_t)__popcll(w), nr = (uint32_t)__popcll(wave_ballot(walking));
    if (nw >= (nr < s.entry_batch ? nr : s.entry_batch) && pend) {
      st.enter(s, L.f);
      pend = false;
    }
  }
  VRT_D f3 dir() const { return mk3(st.rdv[0], st.rdv[1], st.rdv[2]); }
};

template <class Runner, int TV, int WV = 1>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WV))) void k_trace_persistent(SceneDev s, const float* __restrict__ rays,
                                                          uint32_t n, float tmax,
                                                          uint4* __restrict__ hits,
                                                          unsigned long long* __restrict__ counters,
                                                          uint32_t* __restrict__ queue, int steps,
                                                          int min_refill) {
  extern __shared__ uint32_t lds_dyn[];
  const LaneMem L = lds_setup<TV>(s, lds_dyn);
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t lt_mask = lane == 0u ? 0ull : (~0ull >> (64u - lane));
  uint32_t ray = RAY_NONE;
  Runner st;
  uint32_t nb = 0, nv = 0, nh = 0, nr = 0;
  // the wave's current chunk of the queue: [cbase, cbase + cleft)
  uint32_t cbase = 0u, cleft = 0u;
  for (;;) {
    const uint64_t need = __ballot(ray == RAY_NONE);
    const uint64_t busy = __ballot(ray < n);
    if (need != 0ull && (busy == 0ull || __popcll(need) >= min_refill)) {
      if (cleft == 0u) {
        uint32_t b = 0u;
        if (lane == 0u) b = atomicAdd(queue, QUEUE_CHUNK);
        cbase = __shfl(b, 0, 64);
        cleft = QUEUE_CHUNK;
      }
      const uint32_t cnt = (uint32_t)__popcll(need);
      const uint32_t take = cnt < cleft ? cnt : cleft;
      const uint32_t rank = (uint32_t)__popcll(need & lt_mask);
      const uint32_t base = cbase;
      cbase += take;
      cleft -= take;
      if (ray == RAY_NONE && rank < take) {
        const uint32_t my = base + rank;
        if (my < n) {
          const float* r = rays + (uint64_t)my * 6;
          TraceResult o;
          if (st.begin(s, L, mk3(r[0], r[1], r[2]), mk3(r[3], r[4], r[5]), tmax, o)) {
            store_hit(hits, my, o);
            nr += 1u;
          } else {
            ray = my;
          }
        } else {
          ray = RAY_DONE;
        }
      }
    }
    if (__ballot(ray != RAY_DONE) == 0ull) break;
    for (int k = 0; k < steps; ++k) {
      if (ray < n) {
        TraceResult o;
        if (st.step(s, L, o)) {
          store_hit(hits, ray, o);
          nb += o.bsteps;
          nv += o.vsteps;
          nh += o.t >= 0.0f ? 1u : 0u;
          nr += 1u;
          ray = RAY_NONE;
        }
      }
      st.settle(s, L, ray < n);
    }
  }
  count4(counters, nr, nb, nv, nh);
}

// ---------------------------------------------------------------------------
// camera / sky / sampling, common.wgsl:90-139
// ---------------------------------------------------------------------------
VRT_D float4 mat_vec4(const float* m, float x, float y, float z, float w) {
  float4 o;
  o.x = ((m[0] * x + m[4] * y) + m[8] * z) + m[12] * w;
  o.y = ((m[1] * x + m[5] * y) + m[9] * z) + m[13] * w;
  o.z = ((m[2] * x + m[6] * y) + m[10] * z) + m[14] * w;
  o.w = ((m[3] * x + m[7] * y) + m[11] * z) + m[15] * w;
  return o;
}

VRT_D f3 camera_ray(const Uniforms& u, float uvx, float uvy) {
  const float ndcx = uvx * 2.0f - 1.0f;
  const float ndcy = 1.0f - uvy * 2.0f;
  const float4 far = mat_vec4(u.inv_view_proj, ndcx, ndcy, 1.0f, 1.0f);
  const f3 p = mk3(far.x / far.w, far.y / far.w, far.z / far.w);
  return normalize(p - mk3(u.cam_pos[0], u.cam_pos[1], u.cam_pos[2]));
}

VRT_D f3 sky_color(const Uniforms& u, f3 rd, bool disc) {
  const float up = vclamp(rd.y, -1.0f, 1.0f);
  const float tt = pow_r(vmax(up, 0.0f), 0.55f);
  f3 sky = mix(mk3(0.55f, 0.68f, 0.85f), mk3(0.15f, 0.32f, 0.62f), tt);
  if (up < 0.0f) sky = mix(sky, mk3(0.28f, 0.26f, 0.24f), vmin(1.0f, -up * 4.0f));
  const f3 sd = mk3(u.sun_dir[0], u.sun_dir[1], u.sun_dir[2]);
  const float sunAmt = vmax(dot(rd, sd), 0.0f);
  sky = sky + (mk3(1.0f, 0.75f, 0.5f) * pow8(sunAmt)) * 0.2f;
  if (disc && sunAmt > u.sun_dir[3]) {
    const float solidAngle = TWO_PI_F * (1.0f - u.sun_dir[3]);
    sky = sky + mk3(u.sun_radiance[0], u.sun_radiance[1], u.sun_radiance[2]) / vmax(solidAngle, 1e-4f);
  }
  return sky * u.sun_radiance[3];
}

VRT_D f3 sample_sun_dir(const Uniforms& u, uint32_t& rng) {
  const float r0 = rand01(rng);
  const float r1 = rand01(rng);
  const float cosT = mix(1.0f, u.sun_dir[3], r0);
  const float sinT = vsqrt(vmax(0.0f, 1.0f - cosT * cosT));
  const float phi = TWO_PI_F * r1;
  const f3 w = mk3(u.sun_dir[0], u.sun_dir[1], u.sun_dir[2]);
  const f3 a = vabs(w.x) > 0.9f ? mk3(0.0f, 1.0f, 0.0f) : mk3(1.0f, 0.0f, 0.0f);
  const f3 tang = normalize(cross(a, w));
  const f3 bitan = cross(w, tang);
  float s, c;
  sincos_r(phi, &s, &c);
  return normalize((tang * (c * sinT) + bitan * (s * sinT)) + w * cosT);
}

VRT_D f3 cosine_hemisphere(f3 n, uint32_t& rng) {
  const float r0 = rand01(rng);
  const float r1 = rand01(rng);
  const float phi = TWO_PI_F * r0;
  const float sr = vsqrt(r1);
  const f3 a = vabs(n.x) > 0.9f ? mk3(0.0f, 1.0f, 0.0f) : mk3(1.0f, 0.0f, 0.0f);
  const f3 tang = normalize(cross(a, n));
  const f3 bitan = cross(n, tang);
  float s, c;
  sincos_r(phi, &s, &c);
  return normalize((tang * (c * sr) + bitan * (s * sr)) + n * vsqrt(vmax(0.0f, 1.0f - r1)));
}

// ---------------------------------------------------------------------------
// path tracer, one pixel per lane
// ---------------------------------------------------------------------------
struct PixelOut {
  f3 radiance;
};

// Where the first sample of a pixel stores its albedo (rgba8unorm of
// sqrt(albedo)) and G-buffer texel, written as soon as the primary ray is
// resolved so that neither stays live through the bounce loop.
struct GbufOut {
  uchar4* albedo;
  float4* gbuf;
  uint64_t i;
  bool write;
  VRT_D void store(f3 alb_sqrt, float4 g) const {
    if (!write) return;
    albedo[i] = make_uchar4((uint8_t)f32_to_unorm8(alb_sqrt.x), (uint8_t)f32_to_unorm8(alb_sqrt.y),
                            (uint8_t)f32_to_unorm8(alb_sqrt.z), 255);
    gbuf[i] = g;
  }
};

struct RayCount {
  uint32_t rays, bsteps, vsteps, hits;
  VRT_D void add(const TraceResult& h) {
    rays += 1u;
    bsteps += h.bsteps;
    vsteps += h.vsteps;
    hits += h.t >= 0.0f ? 1u : 0u;
  }
};

template <int TV>
VRT_D PixelOut shade_pixel(const SceneDev& s, const LaneMem& L, const Uniforms& u, float fclamp, uint32_t px,
                           uint32_t py, uint32_t frame, uint32_t w, uint32_t h, RayCount& rc, const GbufOut& go) {
  PixelOut res;
  uint32_t rng = init_rng(px, py, frame);
  const float jx = rand01(rng) - 0.5f;
  const float jy = rand01(rng) - 0.5f;
  const float uvx = (((float)px + 0.5f) + jx) / (float)w;
  const float uvy = (((float)py + 0.5f) + jy) / (float)h;
  const f3 ro = mk3(u.cam_pos[0], u.cam_pos[1], u.cam_pos[2]);
  f3 rd = camera_ray(u, uvx, uvy);

  TraceResult h0;
  trace_sel<TV>(s, L, ro, rd, 1e4f, h0);
  rc.add(h0);
  const bool indirect = (u.params0[2] & VRT_FLAG_INDIRECT) != 0u;
  if (h0.t < 0.0f) {
    res.radiance = indirect ? mk3(0.0f, 0.0f, 0.0f) : sky_color(u, rd, true);
    go.store(mk3(1.0f, 1.0f, 1.0f), make_float4(-rd.x, -rd.y, -rd.z, -1.0f));
    return res;
  }
  const Material mat1 = unpack_material(h0.mat);
  const f3 x1 = ro + rd * h0.t;
  const f3 n1 = h0.n;
  go.store(mk3(vsqrt(mat1.albedo.x), vsqrt(mat1.albedo.y), vsqrt(mat1.albedo.z)),
           make_float4(n1.x, n1.y, n1.z, h0.t));

  f3 radiance = mk3(0.0f, 0.0f, 0.0f);
  f3 tp = mk3(1.0f, 1.0f, 1.0f);
  f3 pos = x1, n = n1;
  Material mat = mat1;
  const uint32_t nb = u.params0[1];
  const f3 sunRad = mk3(u.sun_radiance[0], u.sun_radiance[1], u.sun_radiance[2]);
  for (uint32_t bounce = 0; bounce <= nb; ++bounce) {
    const bool direct = !(indirect && bounce == 0u);
    if (mat.emissive > 0.0f && direct) radiance = radiance + (tp * mat.albedo) * (mat.emissive * EMISSIVE_SCALE);
    const f3 surf = pos + n * 1e-3f;
    const f3 l = sample_sun_dir(u, rng);
    const float ndl = dot(n, l);
    if (ndl > 0.0f && l.y > 0.0f && direct) {
      TraceResult hs;
      trace_sel<TV>(s, L, surf, l, 1e4f, hs);
      rc.add(hs);
      if (!(hs.t >= 0.0f)) radiance = radiance + (((tp * mat.albedo) * INV_PI_F) * ndl) * sunRad;
    }
    if (bounce == nb) break;
    rd = cosine_hemisphere(n, rng);
    tp = tp * mat.albedo;
    if (bounce >= 2u) {
      const float p = vclamp(vmax(tp.x, vmax(tp.y, tp.z)), 0.05f, 0.95f);
      if (rand01(rng) > p) break;
      tp = tp / p;
    }
    TraceResult hb;
    trace_sel<TV>(s, L, surf, rd, 1e4f, hb);
    rc.add(hb);
    if (hb.t < 0.0f) {
      radiance = radiance + tp * sky_color(u, rd, false);
      break;
    }
    pos = surf + rd * hb.t;
    n = hb.n;
    mat = unpack_material(hb.mat);
  }
  if (!(fclamp <= 0.0f))
    radiance = mk3(vmin(radiance.x, fclamp), vmin(radiance.y, fclamp), vmin(radiance.z, fclamp));
  const f3 den = mk3(vmax(mat1.albedo.x, 1e-3f), vmax(mat1.albedo.y, 1e-3f), vmax(mat1.albedo.z, 1e-3f));
  res.radiance = radiance / den;
  return res;
}

// pathtrace.wgsl:248-330 as a resumable per-lane sample path: start() sets
// up a sample's primary ray, after() consumes the result of the ray just
// traced and either sets up the next ray (shadow or bounce) or ends the
// sample.  The operations and RNG draws are those of shade_pixel in the same
// order, so a lane that runs start(), then after() once per traced ray,
// produces the same bits.
// Register reuse (the path's values are those of shade_pixel): ro is the
// origin of the next ray -- the camera, then the offset surface point of the
// last vertex, where both its shadow and its bounce ray start -- and rd its
// direction: the sun sample during SHADOW (the previous direction is dead
// once the vertex is known, and the bounce direction is drawn after the
// shadow ray), else the primary or bounce direction.
struct SamplePath {
  enum : uint32_t { PRIMARY, SHADOW, BOUNCE };
  uint32_t rng, bounce, phase;
  f3 ro, rd, radiance, tp, n, alb1;
  Material mat;
  float ndl;
  VRT_D f3 dir() const { return rd; }
  VRT_D void start(const Uniforms& u, uint32_t px, uint32_t py, uint32_t frame, uint32_t w, uint32_t h) {
    rng = init_rng(px, py, frame);
    const float jx = rand01(rng) - 0.5f;
    const float jy = rand01(rng) - 0.5f;
    const float uvx = (((float)px + 0.5f) + jx) / (float)w;
    const float uvy = (((float)py + 0.5f) + jy) / (float)h;
    ro = mk3(u.cam_pos[0], u.cam_pos[1], u.cam_pos[2]);
    rd = camera_ray(u, uvx, uvy);
    radiance = mk3(0.0f, 0.0f, 0.0f);
    tp = mk3(1.0f, 1.0f, 1.0f);
    bounce = 0u;
    phase = PRIMARY;
  }
  // True when the sample is finished; *out is then its (clamped,
  // demodulated) radiance, or the sky for a primary miss.
  VRT_D bool after(const Uniforms& u, float fclamp, const TraceResult& hh, const GbufOut& go, f3* out) {
    const bool indirect = (u.params0[2] & VRT_FLAG_INDIRECT) != 0u;
    const uint32_t nb = u.params0[1];
    bool vertex = false, after_nee = false, done = false;
    f3 pos = ro;
    if (phase == PRIMARY) {
      if (hh.t < 0.0f) {
        *out = indirect ? mk3(0.0f, 0.0f, 0.0f) : sky_color(u, rd, true);
        go.store(mk3(1.0f, 1.0f, 1.0f), make_float4(-rd.x, -rd.y, -rd.z, -1.0f));
        return true;
      }
      mat = unpack_material(hh.mat);
      alb1 = mat.albedo;
      pos = ro + rd * hh.t;
      n = hh.n;
      go.store(mk3(vsqrt(alb1.x), vsqrt(alb1.y), vsqrt(alb1.z)), make_float4(n.x, n.y, n.z, hh.t));
      vertex = true;
    } else if (phase == SHADOW) {
      const f3 sunRad = mk3(u.sun_radiance[0], u.sun_radiance[1], u.sun_radiance[2]);
      if (!(hh.t >= 0.0f)) radiance = radiance + (((tp * mat.albedo) * INV_PI_F) * ndl) * sunRad;
      after_nee = true;
    } else {  // BOUNCE
      if (hh.t < 0.0f) {
        radiance = radiance + tp * sky_color(u, rd, false);
        done = true;
      } else {
        pos = ro + rd * hh.t;
        n = hh.n;
        mat = unpack_material(hh.mat);
        ++bounce;
        vertex = true;
      }
    }
    if (vertex) {  // emission + sun NEE at the path vertex
      const bool direct = !(indirect && bounce == 0u);
      if (mat.emissive > 0.0f && direct) radiance = radiance + (tp * mat.albedo) * (mat.emissive * EMISSIVE_SCALE);
      ro = pos + n * 1e-3f;
      rd = sample_sun_dir(u, rng);
      ndl = dot(n, rd);
      if (ndl > 0.0f && rd.y > 0.0f && direct) {
        phase = SHADOW;
      } else {
        after_nee = true;
      }
    }
    if (after_nee) {  // cosine bounce + Russian roulette
      if (bounce == nb) {
        done = true;
      } else {
        rd = cosine_hemisphere(n, rng);
        tp = tp * mat.albedo;
        bool alive = true;
        if (bounce >= 2u) {
          const float p = vclamp(vmax(tp.x, vmax(tp.y, tp.z)), 0.05f, 0.95f);
          if (rand01(rng) > p) alive = false;
          else tp = tp / p;
        }
        if (alive) {
          phase = BOUNCE;
        } else {
          done = true;
        }
      }
    }
    if (!done) return false;
    f3 r = radiance;
    if (!(fclamp <= 0.0f)) r = mk3(vmin(r.x, fclamp), vmin(r.y, fclamp), vmin(r.z, fclamp));
    const f3 den = mk3(vmax(alb1.x, 1e-3f), vmax(alb1.y, 1e-3f), vmax(alb1.z, 1e-3f));
    *out = r / den;
    return true;
  }
};

// Pixel p of a launch covering TILE x TILE tiles, row-major tiles: keeps the
// 64 lanes of a wave on a compact 16x4 patch (coherent primary rays).
// T = 8: 8x8 tiles of one wave each.
template <int T = 16>
VRT_D bool tile_pixel(uint32_t w, uint32_t h, uint32_t tile, uint32_t& px, uint32_t& py) {
  const uint32_t tiles_x = (w + (uint32_t)T - 1u) / (uint32_t)T;
  px = (tile % tiles_x) * (uint32_t)T + (threadIdx.x % (uint32_t)T);
  py = (tile / tiles_x) * (uint32_t)T + (threadIdx.x / (uint32_t)T);
  return px < w && py < h;
}

// One T x T tile of pathtrace.wgsl:248-330 (one pixel per lane).
template <int TV, int T = 16>
VRT_D void pathtrace_tile(const SceneDev& s, const LaneMem& L, const Uniforms& u, float fclamp, uint32_t spp, uint32_t tile,
                          ushort4* __restrict__ radiance, uchar4* __restrict__ albedo,
                          float4* __restrict__ gbuf, RayCount& rc) {
  const uint32_t w = (uint32_t)u.params1[0], h = (uint32_t)u.params1[1];
  uint32_t px, py;
  if (tile_pixel<T>(w, h, tile, px, py)) {
    const uint64_t i = (uint64_t)py * w + px;
    f3 acc = mk3(0.0f, 0.0f, 0.0f);
    for (uint32_t smp = 0; smp < spp; ++smp) {
      const GbufOut go{albedo, gbuf, i, smp == 0u};  // the first sample's albedo / G-buffer
      PixelOut r;
      r = shade_pixel<TV>(s, L, u, fclamp, px, py, u.params0[0] * spp + smp, w, h, rc, go);
      acc = acc + r.radiance;
    }
    const f3 o = acc / (float)spp;
    radiance[i] = make_ushort4(f32_to_f16(o.x), f32_to_f16(o.y), f32_to_f16(o.z), f32_to_f16(1.0f));
  }
}

// Tile of block b of an n-block launch in the given order (SceneDev::tile_order).
VRT_D uint32_t map_tile(uint32_t b, uint32_t n, uint32_t order) {
  if (order & 2u) b = n - 1u - b;
  if (order & 1u) {
    const uint32_t q = n / 8u, r = n % 8u, x = b % 8u, j = b / 8u;
    b = (x < r ? x * (q + 1u) : r * (q + 1u) + (x - r) * q) + j;
  }
  return b;
}

// WV: minimum waves per SIMD the register allocation must allow (1 = no
// constraint).  The path tracer's natural allocation (~150-165 VGPRs) leaves
// 3 waves per SIMD; 4 or 5 trade a few spills for more waves in flight.
template <int TV, int WV>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WV))) void k_pathtrace(SceneDev s, Uniforms u, float fclamp,
                                                   uint32_t spp, ushort4* __restrict__ radiance,
                                                   uchar4* __restrict__ albedo,
                                                   float4* __restrict__ gbuf,
                                                   unsigned long long* __restrict__ counters) {
  extern __shared__ uint32_t lds_dyn[];
  const LaneMem L = lds_setup<TV>(s, lds_dyn);
  RayCount rc{0u, 0u, 0u, 0u};
  pathtrace_tile<TV>(s, L, u, fclamp, spp, map_tile(blockIdx.x, gridDim.x, s.tile_order), radiance, albedo, gbuf,
                     rc);
  count4(counters, rc.rays, rc.bsteps, rc.vsteps, rc.hits);
}

// The global-memory variant (no LDS to share) on 8x8 tiles of one wave per
// block: four times the blocks of finer cost, so the launch's tail is
// shorter, and square 8x8 patches of coherent primary rays.
template <int WV>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(WV))) void k_pathtrace8(
    SceneDev s, Uniforms u, float fclamp, uint32_t spp, ushort4* __restrict__ radiance, uchar4* __restrict__ albedo,
    float4* __restrict__ gbuf, unsigned long long* __restrict__ counters) {
  const LaneMem L{{nullptr, 0u, 0u, 0u, 0u}};
  RayCount rc{0u, 0u, 0u, 0u};
  pathtrace_tile<TV_FASTG, 8>(s, L, u, fclamp, spp, map_tile(blockIdx.x, gridDim.x, s.tile_order), radiance, albedo,
                              gbuf, rc);
  count4(counters, rc.rays, rc.bsteps, rc.vsteps, rc.hits);
}

// One block (NT threads) renders the TW x TH tile at (x0, y0), its lanes
// taking pixels from the tile's counter (a wave takes as many consecutive
// pixels as it has idle lanes, in 8x8 patches): each loop round traces one
// ray per busy lane -- primary, shadow or bounce, one call site -- and a lane
// whose pixel is finished takes the next one, so lanes do not idle while
// other lanes of the wave finish longer paths.  The counter is a wave-uniform
// register for one-wave blocks, else *s_next in LDS (set to 0 and published
// by the caller).
template <int TV, int TW, int TH, int NT>
VRT_D void refill_tile(const SceneDev& s, const LaneMem& L, const Uniforms& u, float fclamp, uint32_t spp,
                       uint32_t x0, uint32_t y0, ushort4* __restrict__ radiance, uchar4* __restrict__ albedo,
                       float4* __restrict__ gbuf, RayCount& rc, uint32_t* s_next) {
  static_assert(TW % 8 == 0 && TH % 8 == 0, "tiles of 8x8 patches");
  constexpr uint32_t NP = (uint32_t)(TW * TH), PW = (uint32_t)(TW / 8);
  const uint32_t w = (uint32_t)u.params1[0], h = (uint32_t)u.params1[1];
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t lt_mask = lane == 0u ? 0ull : (~0ull >> (64u - lane));
  SamplePath P;
  bool have = false, exhausted = false;
  uint32_t px = 0u, py = 0u, smp = 0u, next = 0u;
  f3 acc = mk3(0.0f, 0.0f, 0.0f);
  for (;;) {
    const uint64_t need = __ballot(!have && !exhausted);
    if (need != 0ull) {
      uint32_t base;
      if constexpr (NT == 64) {
        base = next;
        next += (uint32_t)__popcll(need);
      } else {
        const uint32_t first = (uint32_t)__builtin_ctzll(need);
        base = 0u;
        if (lane == first) base = atomicAdd(s_next, (uint32_t)__popcll(need));
        base = __shfl(base, (int)first, 64);
      }
      if (!have && !exhausted) {
        const uint32_t p = base + (uint32_t)__popcll(need & lt_mask);
        if (p >= NP) {
          exhausted = true;
        } else {
          const uint32_t patch = p >> 6, q = p & 63u;
          px = x0 + (patch % PW) * 8u + (q & 7u);
          py = y0 + (patch / PW) * 8u + (q >> 3);
          if (px < w && py < h) {
            have = true;
            smp = 0u;
            acc = mk3(0.0f, 0.0f, 0.0f);
            P.start(u, px, py, u.params0[0] * spp, w, h);
          }
        }
      }
    }
    if (exhausted) break;
    if (!have) continue;  // a pixel outside the image: take another
    TraceResult hh;
    trace_sel<TV>(s, L, P.ro, P.dir(), 1e4f, hh);
    rc.add(hh);
    const uint64_t i = (uint64_t)py * w + px;
    f3 val;
    if (P.after(u, fclamp, hh, GbufOut{albedo, gbuf, i, smp == 0u}, &val)) {
      acc = acc + val;
      ++smp;
      if (smp < spp) {
        P.start(u, px, py, u.params0[0] * spp + smp, w, h);
      } else {
        const f3 o = acc / (float)spp;
        radiance[i] = make_ushort4(f32_to_f16(o.x), f32_to_f16(o.y), f32_to_f16(o.z), f32_to_f16(1.0f));
        have = false;
      }
    }
  }
}

template <int TV, int WV, int TW, int TH, int NT = 256>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WV))) void k_pathtrace_refill(
    SceneDev s, Uniforms u, float fclamp, uint32_t spp, ushort4* __restrict__ radiance, uchar4* __restrict__ albedo,
    float4* __restrict__ gbuf, unsigned long long* __restrict__ counters) {
  extern __shared__ uint32_t lds_dyn[];
  __shared__ uint32_t s_next;
  if (threadIdx.x == 0u) s_next = 0u;
  const LaneMem L = lds_setup<TV>(s, lds_dyn);
  __syncthreads();
  const uint32_t w = (uint32_t)u.params1[0];
  const uint32_t tiles_x = (w + (uint32_t)TW - 1u) / (uint32_t)TW;
  const uint32_t tile = map_tile(blockIdx.x, gridDim.x, s.tile_order);
  RayCount rc{0u, 0u, 0u, 0u};
  refill_tile<TV, TW, TH, NT>(s, L, u, fclamp, spp, (tile % tiles_x) * (uint32_t)TW, (tile / tiles_x) * (uint32_t)TH,
                              radiance, albedo, gbuf, rc, &s_next);
  count4(counters, rc.rays, rc.bsteps, rc.vsteps, rc.hits);
}

// ---------------------------------------------------------------------------
// temporal.wgsl:17-77
// ---------------------------------------------------------------------------
template <int FMT>
VRT_D float4 load_accum(const void* p, uint64_t i) {
  if (FMT == 1) return ((const float4*)p)[i];
  const ushort4 v = ((const ushort4*)p)[i];
  return make_float4(f16_to_f32(v.x), f16_to_f32(v.y), f16_to_f32(v.z), f16_to_f32(v.w));
}
template <int FMT>
VRT_D void store_accum(void* p, uint64_t i, float4 v) {
  if (FMT == 1) {
    ((float4*)p)[i] = v;
  } else {
    ((ushort4*)p)[i] = make_ushort4(f32_to_f16(v.x), f32_to_f16(v.y), f32_to_f16(v.z), f32_to_f16(v.w));
  }
}

// The stored value of an accumulation texel: FP32, or rounded to f16 (RNE) per
// component -- what load_accum returns after store_accum.
template <int FMT>
VRT_D float4 stored_accum(float4 v) {
  if (FMT == 1) return v;
  return make_float4(f16_to_f32(f32_to_f16(v.x)), f16_to_f32(f32_to_f16(v.y)), f16_to_f32(f32_to_f16(v.z)),
                     f16_to_f32(f32_to_f16(v.w)));
}

// temporal.wgsl:17-77 for pixel (px, py): the accumulated colour and history
// count before the store's rounding.
template <int FMT>
VRT_D float4 temporal_value(const Uniforms& u, uint32_t px, uint32_t py, const ushort4* __restrict__ radiance,
                            const float4* __restrict__ gcur, const void* __restrict__ aprev,
                            const float4* __restrict__ gprev) {
  const uint32_t w = (uint32_t)u.params1[0], h = (uint32_t)u.params1[1];
  const uint64_t i = (uint64_t)py * w + px;
  const ushort4 rv = radiance[i];
  const f3 cur = mk3(f16_to_f32(rv.x), f16_to_f32(rv.y), f16_to_f32(rv.z));
  const float4 g = gcur[i];
  const bool temporalOn = (u.params0[2] & 1u) != 0u;
  if (g.w < 0.0f || !temporalOn) return make_float4(cur.x, cur.y, cur.z, 1.0f);
  const float uvx = ((float)px + 0.5f) / (float)w;
  const float uvy = ((float)py + 0.5f) / (float)h;
  const f3 worldPos = mk3(u.cam_pos[0], u.cam_pos[1], u.cam_pos[2]) + camera_ray(u, uvx, uvy) * g.w;
  const float4 prevClip = mat_vec4(u.prev_view_proj, worldPos.x, worldPos.y, worldPos.z, 1.0f);
  f3 color = cur;
  float history = 1.0f;
  if (prevClip.w > 0.0f) {
    const float ndcx = prevClip.x / prevClip.w;
    const float ndcy = prevClip.y / prevClip.w;
    const float pux = ndcx * 0.5f + 0.5f;
    const float puy = (-ndcy) * 0.5f + 0.5f;
    if (pux > 0.0f && puy > 0.0f && pux < 1.0f && puy < 1.0f) {
      int ppx = (int)(pux * (float)w);
      int ppy = (int)(puy * (float)h);
      ppx = ppx > (int)w - 1 ? (int)w - 1 : ppx;
      ppy = ppy > (int)h - 1 ? (int)h - 1 : ppy;
      const float4 pg = gprev[(uint64_t)ppy * w + ppx];
      const f3 dwp = worldPos - mk3(u.prev_cam_pos[0], u.prev_cam_pos[1], u.prev_cam_pos[2]);
      const float expectedT = vsqrt(dot(dwp, dwp));
      const bool depthOk = pg.w > 0.0f && vabs(pg.w - expectedT) < 0.08f * expectedT + 0.02f;
      const bool normalOk = dot(mk3(pg.x, pg.y, pg.z), mk3(g.x, g.y, g.z)) > 0.85f;
      if (depthOk && normalOk) {
        // textureSampleLevel(linear, clamp-to-edge) restated in FP32, with
        // the texel coordinate rounded to 8 fractional bits as a texture
        // unit's fixed-point filter does (sub_texel)
        float flx, fly, fx, fy;
        sub_texel(pux * (float)w - 0.5f, flx, fx);
        sub_texel(puy * (float)h - 0.5f, fly, fy);
        int x0 = (int)flx, y0 = (int)fly;
        int x1 = x0 + 1, y1 = y0 + 1;
        x0 = x0 < 0 ? 0 : (x0 > (int)w - 1 ? (int)w - 1 : x0);
        x1 = x1 < 0 ? 0 : (x1 > (int)w - 1 ? (int)w - 1 : x1);
        y0 = y0 < 0 ? 0 : (y0 > (int)h - 1 ? (int)h - 1 : y0);
        y1 = y1 < 0 ? 0 : (y1 > (int)h - 1 ? (int)h - 1 : y1);
        const float4 t00 = load_accum<FMT>(aprev, (uint64_t)y0 * w + x0);
        const float4 t10 = load_accum<FMT>(aprev, (uint64_t)y0 * w + x1);
        const float4 t01 = load_accum<FMT>(aprev, (uint64_t)y1 * w + x0);
        const float4 t11 = load_accum<FMT>(aprev, (uint64_t)y1 * w + x1);
        float4 prev;
        prev.x = mix(mix(t00.x, t10.x, fx), mix(t01.x, t11.x, fx), fy);
        prev.y = mix(mix(t00.y, t10.y, fx), mix(t01.y, t11.y, fx), fy);
        prev.z = mix(mix(t00.z, t10.z, fx), mix(t01.z, t11.z, fx), fy);
        prev.w = mix(mix(t00.w, t10.w, fx), mix(t01.w, t11.w, fx), fy);
        history = vmin(prev.w, u.params4[1]) + 1.0f;
        const float alpha = 1.0f / history;
        color = mix(mk3(prev.x, prev.y, prev.z), cur, alpha);
      }
    }
  }
  return make_float4(color.x, color.y, color.z, history);
}

template <int FMT>
VRT_D void temporal_tile(const Uniforms& u, uint32_t tile, const ushort4* __restrict__ radiance,
                         const float4* __restrict__ gcur, const void* __restrict__ aprev,
                         const float4* __restrict__ gprev, void* __restrict__ aout) {
  const uint32_t w = (uint32_t)u.params1[0], h = (uint32_t)u.params1[1];
  uint32_t px, py;
  if (!tile_pixel(w, h, tile, px, py)) return;
  store_accum<FMT>(aout, (uint64_t)py * w + px, temporal_value<FMT>(u, px, py, radiance, gcur, aprev, gprev));
}

template <int FMT>
__global__ __launch_bounds__(256) void k_temporal(Uniforms u, const ushort4* __restrict__ radiance,
                                                  const float4* __restrict__ gcur,
                                                  const void* __restrict__ aprev,
                                                  const float4* __restrict__ gprev,
                                                  void* __restrict__ aout) {
  temporal_tile<FMT>(u, blockIdx.x, radiance, gcur, aprev, gprev, aout);
}

// ---------------------------------------------------------------------------
// atrous.wgsl:26-83.  A 16x16 tile stages its (16+4*step)^2 neighbourhood in
// LDS: per texel the colour with its luminance in the fourth component
// (computed once per texel instead of once per tap), its t, and a code for
// its normal.  Sky texels and texels outside the image are staged with colour
// 0, t = -1 and code NCODE_SKY.
//
// Normal weight.  pow(max(dot(n, nq), 0), 32) (atrous.wgsl:73) is exactly 1
// or +0 when both normals are axis vectors or zero -- what the path tracer
// writes (-vmask * stepDir, voxel.wgsl:126): 1 iff they are the same non-zero
// vector.  Each texel's normal is coded 1..6 (+x,-x,+y,-y,+z,-z), 0 (zero
// vector) or NCODE_GENERAL (anything else); the centre's zero code is mapped
// to a value no tap carries, and the weight becomes a code compare.  A tile
// with any general normal in its neighbourhood (host-supplied G-buffers) takes
// the pow32 form instead, reading the normals from the G-buffer.
//
// Weights: ((wKernel * wNormal) * wDepth) * wLum as atrous.wgsl:76 multiplies
// them, wDepth and wLum each through exp_w; the divisions by the per-pixel
// depth scale and sigmaL as products with their correctly rounded
// reciprocals, taken once per pixel (within 1.5 ulps of the quotient, inside
// the 2.5 ulps WGSL allows division).
// ---------------------------------------------------------------------------
constexpr uint32_t NCODE_GENERAL = 7u, NCODE_SKY = 15u, NCODE_NONE = 31u;

VRT_D uint32_t normal_code(float x, float y, float z) {
  const float ax = vabs(x), ay = vabs(y), az = vabs(z);
  const uint32_t nz = (ax != 0.0f ? 1u : 0u) + (ay != 0.0f ? 1u : 0u) + (az != 0.0f ? 1u : 0u);
  if (nz == 0u) return 0u;
  if (nz != 1u || !(ax == 1.0f || ay == 1.0f || az == 1.0f)) return NCODE_GENERAL;
  const uint32_t axis = ax != 0.0f ? 0u : (ay != 0.0f ? 1u : 2u);
  const float v = axis == 0u ? x : (axis == 1u ? y : z);
  return 1u + 2u * axis + (v < 0.0f ? 1u : 0u);
}

template <int STEP>
struct AtrousLds {
  static constexpr int T = 16, HALO = 2 * STEP, S = T + 2 * HALO;
  float4 c[S * S];  // rgb, luminance
  float2 g[S * S];  // t, normal code (as float)
};

template <int STEP>
VRT_D void atrous_filter(const Uniforms& u, const AtrousLds<STEP>& lds, bool general, int ox, int oy, float4 center,
                         const float4* __restrict__ gbuf, ushort4* __restrict__ dst);

template <int FMT, int STEP>
VRT_D void atrous_tile(const Uniforms& u, uint32_t tile, const void* __restrict__ src,
                       const float4* __restrict__ gbuf, ushort4* __restrict__ dst, AtrousLds<STEP>& lds) {
  constexpr int T = 16;
  constexpr int HALO = 2 * STEP;
  constexpr int S = T + 2 * HALO;
  float4* sc = lds.c;
  float2* sg = lds.g;
  const int w = (int)u.params1[0], h = (int)u.params1[1];
  const int tiles_x = (w + T - 1) / T;
  const int ox = (int)(tile % (uint32_t)tiles_x) * T - HALO;
  const int oy = (int)(tile / (uint32_t)tiles_x) * T - HALO;
  const f3 lw = mk3(0.2126f, 0.7152f, 0.0722f);
  int general = 0;
  for (int k = (int)threadIdx.x; k < S * S; k += 256) {
    const int lx = k % S, ly = k / S;
    const int gx = ox + lx, gy = oy + ly;
    float4 c = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float2 tg = make_float2(-1.0f, __builtin_bit_cast(float, NCODE_SKY));
    if (gx >= 0 && gy >= 0 && gx < w && gy < h) {
      const uint64_t j = (uint64_t)gy * w + gx;
      const float4 g = gbuf[j];
      if (!(g.w < 0.0f)) {  // sky taps add nothing (atrous.wgsl:69)
        c = load_accum<FMT>(src, j);
        c.w = dot(mk3(c.x, c.y, c.z), lw);  // luminance (atrous.wgsl:22-24)
        const uint32_t code = normal_code(g.x, g.y, g.z);
        general |= code == NCODE_GENERAL ? 1 : 0;
        tg = make_float2(g.w, __builtin_bit_cast(float, code));
      }
    }
    sc[k] = c;
    sg[k] = tg;
  }
  general = __syncthreads_or(general);
  const int x = ox + HALO + (int)(threadIdx.x & 15u), y = oy + HALO + (int)(threadIdx.x >> 4);
  if (x >= w || y >= h) return;
  atrous_filter<STEP>(u, lds, general != 0, ox, oy, load_accum<FMT>(src, (uint64_t)y * w + x), gbuf, dst);
}

// The filter of atrous_tile for the calling thread's pixel, from the staged
// neighbourhood (tile origin ox, oy including the halo); `center` is the
// pixel's accumulation value with its history count.  The pixel is inside
// the image.
template <int STEP>
VRT_D void atrous_filter(const Uniforms& u, const AtrousLds<STEP>& lds, bool general, int ox, int oy, float4 center,
                         const float4* __restrict__ gbuf, ushort4* __restrict__ dst) {
  constexpr int HALO = 2 * STEP;
  constexpr int S = 16 + 2 * HALO;
  const float4* sc = lds.c;
  const float2* sg = lds.g;
  const int w = (int)u.params1[0];
  const int tx = (int)(threadIdx.x & 15u), ty = (int)(threadIdx.x >> 4);
  const int x = ox + HALO + tx, y = oy + HALO + ty;
  const int cidx = (ty + HALO) * S + (tx + HALO);
  const uint64_t i = (uint64_t)y * w + x;
  const float2 g = sg[cidx];
  if (g.x < 0.0f) {  // sky: copy (atrous.wgsl:38-41)
    dst[i] = make_ushort4(f32_to_f16(center.x), f32_to_f16(center.y), f32_to_f16(center.z),
                          f32_to_f16(center.w));
    return;
  }
  const float K[3] = {0.375f, 0.25f, 0.0625f};
  const float history = vmax(center.w, 1.0f);
  const float lumC = sc[cidx].w;
  const float relax = 2.0f / vsqrt(history);
  const float sigmaL = vmax(0.08f, relax);
  f3 sum = mk3(center.x, center.y, center.z) * (K[0] * K[0]);
  float wSum = K[0] * K[0];
  const float dscale = 0.05f * g.x + 0.01f;
  const float rdscale = 1.0f / dscale, rsigmaL = 1.0f / sigmaL;
  if (!general) {
    const uint32_t cc0 = __builtin_bit_cast(uint32_t, g.y);
    const uint32_t cc = cc0 == 0u ? NCODE_NONE : cc0;  // a zero normal weighs every tap 0
#pragma unroll
    for (int dy = -2; dy <= 2; ++dy) {
#pragma unroll
      for (int dx = -2; dx <= 2; ++dx) {
        if (dx == 0 && dy == 0) continue;
        // sky and outside taps have code NCODE_SKY and colour 0: weight +0,
        // contribution 0, as if skipped (atrous.wgsl:64-69)
        const int l = (ty + HALO + dy * STEP) * S + (tx + HALO + dx * STEP);
        const float2 gq = sg[l];
        const float4 cq = sc[l];
        const float wKN = __builtin_bit_cast(uint32_t, gq.y) == cc ? K[dx < 0 ? -dx : dx] * K[dy < 0 ? -dy : dy] : 0.0f;
        const float wDepth = exp_w(-(vabs(g.x - gq.x) * rdscale));
        const float wLum = exp_w(-(vabs(lumC - cq.w) * rsigmaL));
        const float wt = (wKN * wDepth) * wLum;
        sum = sum + mk3(cq.x, cq.y, cq.z) * wt;
        wSum = wSum + wt;
      }
    }
  } else {
    const float4 gc = gbuf[i];
    const f3 gn = mk3(gc.x, gc.y, gc.z);
    for (int dy = -2; dy <= 2; ++dy) {
      for (int dx = -2; dx <= 2; ++dx) {
        if (dx == 0 && dy == 0) continue;
        const int l = (ty + HALO + dy * STEP) * S + (tx + HALO + dx * STEP);
        const float2 gq = sg[l];
        if (gq.x < 0.0f) continue;  // sky or outside the image
        const float4 cq = sc[l];
        const float4 nq = gbuf[(uint64_t)(y + dy * STEP) * w + (x + dx * STEP)];
        const float wKernel = K[dx < 0 ? -dx : dx] * K[dy < 0 ? -dy : dy];
        const float wNormal = pow32(vmax(dot(gn, mk3(nq.x, nq.y, nq.z)), 0.0f));
        const float wDepth = exp_w(-(vabs(g.x - gq.x) * rdscale));
        const float wLum = exp_w(-(vabs(lumC - cq.w) * rsigmaL));
        const float wt = ((wKernel * wNormal) * wDepth) * wLum;
        sum = sum + mk3(cq.x, cq.y, cq.z) * wt;
        wSum = wSum + wt;
      }
    }
  }
  const f3 o = sum / vmax(wSum, 1e-6f);
  dst[i] = make_ushort4(f32_to_f16(o.x), f32_to_f16(o.y), f32_to_f16(o.z), f32_to_f16(center.w));
}

template <int FMT, int STEP>
__global__ __launch_bounds__(256) void k_atrous(Uniforms u, const void* __restrict__ src,
                                                const float4* __restrict__ gbuf,
                                                ushort4* __restrict__ dst) {
  __shared__ AtrousLds<STEP> lds;
  atrous_tile<FMT, STEP>(u, blockIdx.x, src, gbuf, dst, lds);
}

// ---------------------------------------------------------------------------
// batched form (vrt_batch.h): one launch per pass over every view's tiles
// ---------------------------------------------------------------------------
// The view's own tile for batch tile b: the views keep their order in the
// batch; with tile_order 2/3 each view's tiles run from its last one.
VRT_D uint32_t view_tile(const ViewDesc& d, uint32_t b) {
  const uint32_t t = b - d.tile0;
  if (!(d.s.tile_order & 2u)) return t;
  const uint32_t n = d.tiles_x * ((d.h + 15u) >> 4);
  return n - 1u - t;
}

template <int TV, int WV>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WV))) void k_pathtrace_batch(
    const ViewDesc* __restrict__ views, uint32_t nviews, unsigned long long* __restrict__ counters) {
  extern __shared__ uint32_t lds_dyn[];
  const uint32_t b = blockIdx.x;
  const uint32_t vi = find_view(views, nviews, b);
  const ViewDesc& d = views[vi];
  const LaneMem L = lds_setup<TV>(d.s, lds_dyn);
  RayCount rc{0u, 0u, 0u, 0u};
  pathtrace_tile<TV>(d.s, L, d.u, d.fclamp, d.spp, view_tile(d, b), (ushort4*)d.radiance, (uchar4*)d.albedo,
                     (float4*)d.gbuf_cur, rc);
  count4(counters, rc.rays, rc.bsteps, rc.vsteps, rc.hits);
  count4(d.vcount, rc.rays, rc.bsteps, rc.vsteps, rc.hits);
}

// Persistent form: a fixed grid of blocks pulls tiles of every view from the
// table's shared queue (one atomic per tile), so views of any size and cost
// share the machine until the last tile; LDS-staged occupancy is reloaded
// only when consecutive tiles belong to different scenes.  The grid does not
// depend on the batch, so the launch can be replayed from a captured graph.
template <int TV, int WV>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WV))) void k_pathtrace_batch_q(
    BatchHead* __restrict__ head, const ViewDesc* __restrict__ views, unsigned long long* __restrict__ counters) {
  extern __shared__ uint32_t lds_dyn[];
  __shared__ uint32_t s_tile;
  const uint32_t tiles = head->tiles, nviews = head->nviews;
  const void* loaded = nullptr;  // scene whose occupancy is in LDS
  LaneMem L{{nullptr, 0u, 0u, 0u, 0u}};
  for (;;) {
    __syncthreads();  // every lane is done with the previous tile (s_tile, LDS)
    if (threadIdx.x == 0u) s_tile = atomicAdd(&head->queue, 1u);
    __syncthreads();
    // block-uniform: read into a scalar register, so the view's descriptor
    // (scene, uniforms) loads as scalars instead of into every lane's VGPRs
    const uint32_t b = __builtin_amdgcn_readfirstlane(s_tile);
    if (b >= tiles) break;
    const uint32_t vi = find_view(views, nviews, b);
    const ViewDesc& d = views[vi];
    if (tv_fast(TV)) {
      if (d.s.pool != loaded) {
        L = lds_setup<TV>(d.s, lds_dyn);
        loaded = d.s.pool;
      }
    }
    RayCount rc{0u, 0u, 0u, 0u};
    pathtrace_tile<TV>(d.s, L, d.u, d.fclamp, d.spp, view_tile(d, b), (ushort4*)d.radiance, (uchar4*)d.albedo,
                       (float4*)d.gbuf_cur, rc);
    count4(counters, rc.rays, rc.bsteps, rc.vsteps, rc.hits);
    count4(d.vcount, rc.rays, rc.bsteps, rc.vsteps, rc.hits);
  }
}

// The view's own refill tile for batch refill tile b, ordered as view_tile
// orders the 16x16 ones.
template <int TH>
VRT_D uint32_t view_rtile(const ViewDesc& d, uint32_t b) {
  const uint32_t t = b - d.rtile0;
  if (!(d.s.tile_order & 2u)) return t;
  const uint32_t n = d.rtx * ((d.h + (uint32_t)TH - 1u) / (uint32_t)TH);
  return n - 1u - t;
}

template <int TV, int TW, int TH, int NT>
VRT_D void batch_refill_tile(const ViewDesc& d, const LaneMem& L, uint32_t b, unsigned long long* counters,
                             uint32_t* s_next) {
  const uint32_t t = view_rtile<TH>(d, b);
  RayCount rc{0u, 0u, 0u, 0u};
  refill_tile<TV, TW, TH, NT>(d.s, L, d.u, d.fclamp, d.spp, (t % d.rtx) * (uint32_t)TW, (t / d.rtx) * (uint32_t)TH,
                              (ushort4*)d.radiance, (uchar4*)d.albedo, (float4*)d.gbuf_cur, rc, s_next);
  count4(counters, rc.rays, rc.bsteps, rc.vsteps, rc.hits);
  count4(d.vcount, rc.rays, rc.bsteps, rc.vsteps, rc.hits);
}

// Refill path tracing over every view's TW x TH tiles, one block per tile.
template <int TV, int WV, int TW, int TH, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WV))) void k_pathtrace_batch_r(
    const ViewDesc* __restrict__ views, uint32_t nviews, unsigned long long* __restrict__ counters) {
  extern __shared__ uint32_t lds_dyn[];
  __shared__ uint32_t s_next;
  const uint32_t b = blockIdx.x;
  const ViewDesc& d = views[find_view_r(views, nviews, b)];
  if (threadIdx.x == 0u) s_next = 0u;
  const LaneMem L = lds_setup<TV>(d.s, lds_dyn);
  if constexpr (NT > 64) __syncthreads();
  batch_refill_tile<TV, TW, TH, NT>(d, L, b, counters, &s_next);
}

// Persistent form of k_pathtrace_batch_r: each block pulls the next refill
// tile of any view from the table's queue (one atomic per tile); LDS-staged
// occupancy is reloaded only when consecutive tiles belong to different
// scenes.
template <int TV, int WV, int TW, int TH, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WV))) void k_pathtrace_batch_qr(
    BatchHead* __restrict__ head, const ViewDesc* __restrict__ views, unsigned long long* __restrict__ counters) {
  extern __shared__ uint32_t lds_dyn[];
  __shared__ uint32_t s_tile, s_next;
  const uint32_t tiles = head->rtiles, nviews = head->nviews;
  const void* loaded = nullptr;  // scene whose occupancy is in LDS
  LaneMem L{{nullptr, 0u, 0u, 0u, 0u}};
  for (;;) {
    // b is block-uniform; readfirstlane puts it in a scalar register, so the
    // view's descriptor (scene, uniforms) loads as scalars instead of into
    // every lane's VGPRs
    uint32_t b;
    if constexpr (NT == 64) {
      b = 0u;
      if (threadIdx.x == 0u) b = atomicAdd(&head->queue, 1u);
      b = __builtin_amdgcn_readfirstlane(__shfl(b, 0, 64));
    } else {
      __syncthreads();  // every lane is done with the previous tile (s_tile, s_next, LDS)
      if (threadIdx.x == 0u) {
        s_tile = atomicAdd(&head->queue, 1u);
        s_next = 0u;
      }
      __syncthreads();
      b = __builtin_amdgcn_readfirstlane(s_tile);
    }
    if (b >= tiles) break;
    const ViewDesc& d = views[find_view_r(views, nviews, b)];
    if (tv_fast(TV)) {
      if (d.s.pool != loaded) {
        L = lds_setup<TV>(d.s, lds_dyn);
        loaded = d.s.pool;
      }
    }
    batch_refill_tile<TV, TW, TH, NT>(d, L, b, counters, &s_next);
  }
}

VRT_D void temporal_batch_tile(const ViewDesc* __restrict__ views, uint32_t nviews, uint32_t b) {
  const uint32_t vi = find_view(views, nviews, b);
  const ViewDesc& d = views[vi];
  const uint32_t tile = b - d.tile0;
  if (d.fmt == 1u)
    temporal_tile<1>(d.u, tile, (const ushort4*)d.radiance, (const float4*)d.gbuf_cur, d.accum_prev,
                     (const float4*)d.gbuf_prev, d.accum_cur);
  else
    temporal_tile<0>(d.u, tile, (const ushort4*)d.radiance, (const float4*)d.gbuf_cur, d.accum_prev,
                     (const float4*)d.gbuf_prev, d.accum_cur);
}
__global__ __launch_bounds__(256) void k_temporal_batch(const ViewDesc* __restrict__ views, uint32_t nviews) {
  temporal_batch_tile(views, nviews, blockIdx.x);
}
// Grid-stride forms for the captured graph: the tile count comes from the table.
__global__ __launch_bounds__(256) void k_temporal_batch_gs(const BatchHead* __restrict__ head,
                                                           const ViewDesc* __restrict__ views) {
  for (uint32_t b = blockIdx.x; b < head->tiles; b += gridDim.x) temporal_batch_tile(views, head->nviews, b);
}

// pass 0: accum -> den0 (step 1), 1: den0 -> den1 (step 2), 2: den1 -> den0
// (step 4), main.js:502-507.  Views without denoise skip every pass.
template <int PASS>
VRT_D void atrous_batch_tile(const ViewDesc* __restrict__ views, uint32_t nviews, uint32_t b,
                             AtrousLds<PASS == 0 ? 1 : (PASS == 1 ? 2 : 4)>& lds) {
  constexpr int STEP = PASS == 0 ? 1 : (PASS == 1 ? 2 : 4);
  const uint32_t vi = find_view(views, nviews, b);
  const ViewDesc& d = views[vi];
  if (!d.denoise) return;
  const uint32_t tile = b - d.tile0;
  if (PASS == 0) {
    if (d.fmt == 1u)
      atrous_tile<1, STEP>(d.u, tile, d.accum_cur, (const float4*)d.gbuf_cur, (ushort4*)d.den0, lds);
    else
      atrous_tile<0, STEP>(d.u, tile, d.accum_cur, (const float4*)d.gbuf_cur, (ushort4*)d.den0, lds);
  } else {
    atrous_tile<0, STEP>(d.u, tile, PASS == 1 ? d.den0 : d.den1, (const float4*)d.gbuf_cur,
                         (ushort4*)(PASS == 1 ? d.den1 : d.den0), lds);
  }
}
template <int PASS>
__global__ __launch_bounds__(256) void k_atrous_batch(const ViewDesc* __restrict__ views, uint32_t nviews) {
  __shared__ AtrousLds<PASS == 0 ? 1 : (PASS == 1 ? 2 : 4)> lds;
  atrous_batch_tile<PASS>(views, nviews, blockIdx.x, lds);
}
template <int PASS>
__global__ __launch_bounds__(256) void k_atrous_batch_gs(const BatchHead* __restrict__ head,
                                                         const ViewDesc* __restrict__ views) {
  __shared__ AtrousLds<PASS == 0 ? 1 : (PASS == 1 ? 2 : 4)> lds;
  for (uint32_t b = blockIdx.x; b < head->tiles; b += gridDim.x) {
    atrous_batch_tile<PASS>(views, head->nviews, b, lds);
    __syncthreads();  // the next tile rewrites the LDS tile
  }
}

// ---------------------------------------------------------------------------
// main.js:628-659: hdr = illum * (alb8/255)^2, computed in double like JS
// ---------------------------------------------------------------------------
template <int FMT>
__global__ __launch_bounds__(256) void k_hdr(uint32_t npx, const void* __restrict__ illum,
                                             const uchar4* __restrict__ albedo,
                                             float* __restrict__ hdr) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npx) return;
  const float4 il = load_accum<FMT>(illum, i);
  const uchar4 a = albedo[i];
  const double ar = (double)a.x / 255.0, ag = (double)a.y / 255.0, ab = (double)a.z / 255.0;
  hdr[i * 3 + 0] = (float)((double)il.x * ar * ar);
  hdr[i * 3 + 1] = (float)((double)il.y * ag * ag);
  hdr[i * 3 + 2] = (float)((double)il.z * ab * ab);
}

// present.wgsl:40-55 in FP32 with this project's pow: illum * (alb8/255)^2 *
// exposure, ACES fit, gamma 1/2.2, stored rgba8unorm.
template <int FMT>
__global__ __launch_bounds__(256) void k_present(uint32_t npx, const void* __restrict__ illum,
                                                 const uchar4* __restrict__ albedo, float exposure,
                                                 uchar4* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npx) return;
  const float4 il = load_accum<FMT>(illum, i);
  const uchar4 a = albedo[i];
  const float inv_gamma = (float)(1.0 / 2.2);
  const float ic[3] = {il.x, il.y, il.z};
  const uint8_t ab[3] = {a.x, a.y, a.z};
  uint32_t o[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const float s = (float)ab[c] / 255.0f;
    const float x = (ic[c] * (s * s)) * exposure;
    const float num = x * (2.51f * x + 0.03f);
    const float den = x * (2.43f * x + 0.59f) + 0.14f;
    o[c] = f32_to_unorm8(pow_r(vclamp(num / den, 0.0f, 1.0f), inv_gamma));
  }
  out[i] = make_uchar4((uint8_t)o[0], (uint8_t)o[1], (uint8_t)o[2], 255);
}

// HDR of every view that asks for it (d.hdr != null): the denoised target
// when denoise is on, else the accumulation, as k_hdr.
VRT_D void hdr_batch_tile(const ViewDesc* __restrict__ views, uint32_t nviews, uint32_t b) {
  const uint32_t vi = find_view(views, nviews, b);
  const ViewDesc& d = views[vi];
  if (d.hdr == nullptr) return;
  uint32_t px, py;
  if (!tile_pixel(d.w, d.h, b - d.tile0, px, py)) return;
  const uint64_t i = (uint64_t)py * d.w + px;
  const float4 il = d.denoise ? load_accum<0>(d.den0, i)
                              : (d.fmt == 1u ? load_accum<1>(d.accum_cur, i) : load_accum<0>(d.accum_cur, i));
  const uchar4 a = ((const uchar4*)d.albedo)[i];
  const double ar = (double)a.x / 255.0, ag = (double)a.y / 255.0, ab = (double)a.z / 255.0;
  d.hdr[i * 3 + 0] = (float)((double)il.x * ar * ar);
  d.hdr[i * 3 + 1] = (float)((double)il.y * ag * ag);
  d.hdr[i * 3 + 2] = (float)((double)il.z * ab * ab);
}
__global__ __launch_bounds__(256) void k_hdr_batch(const ViewDesc* __restrict__ views, uint32_t nviews) {
  hdr_batch_tile(views, nviews, blockIdx.x);
}
__global__ __launch_bounds__(256) void k_hdr_batch_gs(const BatchHead* __restrict__ head,
                                                      const ViewDesc* __restrict__ views) {
  for (uint32_t b = blockIdx.x; b < head->tiles; b += gridDim.x) hdr_batch_tile(views, head->nviews, b);
}

}  // namespace vrt

#include "vrt_pathtrace_wave.h"
#include "vrt_trace_wave.h"

// ---------------------------------------------------------------------------
// host-side launchers (declared in vrt_kernels.h)
// ---------------------------------------------------------------------------
namespace vrt {

static inline uint32_t div_up(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// LDS of the single-loop variants: 16 mask words per lane (256-thread
// blocks), then the padded bitmap, its prefix counts and the row distances
// (bytes), kept within 64 KiB per block.  bit_words (the unpadded bitmap) is
// not staged by any variant; the argument keeps the launchers' signatures.
static inline uint32_t lds_words(int tv, uint32_t, uint32_t pwords) {
  return tv_fast(tv) ? 2u * pwords + (pwords + 3u) / 4u : 0u;
}
static inline uint32_t lds_bytes(int tv, uint32_t bit_words, uint32_t pwords) {
  return tv_fast(tv) ? (16u * 256u + lds_words(tv, bit_words, pwords)) * 4u : 0u;
}
// Variant actually launched: the LDS variants fall back to the single loop on
// global memory when what they stage does not fit.
static inline int lds_variant(int tv, uint32_t bit_words, uint32_t pwords) {
  if (!tv_fast(tv) || lds_words(tv, bit_words, pwords) <= LDS_BIT_WORDS_MAX) return tv;
  return TV_FASTG;
}

hipError_t launch_brick_masks(hipStream_t st, const uint32_t* vox, int nx, int ny, int bgx,
                              int bgy, int bz0, int nbz, uint32_t* masks, uint32_t* flags) {
  const uint64_t waves = (uint64_t)bgx * bgy * nbz;
  hipLaunchKernelGGL(k_brick_masks, dim3(div_up(waves, 4)), dim3(256), 0, st, vox, nx, ny, bgx,
                     bgy, bz0, nbz, masks, flags);
  return hipGetLastError();
}

hipError_t launch_edit_bricks(hipStream_t st, const uint32_t* vox, int nx, int ny, int bgx, int bgy, int bx0,
                              int by0, int bz0, int ex, int ey, int ez, const uint32_t* bricks, uint32_t* pool,
                              uint32_t* changed) {
  const uint64_t waves = (uint64_t)ex * ey * ez;
  hipLaunchKernelGGL(k_edit_bricks, dim3(div_up(waves, 4)), dim3(256), 0, st, vox, nx, ny, bgx, bgy, bx0, by0, bz0,
                     ex, ey, ez, bricks, pool, changed);
  return hipGetLastError();
}

hipError_t launch_brick_words(hipStream_t st, const uint32_t* flags, const uint32_t* slots,
                              uint64_t nb, uint32_t* bricks) {
  hipLaunchKernelGGL(k_brick_words, dim3(div_up(nb, 256)), dim3(256), 0, st, flags, slots, nb,
                     bricks);
  return hipGetLastError();
}

hipError_t launch_brick_bits(hipStream_t st, const uint32_t* bricks, uint64_t nb, uint32_t* bits) {
  hipLaunchKernelGGL(k_brick_bits, dim3(div_up(nb, 256)), dim3(256), 0, st, bricks, nb, bits);
  return hipGetLastError();
}

hipError_t launch_brick_pbits(hipStream_t st, const uint32_t* bricks, int bgx, int bgy, int bgz, uint32_t sh1,
                              uint32_t sh2, uint64_t nbits, uint32_t* bits) {
  hipLaunchKernelGGL(k_brick_pbits, dim3(div_up(nbits, 256)), dim3(256), 0, st, bricks, bgx, bgy, bgz, sh1, sh2,
                     nbits, bits);
  return hipGetLastError();
}

hipError_t launch_row_dist(hipStream_t st, const uint32_t* bricks, int bgx, int bgy, int bgz, uint32_t sh1,
                           uint32_t sh2, uint64_t pwords, uint8_t* tmp, uint8_t* pdist) {
  const uint64_t rows = (uint64_t)bgy * (uint64_t)bgz, nw = (pwords + 3u) & ~3ull;
  uint8_t *occ = tmp, *dy = tmp + rows, *dist = tmp + 2 * rows;
  hipLaunchKernelGGL(k_row_occ, dim3(div_up(rows, 256)), dim3(256), 0, st, bricks, bgx, bgy, bgz, occ);
  hipLaunchKernelGGL(k_row_dy, dim3(div_up(rows, 256)), dim3(256), 0, st, occ, bgy, bgz, dy);
  hipLaunchKernelGGL(k_row_dist, dim3(div_up(rows, 256)), dim3(256), 0, st, dy, bgy, bgz, dist);
  hipLaunchKernelGGL(k_word_dist, dim3(div_up(nw, 256)), dim3(256), 0, st, dist, bgy, bgz, sh1, sh2, pwords, nw,
                     pdist);
  return hipGetLastError();
}

hipError_t launch_brick_vbox(hipStream_t st, const uint32_t* pool, uint64_t nslots, uint32_t* vbox) {
  if (nslots == 0) return hipSuccess;
  hipLaunchKernelGGL(k_brick_vbox, dim3(div_up(nslots, 256)), dim3(256), 0, st, pool, nslots, vbox);
  return hipGetLastError();
}

hipError_t launch_row_span(hipStream_t st, const uint32_t* bricks, int bgx, int bgy, int bgz, uint32_t* span) {
  const uint64_t rows = (uint64_t)bgy * (uint64_t)bgz;
  hipLaunchKernelGGL(k_row_span, dim3(div_up(rows, 256)), dim3(256), 0, st, bricks, bgx, bgy, bgz, span);
  return hipGetLastError();
}

hipError_t launch_popc(hipStream_t st, const uint32_t* words, uint64_t n, uint32_t* counts) {
  hipLaunchKernelGGL(k_popc, dim3(div_up(n, 256)), dim3(256), 0, st, words, n, counts);
  return hipGetLastError();
}

hipError_t launch_pack_pool(hipStream_t st, const uint32_t* vox, int nx, int ny, int bgx, int bgy,
                            int bz0, int nbz, const uint32_t* masks, const uint32_t* bricks,
                            uint32_t* pool) {
  const uint64_t waves = (uint64_t)bgx * bgy * nbz;
  hipLaunchKernelGGL(k_pack_pool, dim3(div_up(waves, 4)), dim3(256), 0, st, vox, nx, ny, bgx, bgy,
                     bz0, nbz, masks, bricks, pool);
  return hipGetLastError();
}

hipError_t launch_trace_rays(hipStream_t st, const SceneDev& s, const float* rays, uint64_t n,
                             float tmax, void* hits, unsigned long long* counters, int blocks, int tv,
                             int waves) {
  if (blocks <= 0) blocks = 256 * 16;
  const uint32_t need = div_up(n, 256);
  const uint32_t nb = need < (uint32_t)blocks ? need : (uint32_t)blocks;
  const dim3 g(nb > 0 ? nb : 1);
  tv = lds_variant(tv, s.bit_words, s.pwords);
  const uint32_t shmem = lds_bytes(tv, s.bit_words, s.pwords);
  const bool w6 = waves == 6;
  if (tv == TV_FAST && w6)
    hipLaunchKernelGGL((k_trace_rays<TV_FAST, 6>), g, dim3(256), shmem, st, s, rays, n, tmax, (uint4*)hits, counters);
  else if (tv == TV_FAST)
    hipLaunchKernelGGL(k_trace_rays<TV_FAST>, g, dim3(256), shmem, st, s, rays, n, tmax, (uint4*)hits, counters);
  else if (tv == TV_FASTG && w6)
    hipLaunchKernelGGL((k_trace_rays<TV_FASTG, 6>), g, dim3(256), 0, st, s, rays, n, tmax, (uint4*)hits, counters);
  else if (tv == TV_FASTG)
    hipLaunchKernelGGL(k_trace_rays<TV_FASTG>, g, dim3(256), 0, st, s, rays, n, tmax, (uint4*)hits, counters);
  else if (tv == TV_FASTSKIP)
    hipLaunchKernelGGL(k_trace_rays<TV_FASTSKIP>, g, dim3(256), shmem, st, s, rays, n, tmax, (uint4*)hits,
                       counters);
  else
    hipLaunchKernelGGL(k_trace_rays<TV_NESTED>, g, dim3(256), 0, st, s, rays, n, tmax, (uint4*)hits, counters);
  return hipGetLastError();
}

// Zero n 32-bit words with a kernel: a launch a caller may capture into a
// graph (no memset node).
__global__ __launch_bounds__(64) void k_zero_words(uint32_t* __restrict__ p, uint32_t n) {
  for (uint32_t i = threadIdx.x; i < n; i += 64u) p[i] = 0u;
}

hipError_t launch_zero_words(hipStream_t st, void* p, uint32_t n) {
  hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(64), 0, st, (uint32_t*)p, n);
  return hipGetLastError();
}

hipError_t launch_trace_persistent(hipStream_t st, const SceneDev& s, const float* rays, uint32_t n,
                                  float tmax, void* hits, unsigned long long* counters,
                                  uint32_t* queue, int blocks, int steps, int min_refill, int tv, int waves,
                                  bool clear_queue) {
  // the persistent kernel runs the single loop in LDS; where the scene's
  // bitmaps do not fit, the one-ray-per-lane kernel on global memory
  tv = lds_variant(tv_fast(tv) ? tv : TV_FAST, s.bit_words, s.pwords);
  if (!tv_fast(tv)) return launch_trace_rays(st, s, rays, n, tmax, hits, counters, 0, TV_FASTG, waves);
  if (clear_queue) {
    const hipError_t e = hipMemsetAsync(queue, 0, sizeof(uint32_t), st);
    if (e != hipSuccess) return e;
  }
  const dim3 g(blocks > 0 ? blocks : 1);
  const int sp = steps < 1 ? 1 : steps, mr = min_refill < 1 ? 1 : min_refill;
  const uint32_t shmem = lds_bytes(tv, s.bit_words, s.pwords);
  if (tv == TV_FASTSKIP)
    hipLaunchKernelGGL((k_trace_persistent<FastRunner<true>, TV_FASTSKIP>), g, dim3(256), shmem, st, s, rays, n,
                       tmax, (uint4*)hits, counters, queue, sp, mr);
  else if (waves == 6)
    hipLaunchKernelGGL((k_trace_persistent<FastRunner<false>, TV_FAST, 6>), g, dim3(256), shmem, st, s, rays, n,
                       tmax, (uint4*)hits, counters, queue, sp, mr);
  else
    hipLaunchKernelGGL((k_trace_persistent<FastRunner<false>, TV_FAST>), g, dim3(256), shmem, st, s, rays, n, tmax,
                       (uint4*)hits, counters, queue, sp, mr);
  return hipGetLastError();
}

hipError_t launch_pathtrace(hipStream_t st, const SceneDev& s, const Uniforms& u, float fclamp,
                            uint32_t spp, void* radiance, void* albedo, void* gbuf,
                            unsigned long long* counters, int tv, int waves, int form) {
  const uint32_t w = (uint32_t)u.params1[0], h = (uint32_t)u.params1[1];
  const uint32_t tiles = div_up(w, 16) * div_up(h, 16);
  tv = lds_variant(tv, s.bit_words, s.pwords);
  if ((form == 4 || form == 5) && tv == TV_FASTG) {
    // refill on one-wave blocks: 16x16 (form 4) or 16x8 (form 5) tiles
    const uint32_t th = form == 4 ? 16u : 8u;
    const dim3 grid(div_up(w, 16) * div_up(h, th));
    auto go = [&](auto kernel) {
      hipLaunchKernelGGL(kernel, grid, dim3(64), 0, st, s, u, fclamp, spp, (ushort4*)radiance, (uchar4*)albedo,
                         (float4*)gbuf, counters);
    };
    const int wv = waves == 4 || waves == 5 ? waves : 1;
    switch ((form == 4 ? 10 : 0) + wv) {
      case 11: go(k_pathtrace_refill<TV_FASTG, 1, 16, 16, 64>); break;
      case 14: go(k_pathtrace_refill<TV_FASTG, 4, 16, 16, 64>); break;
      case 15: go(k_pathtrace_refill<TV_FASTG, 5, 16, 16, 64>); break;
      case 1: go(k_pathtrace_refill<TV_FASTG, 1, 16, 8, 64>); break;
      case 4: go(k_pathtrace_refill<TV_FASTG, 4, 16, 8, 64>); break;
      default: go(k_pathtrace_refill<TV_FASTG, 5, 16, 8, 64>); break;
    }
    return hipGetLastError();
  }
  if ((form == 2 || form == 3) && (tv == TV_FAST || tv == TV_FASTG)) {
    // refill: 32x32 (form 2) or 32x16 (form 3) tiles per block
    const uint32_t th = form == 2 ? 32u : 16u;
    const dim3 grid(div_up(w, 32) * div_up(h, th));
    const uint32_t shmem = lds_bytes(tv, s.bit_words, s.pwords);
    auto go = [&](auto kernel) {
      hipLaunchKernelGGL(kernel, grid, dim3(256), shmem, st, s, u, fclamp, spp, (ushort4*)radiance,
                         (uchar4*)albedo, (float4*)gbuf, counters);
    };
    const int wv = waves == 4 || waves == 5 ? waves : 1;
    const int key = (tv == TV_FASTG ? 100 : 0) + (form == 2 ? 10 : 0) + wv;
    switch (key) {
      case 111: go(k_pathtrace_refill<TV_FASTG, 1, 32, 32>); break;
      case 114: go(k_pathtrace_refill<TV_FASTG, 4, 32, 32>); break;
      case 115: go(k_pathtrace_refill<TV_FASTG, 5, 32, 32>); break;
      case 101: go(k_pathtrace_refill<TV_FASTG, 1, 32, 16>); break;
      case 104: go(k_pathtrace_refill<TV_FASTG, 4, 32, 16>); break;
      case 105: go(k_pathtrace_refill<TV_FASTG, 5, 32, 16>); break;
      case 11: go(k_pathtrace_refill<TV_FAST, 1, 32, 32>); break;
      case 14: go(k_pathtrace_refill<TV_FAST, 4, 32, 32>); break;
      case 15: go(k_pathtrace_refill<TV_FAST, 5, 32, 32>); break;
      case 1: go(k_pathtrace_refill<TV_FAST, 1, 32, 16>); break;
      case 4: go(k_pathtrace_refill<TV_FAST, 4, 32, 16>); break;
      default: go(k_pathtrace_refill<TV_FAST, 5, 32, 16>); break;
    }
    return hipGetLastError();
  }
  if (form == 1 && tv == TV_FASTG) {
    auto go8 = [&](auto kernel) {
      hipLaunchKernelGGL(kernel, dim3(div_up(w, 8) * div_up(h, 8)), dim3(64), 0, st, s, u, fclamp, spp,
                         (ushort4*)radiance, (uchar4*)albedo, (float4*)gbuf, counters);
    };
    if (waves == 4) go8(k_pathtrace8<4>);
    else if (waves == 5) go8(k_pathtrace8<5>);
    else go8(k_pathtrace8<1>);
    return hipGetLastError();
  }
  const uint32_t shmem = lds_bytes(tv, s.bit_words, s.pwords);
  auto go = [&](auto kernel) {
    hipLaunchKernelGGL(kernel, dim3(tiles), dim3(256), shmem, st, s, u, fclamp, spp, (ushort4*)radiance,
                       (uchar4*)albedo, (float4*)gbuf, counters);
  };
  switch (tv * 8 + (waves == 4 || waves == 5 ? waves : 1)) {
    case TV_FASTG * 8 + 4: go(k_pathtrace<TV_FASTG, 4>); break;
    case TV_FASTG * 8 + 5: go(k_pathtrace<TV_FASTG, 5>); break;
    case TV_FASTG * 8 + 1: go(k_pathtrace<TV_FASTG, 1>); break;
    case TV_FAST * 8 + 4: go(k_pathtrace<TV_FAST, 4>); break;
    case TV_FAST * 8 + 5: go(k_pathtrace<TV_FAST, 5>); break;
    case TV_FAST * 8 + 1: go(k_pathtrace<TV_FAST, 1>); break;
    case TV_NESTED * 8 + 4: go(k_pathtrace<TV_NESTED, 4>); break;
    case TV_NESTED * 8 + 5: go(k_pathtrace<TV_NESTED, 5>); break;
    default: go(k_pathtrace<TV_NESTED, 1>); break;
  }
  return hipGetLastError();
}

hipError_t launch_pathtrace_wave(hipStream_t st, const SceneDev& s, const Uniforms& u, float fclamp, uint32_t spp,
                                 void* radiance, void* albedo, void* gbuf, unsigned long long* counters,
                                 uint32_t* queue, int blocks, int min_wait, int shade_min, int waves,
                                 int form) {
  hipError_t e = launch_zero_words(st, queue, QWORDS);
  if (e != hipSuccess) return e;
  const uint32_t mw = min_wait < 1 ? 1u : (min_wait > 64 ? 64u : (uint32_t)min_wait);
  const uint32_t sm = shade_min < 1 ? 1u : (shade_min > 64 ? 64u : (uint32_t)shade_min);
  // the LDS form when the scene's bitmaps fit (four waves per block, so the
  // grid has a quarter of the blocks), else the global-memory form
  const bool lds = form != 1 && lds_variant(TV_FAST, s.bit_words, s.pwords) == TV_FAST;
  const uint32_t nt = lds ? 256u : 64u;
  const uint32_t shmem = lds ? lds_bytes(TV_FAST, s.bit_words, s.pwords) : 0u;
  const int nb = blocks > 0 ? (lds ? (blocks + 3) / 4 : blocks) : 1;
  auto go = [&](auto kernel) {
    hipLaunchKernelGGL(kernel, dim3(nb), dim3(nt), shmem, st, s, u, fclamp, spp, (ushort4*)radiance,
                       (uchar4*)albedo, (float4*)gbuf, counters, queue, mw, sm);
  };
  const int wv = waves == 4 || waves == 5 || waves == 6 ? waves : 1;
  switch (wv * 1000 + (int)nt) {
    case 4064: go(k_pathtrace_wave<4, 64>); break;
    case 5064: go(k_pathtrace_wave<5, 64>); break;
    case 6064: go(k_pathtrace_wave<6, 64>); break;
    case 1064: go(k_pathtrace_wave<1, 64>); break;
    case 4256: go(k_pathtrace_wave<4, 256>); break;
    case 5256: go(k_pathtrace_wave<5, 256>); break;
    case 6256: go(k_pathtrace_wave<6, 256>); break;
    default: go(k_pathtrace_wave<1, 256>); break;
  }
  return hipGetLastError();
}

hipError_t launch_trace_wave(hipStream_t st, const SceneDev& s, const float* rays, uint32_t n, float tmax,
                             void* hits, unsigned long long* counters, uint32_t* queue, int blocks, int min_wait,
                             int fin_min, int waves, bool clear_queue) {
  if (clear_queue) {
    const hipError_t e = launch_zero_words(st, queue, QWORDS);
    if (e != hipSuccess) return e;
  }
  const uint32_t mw = min_wait < 1 ? 1u : (min_wait > 64 ? 64u : (uint32_t)min_wait);
  const uint32_t fm = fin_min < 1 ? 1u : (fin_min > 64 ? 64u : (uint32_t)fin_min);
  // the LDS form when the scene's bitmaps fit (blocks of four waves), else
  // the global-memory form on one-wave blocks (four times as many)
  const bool lds = lds_variant(TV_FAST, s.bit_words, s.pwords) == TV_FAST;
  const uint32_t nt = lds ? 256u : 64u;
  const uint32_t shmem = lds ? lds_bytes(TV_FAST, s.bit_words, s.pwords) : 0u;
  const int nb = blocks > 0 ? (lds ? blocks : 4 * blocks) : 1;
  auto go = [&](auto kernel) {
    hipLaunchKernelGGL(kernel, dim3(nb), dim3(nt), shmem, st, s, rays, n, tmax, (uint4*)hits, counters, queue, mw,
                       fm);
  };
  if (lds) {
    if (waves == 6) go(k_trace_wave<6, 256>);
    else go(k_trace_wave<1, 256>);
  } else {
    if (waves == 6) go(k_trace_wave<6, 64>);
    else go(k_trace_wave<1, 64>);
  }
  return hipGetLastError();
}

hipError_t launch_pathtrace_wave_batch(hipStream_t st, BatchHead* head, const ViewDesc* views,
                                       unsigned long long* counters, int blocks, int min_wait, int shade_min,
                                       int waves) {
  const uint32_t mw = min_wait < 1 ? 1u : (min_wait > 64 ? 64u : (uint32_t)min_wait);
  const uint32_t sm = shade_min < 1 ? 1u : (shade_min > 64 ? 64u : (uint32_t)shade_min);
  const dim3 g(blocks > 0 ? blocks : 1);
  if (waves == 5) hipLaunchKernelGGL(k_pathtrace_wave_batch<5>, g, dim3(64), 0, st, head, views, counters, mw, sm);
  else if (waves == 6) hipLaunchKernelGGL(k_pathtrace_wave_batch<6>, g, dim3(64), 0, st, head, views, counters, mw, sm);
  else if (waves == 4) hipLaunchKernelGGL(k_pathtrace_wave_batch<4>, g, dim3(64), 0, st, head, views, counters, mw, sm);
  else hipLaunchKernelGGL(k_pathtrace_wave_batch<1>, g, dim3(64), 0, st, head, views, counters, mw, sm);
  return hipGetLastError();
}

hipError_t launch_batch_filters(hipStream_t st, BatchHead* head, const ViewDesc* views, uint32_t gs_blocks,
                                bool denoise, bool hdr) {
  const dim3 gg(gs_blocks > 0 ? gs_blocks : 1);
  hipLaunchKernelGGL(k_temporal_batch_gs, gg, dim3(256), 0, st, head, views);
  if (denoise) {
    hipLaunchKernelGGL(k_atrous_batch_gs<0>, gg, dim3(256), 0, st, head, views);
    hipLaunchKernelGGL(k_atrous_batch_gs<1>, gg, dim3(256), 0, st, head, views);
    hipLaunchKernelGGL(k_atrous_batch_gs<2>, gg, dim3(256), 0, st, head, views);
  }
  if (hdr) hipLaunchKernelGGL(k_hdr_batch_gs, gg, dim3(256), 0, st, head, views);
  return hipGetLastError();
}

int refill_form(int tv, int form, uint32_t max_bit_words, uint32_t max_pwords) {
  tv = lds_variant(tv, max_bit_words, max_pwords);
  if ((form == 4 || form == 5) && tv == TV_FASTG) return form;
  if ((form == 2 || form == 3) && (tv == TV_FAST || tv == TV_FASTG)) return form;
  return 0;
}

void refill_tile_dims(int rform, uint32_t* tw, uint32_t* th) {
  *tw = rform == 2 || rform == 3 ? 32u : 16u;
  *th = rform == 2 ? 32u : rform == 5 ? 8u : 16u;
}

// The batched refill kernel K<TV, WV, TW, TH, NT> for (tv, rform, waves).
template <template <int, int, int, int, int> class K, typename Go>
static void refill_batch_switch(int tv, int rform, int waves, Go go) {
  const int wv = waves == 4 || waves == 5 ? waves : 1;
  const int key = (tv == TV_FASTG ? 100 : 0) + rform * 10 + wv;
  switch (key) {
    case 141: go(K<TV_FASTG, 1, 16, 16, 64>{}); break;
    case 144: go(K<TV_FASTG, 4, 16, 16, 64>{}); break;
    case 145: go(K<TV_FASTG, 5, 16, 16, 64>{}); break;
    case 151: go(K<TV_FASTG, 1, 16, 8, 64>{}); break;
    case 154: go(K<TV_FASTG, 4, 16, 8, 64>{}); break;
    case 155: go(K<TV_FASTG, 5, 16, 8, 64>{}); break;
    case 121: go(K<TV_FASTG, 1, 32, 32, 256>{}); break;
    case 124: go(K<TV_FASTG, 4, 32, 32, 256>{}); break;
    case 125: go(K<TV_FASTG, 5, 32, 32, 256>{}); break;
    case 131: go(K<TV_FASTG, 1, 32, 16, 256>{}); break;
    case 134: go(K<TV_FASTG, 4, 32, 16, 256>{}); break;
    case 135: go(K<TV_FASTG, 5, 32, 16, 256>{}); break;
    case 21: go(K<TV_FAST, 1, 32, 32, 256>{}); break;
    case 24: go(K<TV_FAST, 4, 32, 32, 256>{}); break;
    case 25: go(K<TV_FAST, 5, 32, 32, 256>{}); break;
    case 31: go(K<TV_FAST, 1, 32, 16, 256>{}); break;
    case 34: go(K<TV_FAST, 4, 32, 16, 256>{}); break;
    default: go(K<TV_FAST, 5, 32, 16, 256>{}); break;
  }
}

template <int TV, int WV, int TW, int TH, int NT>
struct BatchR {
  static constexpr int nt = NT;
  static constexpr auto fn = k_pathtrace_batch_r<TV, WV, TW, TH, NT>;
};
template <int TV, int WV, int TW, int TH, int NT>
struct BatchQR {
  static constexpr int nt = NT;
  static constexpr auto fn = k_pathtrace_batch_qr<TV, WV, TW, TH, NT>;
};

hipError_t launch_pathtrace_batch(hipStream_t st, const ViewDesc* views, uint32_t nviews, uint32_t tiles,
                                  unsigned long long* counters, int tv, int waves, uint32_t max_bit_words,
                                  uint32_t max_pwords, int rform, uint32_t rtiles) {
  if (rform != 0) {
    tv = lds_variant(tv, max_bit_words, max_pwords);
    const uint32_t shmem = tv == TV_FASTG ? 0u : lds_bytes(tv, max_bit_words, max_pwords);
    refill_batch_switch<BatchR>(tv, rform, waves, [&](auto k) {
      using K = decltype(k);
      hipLaunchKernelGGL(K::fn, dim3(rtiles > 0 ? rtiles : 1), dim3(K::nt), shmem, st, views, nviews, counters);
    });
    return hipGetLastError();
  }
  tv = lds_variant(tv, max_bit_words, max_pwords);
  const uint32_t shmem = lds_bytes(tv, max_bit_words, max_pwords);
  auto go = [&](auto kernel) {
    hipLaunchKernelGGL(kernel, dim3(tiles), dim3(256), shmem, st, views, nviews, counters);
  };
  switch (tv * 8 + (waves == 4 || waves == 5 ? waves : 1)) {
    case TV_FASTG * 8 + 4: go(k_pathtrace_batch<TV_FASTG, 4>); break;
    case TV_FASTG * 8 + 5: go(k_pathtrace_batch<TV_FASTG, 5>); break;
    case TV_FASTG * 8 + 1: go(k_pathtrace_batch<TV_FASTG, 1>); break;
    case TV_FAST * 8 + 4: go(k_pathtrace_batch<TV_FAST, 4>); break;
    case TV_FAST * 8 + 5: go(k_pathtrace_batch<TV_FAST, 5>); break;
    case TV_FAST * 8 + 1: go(k_pathtrace_batch<TV_FAST, 1>); break;
    case TV_NESTED * 8 + 4: go(k_pathtrace_batch<TV_NESTED, 4>); break;
    case TV_NESTED * 8 + 5: go(k_pathtrace_batch<TV_NESTED, 5>); break;
    default: go(k_pathtrace_batch<TV_NESTED, 1>); break;
  }
  return hipGetLastError();
}

hipError_t launch_batch_persistent(hipStream_t st, BatchHead* head, const ViewDesc* views,
                                   unsigned long long* counters, int tv, int waves, uint32_t max_bit_words,
                                   uint32_t max_pwords, uint32_t pt_blocks, uint32_t gs_blocks, bool denoise,
                                   bool hdr, int rform) {
  tv = lds_variant(tv, max_bit_words, max_pwords);
  const uint32_t shmem = lds_bytes(tv, max_bit_words, max_pwords);
  const dim3 gp(pt_blocks > 0 ? pt_blocks : 1), gg(gs_blocks > 0 ? gs_blocks : 1);
  auto go = [&](auto kernel) { hipLaunchKernelGGL(kernel, gp, dim3(256), shmem, st, head, views, counters); };
  const int wvr = waves == 4 || waves == 5 ? waves : 1;
  if (rform != 0)
    refill_batch_switch<BatchQR>(tv, rform, waves, [&](auto k) {
      using K = decltype(k);
      hipLaunchKernelGGL(K::fn, gp, dim3(K::nt), tv == TV_FASTG ? 0u : shmem, st, head, views, counters);
    });
  else switch (tv * 8 + wvr) {
    case TV_FASTG * 8 + 4: go(k_pathtrace_batch_q<TV_FASTG, 4>); break;
    case TV_FASTG * 8 + 5: go(k_pathtrace_batch_q<TV_FASTG, 5>); break;
    case TV_FASTG * 8 + 1: go(k_pathtrace_batch_q<TV_FASTG, 1>); break;
    case TV_FAST * 8 + 4: go(k_pathtrace_batch_q<TV_FAST, 4>); break;
    case TV_FAST * 8 + 5: go(k_pathtrace_batch_q<TV_FAST, 5>); break;
    case TV_FAST * 8 + 1: go(k_pathtrace_batch_q<TV_FAST, 1>); break;
    case TV_NESTED * 8 + 4: go(k_pathtrace_batch_q<TV_NESTED, 4>); break;
    case TV_NESTED * 8 + 5: go(k_pathtrace_batch_q<TV_NESTED, 5>); break;
    default: go(k_pathtrace_batch_q<TV_NESTED, 1>); break;
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_temporal_batch_gs, gg, dim3(256), 0, st, head, views);
  if (denoise) {
    hipLaunchKernelGGL(k_atrous_batch_gs<0>, gg, dim3(256), 0, st, head, views);
    hipLaunchKernelGGL(k_atrous_batch_gs<1>, gg, dim3(256), 0, st, head, views);
    hipLaunchKernelGGL(k_atrous_batch_gs<2>, gg, dim3(256), 0, st, head, views);
  }
  if (hdr) hipLaunchKernelGGL(k_hdr_batch_gs, gg, dim3(256), 0, st, head, views);
  return hipGetLastError();
}

hipError_t launch_hdr_batch(hipStream_t st, const ViewDesc* views, uint32_t nviews, uint32_t tiles) {
  hipLaunchKernelGGL(k_hdr_batch, dim3(tiles), dim3(256), 0, st, views, nviews);
  return hipGetLastError();
}

hipError_t launch_temporal_batch(hipStream_t st, const ViewDesc* views, uint32_t nviews, uint32_t tiles) {
  hipLaunchKernelGGL(k_temporal_batch, dim3(tiles), dim3(256), 0, st, views, nviews);
  return hipGetLastError();
}

hipError_t launch_atrous_batch(hipStream_t st, const ViewDesc* views, uint32_t nviews, uint32_t tiles, int pass) {
  if (pass == 0) hipLaunchKernelGGL(k_atrous_batch<0>, dim3(tiles), dim3(256), 0, st, views, nviews);
  else if (pass == 1) hipLaunchKernelGGL(k_atrous_batch<1>, dim3(tiles), dim3(256), 0, st, views, nviews);
  else hipLaunchKernelGGL(k_atrous_batch<2>, dim3(tiles), dim3(256), 0, st, views, nviews);
  return hipGetLastError();
}

hipError_t launch_temporal(hipStream_t st, const Uniforms& u, const void* radiance,
                           const void* gcur, const void* aprev, const void* gprev, void* aout,
                           int fmt) {
  const uint32_t w = (uint32_t)u.params1[0], h = (uint32_t)u.params1[1];
  const uint32_t tiles = div_up(w, 16) * div_up(h, 16);
  if (fmt == 1)
    hipLaunchKernelGGL(k_temporal<1>, dim3(tiles), dim3(256), 0, st, u, (const ushort4*)radiance,
                       (const float4*)gcur, aprev, (const float4*)gprev, aout);
  else
    hipLaunchKernelGGL(k_temporal<0>, dim3(tiles), dim3(256), 0, st, u, (const ushort4*)radiance,
                       (const float4*)gcur, aprev, (const float4*)gprev, aout);
  return hipGetLastError();
}

template <int FMT>
static void atrous_fmt(hipStream_t st, const Uniforms& u, const void* src, const void* gbuf,
                       int step, void* dst, uint32_t tiles) {
  switch (step) {
    case 1:
      hipLaunchKernelGGL((k_atrous<FMT, 1>), dim3(tiles), dim3(256), 0, st, u, src,
                         (const float4*)gbuf, (ushort4*)dst);
      break;
    case 2:
      hipLaunchKernelGGL((k_atrous<FMT, 2>), dim3(tiles), dim3(256), 0, st, u, src,
                         (const float4*)gbuf, (ushort4*)dst);
      break;
    default:
      hipLaunchKernelGGL((k_atrous<FMT, 4>), dim3(tiles), dim3(256), 0, st, u, src,
                         (const float4*)gbuf, (ushort4*)dst);
      break;
  }
}

hipError_t launch_atrous(hipStream_t st, const Uniforms& u, const void* src, int src_fmt,
                         const void* gbuf, int step, void* dst) {
  const uint32_t w = (uint32_t)u.params1[0], h = (uint32_t)u.params1[1];
  const uint32_t tiles = div_up(w, 16) * div_up(h, 16);
  if (step != 1 && step != 2 && step != 4) return hipErrorInvalidValue;
  if (src_fmt == 1) atrous_fmt<1>(st, u, src, gbuf, step, dst, tiles);
  else atrous_fmt<0>(st, u, src, gbuf, step, dst, tiles);
  return hipGetLastError();
}

hipError_t launch_present(hipStream_t st, uint32_t npx, const void* illum, int fmt, const void* albedo,
                          float exposure, void* out) {
  if (fmt == 1)
    hipLaunchKernelGGL(k_present<1>, dim3(div_up(npx, 256)), dim3(256), 0, st, npx, illum, (const uchar4*)albedo,
                       exposure, (uchar4*)out);
  else
    hipLaunchKernelGGL(k_present<0>, dim3(div_up(npx, 256)), dim3(256), 0, st, npx, illum, (const uchar4*)albedo,
                       exposure, (uchar4*)out);
  return hipGetLastError();
}

hipError_t launch_hdr(hipStream_t st, uint32_t npx, const void* illum, int fmt, const void* albedo,
                      float* hdr) {
  if (fmt == 1)
    hipLaunchKernelGGL(k_hdr<1>, dim3(div_up(npx, 256)), dim3(256), 0, st, npx, illum,
                       (const uchar4*)albedo, hdr);
  else
    hipLaunchKernelGGL(k_hdr<0>, dim3(div_up(npx, 256)), dim3(256), 0, st, npx, illum,
                       (const uchar4*)albedo, hdr);
  return hipGetLastError();
}

}  // namespace vrt
