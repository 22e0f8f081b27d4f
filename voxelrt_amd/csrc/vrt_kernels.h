// vrt_kernels.h -- launchers of the gfx950 kernels (vrt_kernels.hip), called
// by the runtime (vrt_runtime.cpp).  All launches go on the given stream.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/voxelrt.h"

namespace vrt {

using Uniforms = vrt_uniforms;
struct SceneDev;

// Trace variant ids for the one-ray/pixel-per-lane kernels (launch_* `tv`).
// TV_NESTED: nested brick / voxel loops on global memory (vrt_trace.h).
// TV_FAST: the single-loop DDA with both levels in LDS (vrt_dda_fast.h);
// TV_FASTSKIP: the same DDA taking runs of empty brick rows in bulk
// (traversal kernels only); TV_FASTG: the single-loop DDA on global memory
// (no LDS), the form for scenes whose bitmaps do not fit in LDS.
constexpr int TV_NESTED = 0, TV_FAST = 3, TV_FASTSKIP = 5, TV_FASTG = 6;
constexpr bool tv_fast(int tv) { return tv == TV_FAST || tv == TV_FASTSKIP; }
// LDS words the single-loop variants may stage beyond the mask slots: the
// padded bitmap, its prefix counts and the row distances within 48 KiB.
// Larger scenes run TV_FASTG.
constexpr uint32_t LDS_BIT_WORDS_MAX = 12288;
// u32 words of a persistent kernel's work queue (the split queue's eight
// counters, one per 128-byte line; the single-counter kernels use word 0).
constexpr uint32_t QUEUE_WORDS = 256;
// Most 8x8 patches per work item of the batched step-level refill path
// tracer (a wave finishes its paths before it moves to another view's chunk,
// so large batches take large chunks).  A launch takes
// wave_chunk(patches, waves, nviews) patches per item: single patches for one
// view (no wave ever switches views, and single items balance the waves'
// last draws best), else as many as give every wave eight or more items, so
// that the last items do not leave most waves idle (with two per wave, one
// 1440p view took 48 % longer than the single-view kernel).
constexpr uint32_t WAVE_CHUNK = 8;
inline uint32_t wave_chunk(uint64_t patches, uint32_t waves, uint32_t nviews) {
  if (nviews <= 1u) return 1u;
  const uint64_t c = patches / (8ull * (waves ? waves : 1u));
  return c < 1u ? 1u : (c > WAVE_CHUNK ? WAVE_CHUNK : (uint32_t)c);
}

hipError_t launch_brick_masks(hipStream_t st, const uint32_t* vox, int nx, int ny, int bgx,
                              int bgy, int bz0, int nbz, uint32_t* masks, uint32_t* flags);
// An edit's touched bricks [b0, b0 + e) per axis from the dense grid vox:
// pool entries rewritten in place where occupancy is unchanged, *changed set
// (atomicOr 1) where it changed (the caller rebuilds).
hipError_t launch_edit_bricks(hipStream_t st, const uint32_t* vox, int nx, int ny, int bgx, int bgy, int bx0,
                              int by0, int bz0, int ex, int ey, int ez, const uint32_t* bricks, uint32_t* pool,
                              uint32_t* changed);
hipError_t launch_brick_words(hipStream_t st, const uint32_t* flags, const uint32_t* slots,
                              uint64_t nb, uint32_t* bricks);
hipError_t launch_brick_bits(hipStream_t st, const uint32_t* bricks, uint64_t nb, uint32_t* bits);
hipError_t launch_brick_pbits(hipStream_t st, const uint32_t* bricks, int bgx, int bgy, int bgz, uint32_t sh1,
                              uint32_t sh2, uint64_t nbits, uint32_t* bits);
hipError_t launch_popc(hipStream_t st, const uint32_t* words, uint64_t n, uint32_t* counts);
// Row distances of a scene (SceneDev::pdist) from its brick words; tmp holds
// 3 * bgy * bgz bytes, pdist (pwords + 3) & ~3 bytes.
// Per pool slot [0, nslots): the packed box of the brick's solid voxels
// (SceneDev::vbox), from the entry's mask words.
hipError_t launch_brick_vbox(hipStream_t st, const uint32_t* pool, uint64_t nslots, uint32_t* vbox);
// Per brick row (by, bz), r = by + bz * bgy: first | last << 16 of its
// occupied bx, 0xFFFF when empty (bgy * bgz words).
hipError_t launch_row_span(hipStream_t st, const uint32_t* bricks, int bgx, int bgy, int bgz, uint32_t* span);
hipError_t launch_row_dist(hipStream_t st, const uint32_t* bricks, int bgx, int bgy, int bgz, uint32_t sh1,
                           uint32_t sh2, uint64_t pwords, uint8_t* tmp, uint8_t* pdist);
hipError_t launch_pack_pool(hipStream_t st, const uint32_t* vox, int nx, int ny, int bgx, int bgy,
                            int bz0, int nbz, const uint32_t* masks, const uint32_t* bricks,
                            uint32_t* pool);
// waves: 6 = the single-loop kernels built for >= 6 waves per SIMD, else the
// compiler's register allocation
// Zero n 32-bit words at p on st (a kernel, not a memset node).
hipError_t launch_zero_words(hipStream_t st, void* p, uint32_t n);
hipError_t launch_trace_rays(hipStream_t st, const SceneDev& s, const float* rays, uint64_t n,
                             float tmax, void* hits, unsigned long long* counters, int blocks, int tv,
                             int waves);
hipError_t launch_trace_persistent(hipStream_t st, const SceneDev& s, const float* rays, uint32_t n,
                                  float tmax, void* hits, unsigned long long* counters,
                                  uint32_t* queue, int blocks, int steps, int min_refill, int tv, int waves,
                                  bool clear_queue = true);
// Step-level refill traversal (vrt_trace_wave.h): `blocks` blocks of 256
// threads (LDS form) or 4 x blocks one-wave blocks (global-memory form, for
// scenes whose bitmaps do not fit in LDS); a pass when `min_wait` lanes wait
// at a brick boundary or `fin_min` have finished; waves 6 = built for >= 6
// waves per SIMD.  queue: the chunk queue's 256 u32 words (QWORDS), cleared
// here by a kernel when clear_queue is set (else the caller has zeroed them).
hipError_t launch_trace_wave(hipStream_t st, const SceneDev& s, const float* rays, uint32_t n, float tmax,
                             void* hits, unsigned long long* counters, uint32_t* queue, int blocks, int min_wait,
                             int fin_min, int waves, bool clear_queue = true);
hipError_t launch_pathtrace(hipStream_t st, const SceneDev& s, const Uniforms& u, float fclamp,
                            uint32_t spp, void* radiance, void* albedo, void* gbuf,
                            unsigned long long* counters, int tv, int waves, int form);
// Step-level refill path tracer on persistent one-wave blocks
// (vrt_pathtrace_wave.h): `blocks` blocks of 64 threads; a pass when
// `min_wait` lanes of a wave wait at a brick boundary or `shade_min` have
// finished rays; waves = 0 / 4 / 5 / 6 minimum waves per SIMD; form 1 = the
// global-memory DDA even where the LDS form fits; queue: the patch queue's
// 256 u32 words (QWORDS), cleared here by a kernel.
hipError_t launch_pathtrace_wave(hipStream_t st, const SceneDev& s, const Uniforms& u, float fclamp, uint32_t spp,
                                 void* radiance, void* albedo, void* gbuf, unsigned long long* counters,
                                 uint32_t* queue, int blocks, int min_wait, int shade_min, int waves, int form);
struct ViewDesc;
// The refill tile form (2: 32x32, 3: 32x16 on 256-thread blocks; 4: 16x16,
// 5: 16x8 on one-wave blocks) the path tracer runs for variant tv and tuning
// form `form` over scenes of up to the given bitmap sizes, or 0 for the
// 16x16 one-pixel-per-lane tiles.  refill_tile_dims gives a form's tile size.
int refill_form(int tv, int form, uint32_t max_bit_words, uint32_t max_pwords);
void refill_tile_dims(int rform, uint32_t* tw, uint32_t* th);
// With rform != 0 (from refill_form) the grid is `rtiles` blocks over the
// views' refill tiles (ViewDesc::rtile0 / rtx), else `tiles` blocks over
// their 16x16 tiles.
hipError_t launch_pathtrace_batch(hipStream_t st, const ViewDesc* views, uint32_t nviews, uint32_t tiles,
                                  unsigned long long* counters, int tv, int waves, uint32_t max_bit_words,
                                  uint32_t max_pwords, int rform, uint32_t rtiles);
hipError_t launch_temporal_batch(hipStream_t st, const ViewDesc* views, uint32_t nviews, uint32_t tiles);
hipError_t launch_hdr_batch(hipStream_t st, const ViewDesc* views, uint32_t nviews, uint32_t tiles);
struct BatchHead;
// Persistent batched form: the path tracer on a fixed grid of pt_blocks pulling
// tiles from head->queue (with rform != 0 the refill tiles of that form,
// head->rtiles of them), then temporal / a-trous x3 (denoise) / hdr on a
// fixed grid-stride grid of gs_blocks.  No argument depends on the batch, so
// the sequence can be captured once into a graph and replayed.
hipError_t launch_batch_persistent(hipStream_t st, BatchHead* head, const ViewDesc* views,
                                   unsigned long long* counters, int tv, int waves, uint32_t max_bit_words,
                                   uint32_t max_pwords, uint32_t pt_blocks, uint32_t gs_blocks, bool denoise,
                                   bool hdr, int rform);
hipError_t launch_atrous_batch(hipStream_t st, const ViewDesc* views, uint32_t nviews, uint32_t tiles, int pass);
// The step-level refill path tracer over a batch table (vrt_pathtrace_wave.h
// k_pathtrace_wave_batch): `blocks` one-wave blocks drawing chunks of the
// views' 8x8 patches from head->wq (zero at launch; the table upload zeroes
// it); then the filter passes on a fixed grid-stride grid of gs_blocks.
hipError_t launch_pathtrace_wave_batch(hipStream_t st, BatchHead* head, const ViewDesc* views,
                                       unsigned long long* counters, int blocks, int min_wait, int shade_min,
                                       int waves);
hipError_t launch_batch_filters(hipStream_t st, BatchHead* head, const ViewDesc* views, uint32_t gs_blocks,
                                bool denoise, bool hdr);
hipError_t launch_temporal(hipStream_t st, const Uniforms& u, const void* radiance,
                           const void* gcur, const void* aprev, const void* gprev, void* aout,
                           int fmt);
hipError_t launch_atrous(hipStream_t st, const Uniforms& u, const void* src, int src_fmt,
                         const void* gbuf, int step, void* dst);
hipError_t launch_present(hipStream_t st, uint32_t npx, const void* illum, int fmt, const void* albedo,
                          float exposure, void* out);
hipError_t launch_hdr(hipStream_t st, uint32_t npx, const void* illum, int fmt, const void* albedo,
                      float* hdr);

}  // namespace vrt
