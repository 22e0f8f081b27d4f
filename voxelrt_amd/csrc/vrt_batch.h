// vrt_batch.h -- the batched form: many views, of different scenes, sizes,
// bounce counts and history formats, rendered by one launch per pass.
//
// The host writes one ViewDesc per view into a device table.  Every pass
// tiles each view into 16x16 pixel tiles; tile0 is the view's first tile in
// the batch (an exclusive prefix over the views' tile counts).  A block finds
// its view by binary search over tile0 -- the index is uniform across the
// block, so the descriptor is read with scalar loads -- and then runs exactly
// the per-pixel code of the single-view kernels: results are bit-identical
// to rendering each view alone.
#pragma once
#include "vrt_kernels.h"
#include "vrt_trace.h"

namespace vrt {

struct ViewDesc {
  SceneDev s;
  vrt_uniforms u;
  float fclamp;
  uint32_t spp, fmt, tile0;
  uint32_t w, h, tiles_x, denoise;
  uint32_t rtile0;  // first of the view's path-trace tiles of the refill forms
  uint32_t rtx;     // their count across the view
  uint32_t wchunk0; // first of the view's patch chunks of the step-level refill form
  uint32_t pad0;
  void* radiance;    // rgba16f
  void* albedo;      // rgba8unorm
  void* gbuf_cur;    // rgba32f
  void* gbuf_prev;
  void* accum_prev;  // rgba16f | rgba32f
  void* accum_cur;
  void* den0;        // rgba16f
  void* den1;
  unsigned long long* vcount;  // this view's rays / brick steps / voxel steps / hits, or null
  float* hdr;                  // rgb f32 HDR output (main.js:628-659), or null
};

// Header of the device descriptor table, at offset 0; the ViewDesc array
// follows at BATCH_HEAD_BYTES.  `queue` is the shared tile queue of the
// persistent path-trace launch and `wq` the split queue of the step-level
// refill form's chunks; the host uploads both as 0 with every table.
struct BatchHead {
  uint32_t nviews;
  uint32_t tiles;
  uint32_t queue;
  uint32_t rtiles;   // path-trace tiles of the refill forms (32x32, 32x16, 16x16 or 16x8)
  uint32_t wchunks;  // patch chunks of the step-level refill form
  uint32_t wchunk;   // 8x8 patches per chunk, 1..WAVE_CHUNK
  uint32_t pad[10];
  uint32_t wq[QUEUE_WORDS];
};
constexpr uint32_t BATCH_HEAD_BYTES = 64 + QUEUE_WORDS * 4;
static_assert(sizeof(BatchHead) == BATCH_HEAD_BYTES, "BatchHead layout");

// Index of the view owning tile `b` (views sorted by tile0, tile0[0] = 0).
VRT_D uint32_t find_view(const ViewDesc* __restrict__ views, uint32_t nviews, uint32_t b) {
  uint32_t lo = 0, hi = nviews;  // invariant: views[lo].tile0 <= b < views[hi].tile0
  while (hi - lo > 1u) {
    const uint32_t mid = (lo + hi) >> 1;
    if (views[mid].tile0 <= b) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Index of the view owning refill tile `b` (views sorted by rtile0).
VRT_D uint32_t find_view_r(const ViewDesc* __restrict__ views, uint32_t nviews, uint32_t b) {
  uint32_t lo = 0, hi = nviews;
  while (hi - lo > 1u) {
    const uint32_t mid = (lo + hi) >> 1;
    if (views[mid].rtile0 <= b) lo = mid;
    else hi = mid;
  }
  return lo;
}

}  // namespace vrt
