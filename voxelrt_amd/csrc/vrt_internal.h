// vrt_internal.h -- host-side state of libvoxelrt_hip (not part of the ABI).
#pragma once
#include <array>
#include <hip/hip_runtime.h>

#include <chrono>
#include <map>
#include <utility>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/voxelrt.h"
#include "vrt_trace.h"

namespace vrt {

// Owning device allocation (hipMalloc / hipFree).
struct DeviceBuffer {
  void* ptr = nullptr;
  size_t bytes = 0;
  DeviceBuffer() = default;
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  DeviceBuffer(DeviceBuffer&& o) noexcept : ptr(o.ptr), bytes(o.bytes) {
    o.ptr = nullptr;
    o.bytes = 0;
  }
  DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
    if (this != &o) {
      release();
      ptr = o.ptr;
      bytes = o.bytes;
      o.ptr = nullptr;
      o.bytes = 0;
    }
    return *this;
  }
  ~DeviceBuffer() { release(); }
  hipError_t alloc(size_t n);
  hipError_t ensure(size_t n);  // grow-only
  void release();
};

// One stream's batched-form launch state: a pinned staging ring for the
// view-descriptor table (vrt_batch.h) and its device copy.  A staging buffer
// is rewritten only after its previous copy has executed (event per buffer),
// so the host can run ahead of the device by up to STAGES batches.
struct BatchLane {
  static constexpr uint32_t STAGES = 4;
  hipStream_t stream = nullptr;  // owned when own_stream
  bool own_stream = false;
  void* desc_host[STAGES] = {};
  hipEvent_t desc_done[STAGES] = {};
  uint32_t desc_cap = 0;  // views per staging buffer
  uint32_t desc_next = 0;
  DeviceBuffer desc_dev;  // BatchHead, then the ViewDesc array (vrt_batch.h)
  // instantiated graphs of the persistent batched passes, by launch shape
  std::map<std::pair<uint64_t, uint64_t>, hipGraphExec_t> graphs;
  BatchLane() = default;
  BatchLane(const BatchLane&) = delete;
  BatchLane& operator=(const BatchLane&) = delete;
  ~BatchLane();
  hipError_t ensure(uint32_t nviews);
  void drop_graphs();
};

struct Server;  // serving loop (vrt_serve.cpp)

}  // namespace vrt

struct vrt_scene {
  uint32_t nx = 0, ny = 0, nz = 0;
  uint64_t nbricks = 0, nocc = 0;
  uint64_t bytes = 0;
  uint64_t key = 0;
  uint64_t lru = 0;
  uint64_t uses = 0;  // queued + running serve requests
  uint64_t edits_in_place = 0, edits_rebuilt = 0;  // vrt_scene_edit calls by how they were applied
  uint64_t pins = 0;       // C ABI calls using the scene right now (not evictable meanwhile)
  bool evictable = false;  // VRT_SCENE_EVICTABLE
  bool evicted = false;    // device buffers moved to `host` (budget released)
  std::vector<std::vector<uint8_t>> host;  // an evicted scene's buffers, in buffers() order
  vrt::DeviceBuffer d_bricks;  // one word per 8^3 brick: 0 or 1 + pool slot
  vrt::DeviceBuffer d_pool;    // nocc x 528 words (mask + voxels)
  vrt::DeviceBuffer d_bits;    // one occupancy bit per brick (LDS-staged by the kernels)
  vrt::DeviceBuffer d_prefix;  // per bitmap word: occupied bricks before it
  vrt::DeviceBuffer d_pbits, d_pprefix;  // the same over the power-of-two padded brick grid
  vrt::DeviceBuffer d_pdist;             // row distance per padded bitmap word (bytes)
  vrt::DeviceBuffer d_dense;  // editable scenes: the dense packed grid, rebuilt from on edits
  vrt::DeviceBuffer d_vbox;   // per pool slot: the box of the brick's solid voxels (SceneDev::vbox)
  uint32_t psh1 = 0, psh2 = 0, psh3 = 0;  // padded bit index = bx | by << psh1 | bz << psh2; 2^psh3 bits
  int box_lo[3] = {0, 0, 0}, box_hi[3] = {-1, -1, -1};  // occupied bricks' box (scene_build)
  vrt::SceneDev dev() const;
  // every device buffer of the scene (those of size 0 are not allocated)
  std::array<vrt::DeviceBuffer*, 9> buffers() {
    return {&d_bricks, &d_pool, &d_bits, &d_prefix, &d_pbits, &d_pprefix, &d_pdist, &d_dense, &d_vbox};
  }
};

// Per-view resident frame targets (main.js:341-349): radiance rgba16f,
// albedo rgba8unorm, gbuf x2 rgba32f, accum x2 (rgba16f | rgba32f),
// denoise x2 rgba16f.
struct ViewState {
  uint32_t w = 0, h = 0, fmt = 0;
  uint64_t bytes = 0;
  uint64_t frames = 0;  // frames rendered: parity = frames & 1 (main.js:695)
  uint64_t lru = 0;     // context clock at last use
  vrt::DeviceBuffer radiance, albedo;
  vrt::DeviceBuffer gbuf[2], accum[2], den[2];
};

struct vrt_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[8] = {};
  std::string err;
  std::string err_ret;  // vrt_last_error's copy
  uint64_t budget = 0, resident = 0;
  uint64_t clock = 0;
  uint64_t frames = 0;
  double ms_pathtrace = 0, ms_temporal = 0, ms_atrous = 0, ms_frame = 0;  // last render call
  // per-render-call timing events (pathtrace start, temporal start, a-trous
  // start, end), read back lazily by harvest()
  static constexpr uint32_t RING = 64;
  struct TimingSlot {
    hipEvent_t e[4] = {};
  };
  TimingSlot ring[RING];
  uint64_t ring_head = 0, ring_count = 0;
  uint64_t timed_calls = 0;
  double sum_pathtrace = 0, sum_temporal = 0, sum_atrous = 0, sum_frame = 0;
  int harvest(uint32_t max_slots);  // oldest pending first; waits for each
  // batched form of vrt_render, on the context's stream
  vrt::BatchLane lane0;
  // Enqueue the batched passes for nviews views on `lane`: pathtrace ->
  // temporal -> a-trous x3 (views with denoise) -> hdr (views whose hdr
  // pointer is set).  ev[0..3]: pass boundaries.  vcount/hdr: per-view
  // device counters (4 x u64) and HDR outputs, either array may be null.
  int render_batch(vrt::BatchLane& lane, const vrt_view* views, uint32_t nviews, ViewState** states,
                   hipEvent_t* ev, unsigned long long* const* vcount, float* const* hdr);
  // serialises the C ABI with the serving worker thread
  std::mutex mu;
  std::unique_ptr<vrt::Server> server;
  // make the context's stream wait for every running serve batch
  int order_after_serve();
  int trace_blocks = 0;  // 0 = default grid for k_trace_rays
  int cus = 256;         // compute units of the device (persistent grids)
  vrt_tuning tuning{};   // kernel variants + launch shapes (vrt_tuning_set / vrt_autotune)
  vrt::DeviceBuffer d_counters;  // u64 x4: rays, brick steps, voxel steps, hits
  vrt::DeviceBuffer d_queue;     // work queues of the persistent kernels: traversal (words 0..255), path
                                 // tracer (256..511); word 512: the scene edit's occupancy-changed flag
  vrt::DeviceBuffer scratch_rays, scratch_hits, scratch_hdr, scratch_ldr;
  std::vector<vrt_scene*> scenes;
  std::unordered_map<uint32_t, ViewState> views;

  ~vrt_ctx();
  int fail(int code, const char* fmt, ...);
  // budget: reserve evicts idle evictable scenes (least recently used first)
  // when the bytes do not fit otherwise
  int reserve(uint64_t bytes);
  // scene residency (VRT_SCENE_EVICTABLE): evict = device buffers to host
  // memory and the budget released; restore = back to the device;
  // make_scene_room evicts idle evictable scenes (no serve request queued or
  // running, no C ABI call using them, not `keep`) least recently used first
  // until `need` more bytes fit, and says whether they do
  int evict_scene(vrt_scene* s);
  int restore_scene(vrt_scene* s);
  bool make_scene_room(uint64_t need, const vrt_scene* keep);
  uint64_t scene_evictions = 0, scene_restores = 0;
  void unreserve(uint64_t bytes);
  int counters_to_host(uint64_t out[4]);
  int validate_view(const vrt_view& v);
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  double now_ms() const {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  // device buffers of a view state (budget reserved; history zeroed), not
  // entered in `views`
  int alloc_view_state(ViewState& v, uint32_t w, uint32_t h, uint32_t fmt, hipStream_t st);
  int view_state(uint32_t id, uint32_t w, uint32_t h, uint32_t fmt, ViewState** out,
                 hipStream_t st = nullptr);  // st: stream for the initial clears (default: the context's)
  // A scene's device view with this context's launch-time tuning applied.
  vrt::SceneDev tuned(vrt::SceneDev s) const {
    s.entry_batch = tuning.dda_batch ? tuning.dda_batch : 1u;
    s.tile_order = tuning.pt_order;
    return s;
  }
  // counters / queue: where the kernels count and keep their ray-queue head
  // (null = the context's own); clear_queue: reset the head with a memset
  // (the allocation-free entry point clears its scratch with a kernel)
  hipError_t launch_trace(const vrt::SceneDev& s, const float* rays, uint64_t n, float tmax,
                          void* hits, unsigned long long* counters = nullptr, uint32_t* queue = nullptr,
                          bool clear_queue = true);
  // device address of the allocation-free entry point's scratch on this
  // context's device (set by vrt_create)
  void* noalloc_scratch = nullptr;
  // device buffers handed out by vrt_device_alloc (bytes, counted in the budget)
  std::unordered_map<void*, uint64_t> user_buffers;
  hipError_t launch_pathtrace(const vrt::SceneDev& s, const vrt_uniforms& u, float fclamp,
                              uint32_t spp, void* radiance, void* albedo, void* gbuf);
};
