// vrt_serve.h -- the serving loop behind vrt_submit / vrt_poll (continuous
// batching).  Host-side only; not part of the ABI.
//
// One worker thread per context.  Requests wait in a FIFO.  Whenever one of
// the serve streams ("lanes") has no batch on the device, the worker takes
// queued requests in order -- at most one frame per view, at most
// max_batch_views requests and max_batch_pixels pixel-samples -- admits each
// by HBM footprint (evicting pooled result buffers, idle view state and idle
// evictable scenes, in that order, when
// needed), and enqueues the batched passes for them on that lane.  Lanes run
// concurrently on the device, so a request that arrives while one batch runs
// starts in the next free lane without waiting for the running batch.  The
// worker polls the lanes' completion events, stamps the tickets and wakes
// vrt_wait.  All state below is guarded by vrt_ctx::mu.
#pragma once
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/voxelrt.h"
#include "vrt_internal.h"

namespace vrt {

struct ServeTicket {
  enum State { QUEUED, RUNNING, DONE, FAILED };
  State state = QUEUED;
  vrt_request req{};
  int error = VRT_OK;
  std::string msg;
  void* result = nullptr;     // device HDR (w*h*3 f32), from the result pool
  uint64_t result_bytes = 0;  // size of that pooled buffer
  uint64_t counts[4] = {0, 0, 0, 0};
  vrt_timing timing{};
};

// Per-view bookkeeping of the serving loop (view state itself lives in
// vrt_ctx::views).
struct ViewUse {
  uint32_t pending = 0;  // queued + running requests
  int lane = -1;         // lane running one of its frames, -1 = none
};

struct ServeLane {
  BatchLane lane;
  hipEvent_t ev[4] = {};            // pass boundaries of the running batch
  hipEvent_t ev_done = nullptr;     // after hdr + counter read-back
  hipEvent_t main_ready = nullptr;  // context stream -> lane ordering
  DeviceBuffer d_vcount;            // 4 x u64 per request
  unsigned long long* h_vcount = nullptr;  // pinned read-back of d_vcount
  bool running = false;
  uint64_t batch_id = 0;
  double launch_ms = 0;
  std::vector<uint64_t> tickets;
  ServeLane() = default;
  ServeLane(const ServeLane&) = delete;
  ServeLane& operator=(const ServeLane&) = delete;
  ~ServeLane();
};

struct Server {
  vrt_ctx* ctx = nullptr;
  vrt_serve_config cfg{};
  std::thread worker;
  std::condition_variable cv_work;  // submit / cancel / stop -> worker
  std::condition_variable cv_done;  // worker -> vrt_wait
  bool stop = false;
  bool started = false;
  uint64_t next_ticket = 1;
  uint64_t next_batch = 1;
  std::deque<uint64_t> queue;
  std::unordered_map<uint64_t, ServeTicket> tickets;
  std::unordered_map<uint32_t, ViewUse> vuse;
  std::vector<std::unique_ptr<ServeLane>> lanes;
  std::multimap<uint64_t, void*> pool;  // free result buffers by size
  hipStream_t copy_stream = nullptr;    // vrt_poll read-backs
  vrt_serve_stats st{};

  explicit Server(vrt_ctx* c);
  ~Server();
  int start();  // lanes, streams, worker; called with ctx->mu held
  void run();   // worker body
  bool harvest();
  bool launch(uint32_t li);
  bool make_room(uint64_t need);
  int acquire_result(uint64_t bytes, void** ptr, uint64_t* got);
  void release_result(void* ptr, uint64_t bytes);
  void fail_ticket(uint64_t id, int code, const std::string& msg);
  void retire(uint64_t id);  // drop the ticket's scene / view references
  uint64_t running_lanes() const;
};

}  // namespace vrt
