// vrt_serve.cpp -- serving loop (continuous batching) and its C ABI:
// vrt_serve_configure, vrt_submit, vrt_poll, vrt_wait, vrt_cancel,
// vrt_serve_stats_get.  See vrt_serve.h for the scheduling rules.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <unordered_set>

#include "../../include/voxelrt.h"
#include "vrt_internal.h"
#include "vrt_serve.h"

namespace vrt {

namespace {

constexpr uint64_t RESULT_ALIGN = 64 * 1024;  // result buffers are rounded up to this

uint64_t view_state_bytes(uint32_t w, uint32_t h, uint32_t fmt) {
  const uint64_t px = (uint64_t)w * h;
  return px * (8 + 4 + 16 * 2 + (fmt == VRT_ACCUM_RGBA32F ? 16 : 8) * 2 + 8 * 2);
}

uint32_t view_w(const vrt_view& v) { return (uint32_t)v.u.params1[0]; }
uint32_t view_h(const vrt_view& v) { return (uint32_t)v.u.params1[1]; }

}  // namespace

ServeLane::~ServeLane() {
  for (auto& e : ev)
    if (e) (void)hipEventDestroy(e);
  if (ev_done) (void)hipEventDestroy(ev_done);
  if (main_ready) (void)hipEventDestroy(main_ready);
  if (h_vcount) (void)hipHostFree(h_vcount);
}

// Defaults: four streams and small batches keep a request's wait behind a
// running batch short (measured on 16 scenes 64^3..1024^3 at 0.8 x plain
// capacity: the same saturated rate as two streams with 64-view batches, and
// p99 at the heaviest requests' own service time instead of above it).
Server::Server(vrt_ctx* c) : ctx(c) {
  cfg.streams = 4;
  cfg.max_batch_views = 8;
  cfg.max_batch_pixels = 4ull << 20;
}

Server::~Server() {
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    stop = true;
  }
  cv_work.notify_all();
  if (worker.joinable()) worker.join();
  (void)hipSetDevice(ctx->device);
  for (auto& l : lanes)
    if (l->lane.stream) (void)hipStreamSynchronize(l->lane.stream);
  for (auto& kv : tickets)
    if (kv.second.result) {
      (void)hipFree(kv.second.result);
      ctx->unreserve(kv.second.result_bytes);
    }
  for (auto& kv : pool) {
    (void)hipFree(kv.second);
    ctx->unreserve(kv.first);
  }
  pool.clear();
  lanes.clear();
  if (copy_stream) (void)hipStreamDestroy(copy_stream);
}

int Server::start() {
  if (started) return VRT_OK;
  for (uint32_t i = 0; i < cfg.streams; ++i) {
    auto l = std::make_unique<ServeLane>();
    hipError_t e = hipStreamCreateWithFlags(&l->lane.stream, hipStreamNonBlocking);
    if (e == hipSuccess) l->lane.own_stream = true;
    for (int k = 0; k < 4 && e == hipSuccess; ++k) e = hipEventCreate(&l->ev[k]);
    if (e == hipSuccess) e = hipEventCreate(&l->ev_done);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&l->main_ready, hipEventDisableTiming);
    if (e == hipSuccess) e = l->d_vcount.alloc((uint64_t)cfg.max_batch_views * 4 * sizeof(unsigned long long));
    if (e == hipSuccess)
      e = hipHostMalloc((void**)&l->h_vcount, (uint64_t)cfg.max_batch_views * 4 * sizeof(unsigned long long),
                        hipHostMallocDefault);
    if (e == hipSuccess) e = l->lane.ensure(cfg.max_batch_views);
    if (e != hipSuccess) {
      lanes.clear();
      return ctx->fail(VRT_ERR_DEVICE, "serve lane setup failed: %s", hipGetErrorString(e));
    }
    lanes.push_back(std::move(l));
  }
  hipError_t e = hipStreamCreateWithFlags(&copy_stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    lanes.clear();
    return ctx->fail(VRT_ERR_DEVICE, "serve copy stream: %s", hipGetErrorString(e));
  }
  try {
    worker = std::thread([this] { run(); });
  } catch (...) {
    lanes.clear();
    return ctx->fail(VRT_ERR_DEVICE, "cannot start the serving worker thread");
  }
  started = true;
  return VRT_OK;
}

uint64_t Server::running_lanes() const {
  uint64_t n = 0;
  for (const auto& l : lanes) n += l->running ? 1u : 0u;
  return n;
}

void Server::retire(uint64_t id) {
  auto it = tickets.find(id);
  if (it == tickets.end()) return;
  const vrt_view& v = it->second.req.view;
  vrt_scene* sc = const_cast<vrt_scene*>(v.scene);
  if (sc && sc->uses > 0) sc->uses--;
  auto u = vuse.find(v.view_id);
  if (u != vuse.end() && u->second.pending > 0 && --u->second.pending == 0 && u->second.lane < 0) vuse.erase(u);
}

void Server::fail_ticket(uint64_t id, int code, const std::string& msg) {
  auto it = tickets.find(id);
  if (it == tickets.end()) return;
  retire(id);
  ServeTicket& t = it->second;
  t.state = ServeTicket::FAILED;
  t.error = code;
  t.msg = msg;
  t.timing.done_ms = ctx->now_ms();
  t.timing.latency_ms = t.timing.done_ms - t.timing.submit_ms;
  st.failed++;
}

void Server::release_result(void* ptr, uint64_t bytes) {
  if (ptr) pool.emplace(bytes, ptr);
}

int Server::acquire_result(uint64_t bytes, void** ptr, uint64_t* got) {
  auto it = pool.lower_bound(bytes);
  if (it != pool.end() && it->first <= 2 * bytes) {
    *ptr = it->second;
    *got = it->first;
    pool.erase(it);
    return VRT_OK;
  }
  int rc = ctx->reserve(bytes);
  if (rc) return rc;
  hipError_t e = hipMalloc(ptr, bytes);
  if (e != hipSuccess) {
    ctx->unreserve(bytes);
    *ptr = nullptr;
    return ctx->fail(VRT_ERR_NO_MEMORY, "result buffer: %s", hipGetErrorString(e));
  }
  *got = bytes;
  return VRT_OK;
}

// Free pooled result buffers, then idle view state (no queued or running
// request, least recently used first), until `need` more bytes fit.
bool Server::make_room(uint64_t need) {
  while (ctx->resident + need > ctx->budget) {
    if (!pool.empty()) {
      auto it = std::prev(pool.end());
      (void)hipFree(it->second);
      ctx->unreserve(it->first);
      pool.erase(it);
      continue;
    }
    uint32_t victim = 0;
    uint64_t best = UINT64_MAX;
    bool found = false;
    for (auto& kv : ctx->views) {
      if (vuse.count(kv.first)) continue;  // queued or running
      if (kv.second.lru < best) {
        best = kv.second.lru;
        victim = kv.first;
        found = true;
      }
    }
    if (!found) return ctx->make_scene_room(need, nullptr);  // last: idle evictable scenes
    auto it = ctx->views.find(victim);
    ctx->unreserve(it->second.bytes);
    ctx->views.erase(it);
    st.evictions++;
  }
  return true;
}

// Build and enqueue one batch on lane li from the queue.  Returns true when
// something changed (a batch launched or a request failed).
bool Server::launch(uint32_t li) {
  ServeLane& L = *lanes[li];
  std::vector<uint64_t> pick;
  std::unordered_set<uint32_t> in_batch, blocked;
  uint64_t cost_sum = 0;
  bool changed = false;
  for (auto it = queue.begin(); it != queue.end() && pick.size() < cfg.max_batch_views;) {
    const uint64_t id = *it;
    ServeTicket& t = tickets.at(id);
    const vrt_view& v = t.req.view;
    const uint32_t vid = v.view_id;
    // one frame of a view per batch; a view running elsewhere waits for that
    // lane (its history must be complete), and its later frames queue behind
    if (blocked.count(vid) || in_batch.count(vid) || vuse[vid].lane >= 0) {
      blocked.insert(vid);
      ++it;
      continue;
    }
    const uint32_t w = view_w(v), h = view_h(v);
    const uint64_t px = (uint64_t)w * h;
    const uint64_t cost = px * v.spp;
    if (!pick.empty() && cost_sum + cost > cfg.max_batch_pixels) break;
    // HBM footprint: view state when it is not resident in this shape, plus
    // the result buffer unless a pooled one fits
    uint64_t need = 0;
    auto vs = ctx->views.find(vid);
    const bool reuse_view = vs != ctx->views.end() && vs->second.w == w && vs->second.h == h &&
                            vs->second.fmt == v.accum_format;
    if (!reuse_view) {
      need += view_state_bytes(w, h, v.accum_format);
      if (vs != ctx->views.end()) {  // reshaped: the old state goes first
        ctx->unreserve(vs->second.bytes);
        ctx->views.erase(vs);
      }
    }
    const uint64_t rbytes = (px * 12 + RESULT_ALIGN - 1) / RESULT_ALIGN * RESULT_ALIGN;
    // a pooled result buffer that fits is taken out of the pool first, so
    // that make_room (which frees pooled buffers) cannot free it
    void* pooled = nullptr;
    uint64_t pooled_bytes = 0;
    {
      auto pit = pool.lower_bound(rbytes);
      if (pit != pool.end() && pit->first <= 2 * rbytes) {
        pooled = pit->second;
        pooled_bytes = pit->first;
        pool.erase(pit);
      } else {
        need += rbytes;
      }
    }
    // an evicted scene comes back first (it has queued requests, so nothing
    // evicts it again before the launch)
    vrt_scene* scn = const_cast<vrt_scene*>(v.scene);
    bool room = make_room(need + (scn->evicted ? scn->bytes : 0));
    ViewState* state = nullptr;
    int rc = VRT_OK;
    if (room && scn->evicted) rc = ctx->restore_scene(scn);
    if (room && rc == VRT_OK) {
      rc = ctx->view_state(vid, w, h, v.accum_format, &state, L.lane.stream);
      if (rc == VRT_OK && pooled) {
        t.result = pooled;
        t.result_bytes = pooled_bytes;
        pooled = nullptr;
      } else if (rc == VRT_OK) {
        rc = acquire_result(rbytes, &t.result, &t.result_bytes);
      }
    }
    // out of budget after all (e.g. a reserve elsewhere in between): the
    // same as no room -- wait for running batches, or fail below
    if (rc == VRT_ERR_BUDGET) room = false;
    if (pooled) release_result(pooled, pooled_bytes);  // not used: back to the pool
    if (!room) {
      // nothing left to evict: wait for running batches to finish, or fail
      // the request if nothing can ever free the space
      if (pick.empty() && running_lanes() == 0) {
        const uint64_t held = ctx->resident;
        const uint64_t bud = ctx->budget;
        it = queue.erase(it);
        char msg[256];
        snprintf(msg, sizeof msg,
                 "HBM budget: request needs %llu B, %llu of %llu B held by scenes and unretrieved results",
                 (unsigned long long)need, (unsigned long long)held, (unsigned long long)bud);
        fail_ticket(id, VRT_ERR_BUDGET, msg);
        changed = true;
        continue;
      }
      blocked.insert(vid);
      ++it;
      continue;
    }
    if (rc != VRT_OK) {
      it = queue.erase(it);
      fail_ticket(id, rc, ctx->err);
      changed = true;
      continue;
    }
    pick.push_back(id);
    in_batch.insert(vid);
    cost_sum += cost;
    it = queue.erase(it);
  }
  if (pick.empty()) return changed;

  const uint32_t n = (uint32_t)pick.size();
  std::vector<vrt_view> views(n);
  std::vector<ViewState*> states(n);
  std::vector<unsigned long long*> vcount(n);
  std::vector<float*> hdr(n);
  const double now = ctx->now_ms();
  for (uint32_t k = 0; k < n; ++k) {
    ServeTicket& t = tickets.at(pick[k]);
    views[k] = t.req.view;
    ViewState& vs = ctx->views.at(views[k].view_id);
    vs.lru = ++ctx->clock;
    states[k] = &vs;
    vcount[k] = (unsigned long long*)L.d_vcount.ptr + 4 * k;
    hdr[k] = (float*)t.result;
  }
  auto fail_all = [&](const char* what) {
    for (uint64_t id : pick) fail_ticket(id, VRT_ERR_DEVICE, std::string(what) + ": " + ctx->err);
    return true;
  };
  const hipStream_t s = L.lane.stream;
  // order after everything already enqueued on the context's stream
  // (scene uploads, vrt_render frames of the same views)
  if (hipEventRecord(L.main_ready, ctx->stream) != hipSuccess || hipStreamWaitEvent(s, L.main_ready, 0) != hipSuccess ||
      hipMemsetAsync(L.d_vcount.ptr, 0, (uint64_t)n * 4 * sizeof(unsigned long long), s) != hipSuccess)
    return fail_all("serve launch");
  if (ctx->render_batch(L.lane, views.data(), n, states.data(), L.ev, vcount.data(), hdr.data()) != VRT_OK)
    return fail_all("serve batch");
  if (hipMemcpyAsync(L.h_vcount, L.d_vcount.ptr, (uint64_t)n * 4 * sizeof(unsigned long long),
                     hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipEventRecord(L.ev_done, s) != hipSuccess)
    return fail_all("serve read-back");
  for (uint32_t k = 0; k < n; ++k) {
    states[k]->frames++;
    ServeTicket& t = tickets.at(pick[k]);
    t.state = ServeTicket::RUNNING;
    t.timing.launch_ms = now;
    t.timing.batch_views = n;
    t.timing.stream = li;
    vuse[views[k].view_id].lane = (int)li;
  }
  L.running = true;
  L.batch_id = next_batch++;
  L.launch_ms = now;
  L.tickets = std::move(pick);
  for (uint64_t id : L.tickets) tickets.at(id).timing.batch_id = L.batch_id;
  st.batches++;
  st.batched_requests += n;
  st.max_running_batches = std::max<uint64_t>(st.max_running_batches, running_lanes());
  return true;
}

// Complete the batches whose last event has fired.
bool Server::harvest() {
  bool any = false;
  for (uint32_t li = 0; li < lanes.size(); ++li) {
    ServeLane& L = *lanes[li];
    if (!L.running) continue;
    const hipError_t q = hipEventQuery(L.ev_done);
    if (q == hipErrorNotReady) continue;
    const double now = ctx->now_ms();
    float gpu_ms = 0.0f;
    if (q == hipSuccess) (void)hipEventElapsedTime(&gpu_ms, L.ev[0], L.ev_done);
    for (uint32_t k = 0; k < L.tickets.size(); ++k) {
      const uint64_t id = L.tickets[k];
      auto it = tickets.find(id);
      if (it == tickets.end()) continue;
      ServeTicket& t = it->second;
      auto u = vuse.find(t.req.view.view_id);
      if (u != vuse.end()) u->second.lane = -1;
      if (q != hipSuccess) {
        fail_ticket(id, VRT_ERR_DEVICE, std::string("serve batch failed: ") + hipGetErrorString(q));
        continue;
      }
      for (int c = 0; c < 4; ++c) t.counts[c] = L.h_vcount[4 * k + c];
      t.timing.done_ms = now;
      t.timing.latency_ms = now - t.timing.submit_ms;
      t.timing.gpu_ms = gpu_ms;
      t.state = ServeTicket::DONE;
      retire(id);
      st.completed++;
    }
    ctx->frames += L.tickets.size();
    L.tickets.clear();
    L.running = false;
    any = true;
  }
  if (any) cv_done.notify_all();
  return any;
}

void Server::run() {
  (void)hipSetDevice(ctx->device);
  std::unique_lock<std::mutex> lk(ctx->mu);
  while (!stop) {
    bool progress = harvest();
    for (uint32_t li = 0; li < lanes.size() && !queue.empty(); ++li)
      if (!lanes[li]->running) progress |= launch(li);
    if (progress) {
      cv_done.notify_all();
      continue;
    }
    if (running_lanes() > 0)
      cv_work.wait_for(lk, std::chrono::microseconds(50));
    else
      cv_work.wait(lk);
  }
  // shutting down: let running batches finish, fail what never ran
  lk.unlock();
  for (auto& l : lanes)
    if (l->running) (void)hipStreamSynchronize(l->lane.stream);
  lk.lock();
  harvest();
  while (!queue.empty()) {
    fail_ticket(queue.front(), VRT_ERR_CANCELLED, "context destroyed");
    queue.pop_front();
  }
  cv_done.notify_all();
}

}  // namespace vrt

using vrt::Server;
using vrt::ServeTicket;

// Poll/wait body; lk holds ctx->mu.  The read-back runs without the lock.
static int take_result(vrt_ctx* ctx, std::unique_lock<std::mutex>& lk, vrt_ticket id, vrt_frame* out,
                       vrt_timing* timing) {
  Server& S = *ctx->server;
  auto it = S.tickets.find(id);
  if (it == S.tickets.end()) return ctx->fail(VRT_ERR_INVALID, "unknown ticket %llu", (unsigned long long)id);
  ServeTicket& t = it->second;
  if (t.state == ServeTicket::QUEUED || t.state == ServeTicket::RUNNING) {
    ctx->err = "ticket not finished";
    return VRT_ERR_NOT_READY;
  }
  if (timing) *timing = t.timing;
  if (t.state == ServeTicket::FAILED) {
    const int rc = t.error;
    ctx->err = t.msg;
    if (t.result) S.release_result(t.result, t.result_bytes);
    S.tickets.erase(it);
    return rc;
  }
  void* res = t.result;
  const uint64_t rbytes = t.result_bytes;
  const uint64_t px = (uint64_t)(uint32_t)t.req.view.u.params1[0] * (uint32_t)t.req.view.u.params1[1];
  uint64_t counts[4];
  std::memcpy(counts, t.counts, sizeof counts);
  S.tickets.erase(it);
  int rc = VRT_OK;
  if (out) {
    out->rays = counts[0];
    out->brick_steps = counts[1];
    out->voxel_steps = counts[2];
    if (out->hdr) {
      const hipStream_t cs = S.copy_stream;
      lk.unlock();
      hipError_t e = hipMemcpyAsync(out->hdr, res, px * 12, hipMemcpyDeviceToHost, cs);
      if (e == hipSuccess) e = hipStreamSynchronize(cs);
      lk.lock();
      if (e != hipSuccess) rc = ctx->fail(VRT_ERR_DEVICE, "result read-back: %s", hipGetErrorString(e));
    }
  }
  S.release_result(res, rbytes);
  return rc;
}

extern "C" {

int vrt_serve_configure(vrt_ctx* ctx, const vrt_serve_config* cfg) {
  if (!ctx) return VRT_ERR_INVALID;
  std::lock_guard<std::mutex> lk(ctx->mu);
  try {
    if (ctx->server && ctx->server->started)
      return ctx->fail(VRT_ERR_INVALID, "serving already started: configure before the first vrt_submit");
    if (!ctx->server) ctx->server = std::make_unique<Server>(ctx);
    if (cfg) {
      if (cfg->streams < 1 || cfg->streams > 4 || cfg->max_batch_views < 1 || cfg->max_batch_views > 1024 ||
          cfg->max_batch_pixels < 1)
        return ctx->fail(VRT_ERR_INVALID, "serve config out of range");
      ctx->server->cfg = *cfg;
    }
    return VRT_OK;
  } catch (...) {
    return ctx->fail(VRT_ERR_NO_MEMORY, "host allocation failed");
  }
}

int vrt_submit(vrt_ctx* ctx, const vrt_request* req, vrt_ticket* out) {
  if (!ctx || !req || !out) return ctx ? ctx->fail(VRT_ERR_INVALID, "null argument") : VRT_ERR_INVALID;
  std::lock_guard<std::mutex> lk(ctx->mu);
  try {
    *out = 0;
    if (req->flags != 0) return ctx->fail(VRT_ERR_INVALID, "request flags must be 0");
    int rc = ctx->validate_view(req->view);
    if (rc) return rc;
    (void)hipSetDevice(ctx->device);
    if (!ctx->server) ctx->server = std::make_unique<Server>(ctx);
    Server& S = *ctx->server;
    rc = S.start();
    if (rc) return rc;
    const uint64_t id = S.next_ticket++;
    ServeTicket& t = S.tickets[id];
    t.req = *req;
    t.timing.submit_ms = ctx->now_ms();
    const_cast<vrt_scene*>(req->view.scene)->uses++;
    S.vuse[req->view.view_id].pending++;
    S.queue.push_back(id);
    S.st.submitted++;
    *out = id;
  } catch (...) {
    return ctx->fail(VRT_ERR_NO_MEMORY, "host allocation failed");
  }
  ctx->server->cv_work.notify_one();
  return VRT_OK;
}

int vrt_poll(vrt_ctx* ctx, vrt_ticket t, vrt_frame* out, vrt_timing* timing) {
  if (!ctx) return VRT_ERR_INVALID;
  std::unique_lock<std::mutex> lk(ctx->mu);
  if (!ctx->server) return ctx->fail(VRT_ERR_INVALID, "unknown ticket %llu", (unsigned long long)t);
  try {
    return take_result(ctx, lk, t, out, timing);
  } catch (...) {
    return ctx->fail(VRT_ERR_NO_MEMORY, "host allocation failed");
  }
}

int vrt_wait(vrt_ctx* ctx, vrt_ticket t, uint32_t timeout_ms, vrt_frame* out, vrt_timing* timing) {
  if (!ctx) return VRT_ERR_INVALID;
  std::unique_lock<std::mutex> lk(ctx->mu);
  if (!ctx->server) return ctx->fail(VRT_ERR_INVALID, "unknown ticket %llu", (unsigned long long)t);
  Server& S = *ctx->server;
  auto finished = [&] {
    auto it = S.tickets.find(t);
    return it == S.tickets.end() || it->second.state == ServeTicket::DONE ||
           it->second.state == ServeTicket::FAILED;
  };
  if (timeout_ms == 0)
    S.cv_done.wait(lk, finished);
  else
    S.cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), finished);
  try {
    return take_result(ctx, lk, t, out, timing);
  } catch (...) {
    return ctx->fail(VRT_ERR_NO_MEMORY, "host allocation failed");
  }
}

int vrt_cancel(vrt_ctx* ctx, vrt_ticket t) {
  if (!ctx) return VRT_ERR_INVALID;
  std::lock_guard<std::mutex> lk(ctx->mu);
  if (!ctx->server) return ctx->fail(VRT_ERR_INVALID, "unknown ticket %llu", (unsigned long long)t);
  Server& S = *ctx->server;
  auto it = S.tickets.find(t);
  if (it == S.tickets.end()) return ctx->fail(VRT_ERR_INVALID, "unknown ticket %llu", (unsigned long long)t);
  if (it->second.state != ServeTicket::QUEUED)
    return ctx->fail(VRT_ERR_INVALID, "ticket %llu already launched", (unsigned long long)t);
  auto q = std::find(S.queue.begin(), S.queue.end(), t);
  if (q != S.queue.end()) S.queue.erase(q);
  S.retire(t);
  S.tickets.erase(it);
  S.st.cancelled++;
  return VRT_OK;
}

int vrt_serve_stats_get(vrt_ctx* ctx, vrt_serve_stats* out) {
  if (!ctx || !out) return VRT_ERR_INVALID;
  std::lock_guard<std::mutex> lk(ctx->mu);
  std::memset(out, 0, sizeof *out);
  if (!ctx->server) return VRT_OK;
  Server& S = *ctx->server;
  *out = S.st;
  out->queued = S.queue.size();
  uint64_t running = 0;
  for (const auto& l : S.lanes) running += l->running ? l->tickets.size() : 0;
  out->running = running;
  return VRT_OK;
}

}  // extern "C"

// The context's stream waits for every running serve batch.  Called with
// ctx->mu held by entry points that enqueue work on the context's stream.
int vrt_ctx::order_after_serve() {
  if (!server) return VRT_OK;
  for (auto& l : server->lanes)
    if (l->running) {
      hipError_t e = hipStreamWaitEvent(stream, l->ev_done, 0);
      if (e != hipSuccess) return fail(VRT_ERR_DEVICE, "stream ordering: %s", hipGetErrorString(e));
    }
  return VRT_OK;
}
