// vrt_runtime.cpp -- host runtime and C ABI of libvoxelrt_hip.
//
// Owns the HIP device state: one stream per context (default priority), the
// resident scenes (brick words + compacted brick pool), the per-view frame
// targets (G-buffer / accumulation ping-pong and denoise targets, the
// reference's textures of main.js:341-349), and the frame orchestration of
// main.js:680-747 (pathtrace -> temporal -> a-trous x3).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/voxelrt.h"
#include "vrt_internal.h"
#include "vrt_kernels.h"
#include "vrt_trace.h"
#include "vrt_batch.h"
#include "vrt_serve.h"

namespace vrt {
hipError_t exclusive_scan_u32(hipStream_t st, void* d_tmp, size_t* tmp_bytes, const uint32_t* in,
                              uint32_t* out, uint64_t n);
}

using vrt::DeviceBuffer;

// ---------------------------------------------------------------------------
// error helpers
// ---------------------------------------------------------------------------
int vrt_ctx::fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  err = buf;
  return code;
}

#define HIP_TRY(ctx, expr)                                                                     \
  do {                                                                                         \
    hipError_t e_ = (expr);                                                                    \
    if (e_ != hipSuccess)                                                                      \
      return (ctx)->fail(e_ == hipErrorOutOfMemory ? VRT_ERR_NO_MEMORY : VRT_ERR_DEVICE,       \
                         "%s failed: %s", #expr, hipGetErrorString(e_));                       \
  } while (0)

#define GUARD_BEGIN try {
#define GUARD_END(ctx)                                                                         \
  }                                                                                            \
  catch (const std::bad_alloc&) {                                                              \
    return (ctx) ? (ctx)->fail(VRT_ERR_NO_MEMORY, "host allocation failed") : VRT_ERR_NO_MEMORY; \
  }                                                                                            \
  catch (...) {                                                                                \
    return (ctx) ? (ctx)->fail(VRT_ERR_DEVICE, "unexpected exception") : VRT_ERR_DEVICE;       \
  }

namespace vrt {

hipError_t DeviceBuffer::alloc(size_t n) {
  release();
  if (n == 0) return hipSuccess;
  hipError_t e = hipMalloc(&ptr, n);
  if (e != hipSuccess) {
    ptr = nullptr;
    return e;
  }
  bytes = n;
  return hipSuccess;
}
void DeviceBuffer::release() {
  if (ptr) (void)hipFree(ptr);
  ptr = nullptr;
  bytes = 0;
}
hipError_t DeviceBuffer::ensure(size_t n) {
  if (n <= bytes) return hipSuccess;
  return alloc(n);
}

}  // namespace vrt

static uint64_t view_bytes(uint32_t w, uint32_t h, uint32_t fmt) {
  const uint64_t px = (uint64_t)w * h;
  return px * (8 + 4 + 16 * 2 + (fmt == VRT_ACCUM_RGBA32F ? 16 : 8) * 2 + 8 * 2);
}

int vrt_ctx::reserve(uint64_t bytes) {
  if (resident + bytes > budget && !make_scene_room(bytes, nullptr))
    return fail(VRT_ERR_BUDGET, "HBM budget exceeded: %llu resident + %llu requested > %llu",
                (unsigned long long)resident, (unsigned long long)bytes, (unsigned long long)budget);
  resident += bytes;
  return VRT_OK;
}

int vrt_ctx::evict_scene(vrt_scene* s) {
  if (s->evicted) return VRT_OK;
  // frames enqueued on the context's stream may still read it (serve batches
  // cannot: the scene has no queued or running request)
  HIP_TRY(this, hipStreamSynchronize(stream));
  auto b = s->buffers();
  std::vector<std::vector<uint8_t>> host(b.size());
  for (size_t i = 0; i < b.size(); ++i)
    if (b[i]->ptr) {
      host[i].resize(b[i]->bytes);
      HIP_TRY(this, hipMemcpyAsync(host[i].data(), b[i]->ptr, b[i]->bytes, hipMemcpyDeviceToHost, stream));
    }
  HIP_TRY(this, hipStreamSynchronize(stream));
  for (auto* d : b) d->release();
  s->host = std::move(host);
  s->evicted = true;
  unreserve(s->bytes);
  ++scene_evictions;
  return VRT_OK;
}

int vrt_ctx::restore_scene(vrt_scene* s) {
  if (!s->evicted) return VRT_OK;
  ++s->pins;  // not a candidate while room is made for it
  const int rc = reserve(s->bytes);
  --s->pins;
  if (rc) return rc;
  auto b = s->buffers();
  for (size_t i = 0; i < b.size(); ++i) {
    if (s->host[i].empty()) continue;
    hipError_t e = b[i]->alloc(s->host[i].size());
    if (e == hipSuccess)
      e = hipMemcpyAsync(b[i]->ptr, s->host[i].data(), s->host[i].size(), hipMemcpyHostToDevice, stream);
    if (e != hipSuccess) {
      for (auto* d : b) d->release();
      unreserve(s->bytes);
      return fail(e == hipErrorOutOfMemory ? VRT_ERR_NO_MEMORY : VRT_ERR_DEVICE, "scene restore: %s",
                  hipGetErrorString(e));
    }
  }
  HIP_TRY(this, hipStreamSynchronize(stream));  // before the host copies go
  s->host.clear();
  s->evicted = false;
  s->lru = ++clock;
  ++scene_restores;
  return VRT_OK;
}

// The scenes a C ABI call uses, restored to the device when evicted and
// pinned (not evictable) until the call returns.
struct ScenePins {
  vrt_ctx* ctx;
  std::vector<vrt_scene*> held;
  explicit ScenePins(vrt_ctx* c) : ctx(c) {}
  ScenePins(const ScenePins&) = delete;
  ScenePins& operator=(const ScenePins&) = delete;
  int add(const vrt_scene* sc) {
    vrt_scene* s = const_cast<vrt_scene*>(sc);
    if (std::find(held.begin(), held.end(), s) != held.end()) return VRT_OK;
    held.push_back(s);
    ++s->pins;
    s->lru = ++ctx->clock;
    return ctx->restore_scene(s);
  }
  ~ScenePins() {
    for (vrt_scene* s : held) --s->pins;
  }
};

bool vrt_ctx::make_scene_room(uint64_t need, const vrt_scene* keep) {
  while (resident + need > budget) {
    vrt_scene* victim = nullptr;
    for (vrt_scene* s : scenes)
      if (s->evictable && !s->evicted && s->uses == 0 && s->pins == 0 && s != keep &&
          (!victim || s->lru < victim->lru))
        victim = s;
    if (!victim || evict_scene(victim) != VRT_OK) return false;
  }
  return true;
}
void vrt_ctx::unreserve(uint64_t bytes) { resident = bytes > resident ? 0 : resident - bytes; }

int vrt_ctx::counters_to_host(uint64_t out[4]) {
  HIP_TRY(this, hipMemcpyAsync(out, d_counters.ptr, 4 * sizeof(uint64_t), hipMemcpyDeviceToHost,
                               stream));
  HIP_TRY(this, hipStreamSynchronize(stream));
  return VRT_OK;
}

vrt::SceneDev vrt_scene::dev() const {
  vrt::SceneDev s;
  s.bricks = (const uint32_t*)d_bricks.ptr;
  s.pool = (const uint32_t*)d_pool.ptr;
  s.bits = (const uint32_t*)d_bits.ptr;
  s.prefix = (const uint32_t*)d_prefix.ptr;
  s.pbits = (const uint32_t*)d_pbits.ptr;
  s.pprefix = (const uint32_t*)d_pprefix.ptr;
  s.pdist = (const uint8_t*)d_pdist.ptr;
  s.vbox = (const uint32_t*)d_vbox.ptr;
  s.entry_batch = 1u;
  s.tile_order = 0u;
  s.pwords = psh3 >= 5 ? (1u << (psh3 - 5)) : 1u;
  s.psh1 = psh1;
  s.psh2 = psh2;
  s.bit_words = (uint32_t)((nbricks + 31) / 32);
  s.nx = (int)nx;
  s.ny = (int)ny;
  s.nz = (int)nz;
  s.bgx = (int)(nx / 8);
  s.bgy = (int)(ny / 8);
  s.bgz = (int)(nz / 8);
  s.voxel_size = VRT_VOXEL_SIZE;
  s.inv_voxel = 1.0f / VRT_VOXEL_SIZE;
  s.entry_batch = 1u;
  for (int c = 0; c < 3; ++c) {
    s.box_lo[c] = box_lo[c];
    s.box_hi[c] = box_hi[c];
  }
  s.zero = 0.0f;
  return s;
}

// ---------------------------------------------------------------------------
// view state
// ---------------------------------------------------------------------------
int vrt_ctx::view_state(uint32_t id, uint32_t w, uint32_t h, uint32_t fmt, ViewState** out, hipStream_t st) {
  if (!st) st = stream;
  auto it = views.find(id);
  if (it != views.end()) {
    ViewState& v = it->second;
    if (v.w == w && v.h == h && v.fmt == fmt) {
      *out = &v;
      return VRT_OK;
    }
    unreserve(v.bytes);
    views.erase(it);
  }
  ViewState& v = views[id];
  int rc = alloc_view_state(v, w, h, fmt, st);
  if (rc) {
    views.erase(id);
    return rc;
  }
  *out = &v;
  return VRT_OK;
}

int vrt_ctx::alloc_view_state(ViewState& v, uint32_t w, uint32_t h, uint32_t fmt, hipStream_t st) {
  const uint64_t bytes = view_bytes(w, h, fmt);
  int rc = reserve(bytes);
  if (rc) return rc;
  v.w = w;
  v.h = h;
  v.fmt = fmt;
  v.bytes = bytes;
  v.frames = 0;
  const uint64_t px = (uint64_t)w * h;
  const uint64_t abytes = px * (fmt == VRT_ACCUM_RGBA32F ? 16 : 8);
  hipError_t e = hipSuccess;
  if (e == hipSuccess) e = v.radiance.alloc(px * 8);
  if (e == hipSuccess) e = v.albedo.alloc(px * 4);
  for (int k = 0; k < 2 && e == hipSuccess; ++k) e = v.gbuf[k].alloc(px * 16);
  for (int k = 0; k < 2 && e == hipSuccess; ++k) e = v.accum[k].alloc(abytes);
  for (int k = 0; k < 2 && e == hipSuccess; ++k) e = v.den[k].alloc(px * 8);
  // WebGPU zero-initialises textures: history reads before the first write
  // see zeros (and gbuf.w = 0 fails the depth test).
  for (int k = 0; k < 2 && e == hipSuccess; ++k) e = hipMemsetAsync(v.gbuf[k].ptr, 0, px * 16, st);
  for (int k = 0; k < 2 && e == hipSuccess; ++k) e = hipMemsetAsync(v.accum[k].ptr, 0, abytes, st);
  if (e != hipSuccess) {
    unreserve(bytes);
    return fail(e == hipErrorOutOfMemory ? VRT_ERR_NO_MEMORY : VRT_ERR_DEVICE,
                "view state allocation failed: %s", hipGetErrorString(e));
  }
  return VRT_OK;
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int vrt_device_count(int32_t* count) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (count) *count = (e == hipSuccess) ? n : 0;
  return e == hipSuccess ? VRT_OK : VRT_ERR_NO_DEVICE;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Scratch of the allocation-free traversal entry point
// (vrt_trace_rays_device_noalloc): the four 64-bit counters its kernels add
// rays, brick steps, voxel steps and hits into (bytes 0..31) and the
// persistent kernels' ray queue (bytes 256..1279: eight 32-bit counters, one
// per 128-byte line); 1280 bytes, one copy per device, held until the
// process ends.  Host code takes its device
// address per device once (scratch_address, from vrt_create, with the device
// current) and uses only that address afterwards.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned long long scratch[160];
constexpr uint32_t NOALLOC_QUEUE_WORD = 64;      // u32 index of the ray queue in scratch
constexpr uint32_t NOALLOC_SCRATCH_WORDS = 320;  // u32 words the entry point clears per call

static std::mutex g_scratch_mu;
static std::unordered_map<int, void*> g_scratch_addr;  // device ordinal -> address of scratch

static hipError_t scratch_address(int device, void** out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch_addr.find(device);
  if (it == g_scratch_addr.end()) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    it = g_scratch_addr.emplace(device, p).first;
  }
  *out = it->second;
  return hipSuccess;
}

extern "C" {

int vrt_create(const vrt_create_info* info, vrt_ctx** out) {
  if (!out) return VRT_ERR_INVALID;
  *out = nullptr;
  vrt_ctx* ctx = nullptr;
  try {
    ctx = new vrt_ctx();
  } catch (...) {
    return VRT_ERR_NO_MEMORY;
  }
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    delete ctx;
    return VRT_ERR_NO_DEVICE;
  }
  ctx->device = info ? info->device : 0;
  if (ctx->device < 0 || ctx->device >= n) {
    delete ctx;
    return VRT_ERR_INVALID;
  }
  if (hipSetDevice(ctx->device) != hipSuccess ||
      hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return VRT_ERR_DEVICE;
  }
  ctx->lane0.stream = ctx->stream;
  // the allocation-free entry point's scratch on this device; without it
  // that entry point reports an error and everything else works
  if (scratch_address(ctx->device, &ctx->noalloc_scratch) != hipSuccess) {
    ctx->noalloc_scratch = nullptr;
    (void)hipGetLastError();
  }
  {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device) == hipSuccess && cus > 0)
      ctx->cus = cus;
  }
  for (auto& e : ctx->ev)
    if (hipEventCreate(&e) != hipSuccess) {
      delete ctx;
      return VRT_ERR_DEVICE;
    }
  for (auto& slot : ctx->ring)
    for (auto& e : slot.e)
      if (hipEventCreate(&e) != hipSuccess) {
        delete ctx;
        return VRT_ERR_DEVICE;
      }
  size_t freeb = 0, total = 0;
  (void)hipMemGetInfo(&freeb, &total);
  ctx->budget = (info && info->hbm_budget) ? info->hbm_budget : (uint64_t)(0.9 * (double)freeb);
  ctx->tuning.trace_kernel = 7;
  ctx->tuning.trace_blocks = 1024;
  ctx->tuning.trace_steps = 32;
  ctx->tuning.trace_min_refill = 16;
  ctx->tuning.dda_batch = 16;
  ctx->tuning.pt_kernel = 8;
  ctx->tuning.pt_order = 2;  // last tile first: the bottom rows (near ground) usually cost most
  ctx->tuning.pt_blocks = 0;
  ctx->tuning.pt_waves = 4;
  ctx->tuning.pt_min_refill = 16;
  ctx->tuning.pt_shade_min = 40;
  if (ctx->d_queue.alloc((2 * vrt::QUEUE_WORDS + 64) * sizeof(uint32_t)) != hipSuccess ||
      ctx->d_counters.alloc(4 * sizeof(uint64_t)) != hipSuccess ||
      hipMemsetAsync(ctx->d_counters.ptr, 0, 4 * sizeof(uint64_t), ctx->stream) != hipSuccess) {
    delete ctx;
    return VRT_ERR_DEVICE;
  }
  *out = ctx;
  return VRT_OK;
}

void vrt_destroy(vrt_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  delete ctx;
}

const char* vrt_last_error(const vrt_ctx* ctx) {
  if (!ctx) return "null context";
  vrt_ctx* c = const_cast<vrt_ctx*>(ctx);
  std::lock_guard<std::mutex> lock_(c->mu);
  c->err_ret = c->err;  // a stable copy: the serving worker may set err later
  return c->err_ret.c_str();
}

// ---------------------------------------------------------------------------
// scenes
// ---------------------------------------------------------------------------
}  // extern "C"

// Build (or rebuild) a scene's device structures from a dense packed-u32
// grid: host memory (hvox, staged in slabs of whole brick layers) or device
// memory (dvox, the editable scene's resident copy).  Pass 1 builds the
// 512-bit masks and occupancy flags, a device exclusive scan assigns pool
// slots in brick order, pass 2 packs masks + voxels into the pool, then the
// bitmaps and their prefix counts are built.  On a rebuild the pool grows
// when the occupied-brick count needs it; the other buffers depend on the
// dimensions only.  Ends with the context's stream synchronised.
static int scene_build(vrt_ctx* ctx, vrt_scene* sc, const uint32_t* hvox, const uint32_t* dvox) {
  const uint32_t nx = sc->nx, ny = sc->ny, nz = sc->nz;
  const int bgx = (int)nx / 8, bgy = (int)ny / 8, bgz = (int)nz / 8;
  const uint64_t nb = (uint64_t)bgx * bgy * bgz;
  const uint64_t layer_bytes = (uint64_t)nx * ny * 8 * 4;  // one brick layer of voxels
  const uint64_t stage_cap = 512ull << 20;
  const int slab = dvox ? bgz : (int)std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)bgz, stage_cap / layer_bytes));
  const bool single = slab >= bgz;
  vrt::DeviceBuffer stage, masks, flags, slots, tmp;
  if (!dvox) HIP_TRY(ctx, stage.alloc(layer_bytes * slab));
  const uint32_t* src = dvox ? dvox : (const uint32_t*)stage.ptr;
  HIP_TRY(ctx, masks.alloc(nb * 64));
  HIP_TRY(ctx, flags.alloc(nb * 4));
  HIP_TRY(ctx, slots.alloc(nb * 4));
  // pass 1: masks + occupancy flags
  for (int bz0 = 0; bz0 < bgz; bz0 += slab) {
    const int nbz = std::min(slab, bgz - bz0);
    if (!dvox)
      HIP_TRY(ctx, hipMemcpyAsync(stage.ptr, hvox + (uint64_t)bz0 * 8 * nx * ny, layer_bytes * nbz,
                                  hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, vrt::launch_brick_masks(ctx->stream, src + (dvox ? (uint64_t)bz0 * 8 * nx * ny : 0), (int)nx,
                                         (int)ny, bgx, bgy, bz0, nbz, (uint32_t*)masks.ptr, (uint32_t*)flags.ptr));
  }
  // slots = exclusive prefix of flags
  size_t tmp_bytes = 0;
  HIP_TRY(ctx, vrt::exclusive_scan_u32(ctx->stream, nullptr, &tmp_bytes, (const uint32_t*)flags.ptr,
                                       (uint32_t*)slots.ptr, nb));
  HIP_TRY(ctx, tmp.alloc(std::max<size_t>(tmp_bytes, 16)));
  HIP_TRY(ctx, vrt::exclusive_scan_u32(ctx->stream, tmp.ptr, &tmp_bytes, (const uint32_t*)flags.ptr,
                                       (uint32_t*)slots.ptr, nb));
  uint32_t last[2] = {0, 0};
  HIP_TRY(ctx, hipMemcpyAsync(&last[0], (uint32_t*)slots.ptr + (nb - 1), 4, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipMemcpyAsync(&last[1], (uint32_t*)flags.ptr + (nb - 1), 4, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  const uint64_t nocc = (uint64_t)last[0] + last[1];
  sc->nocc = nocc;
  // padded bitmap over a power-of-two brick grid: ceil(log2) bits per axis
  auto lg = [](uint32_t v) {
    uint32_t k = 0;
    while ((1u << k) < v) ++k;
    return k;
  };
  sc->psh1 = lg((uint32_t)bgx);
  sc->psh2 = sc->psh1 + lg((uint32_t)bgy);
  sc->psh3 = sc->psh2 + lg((uint32_t)bgz);
  const uint64_t pbits_n = 1ull << sc->psh3;
  const uint64_t pool_need = std::max<uint64_t>(nocc, 1) * vrt::POOL_WORDS * 4;
  // editable scenes keep 25 % headroom in the pool so that edits rarely reallocate
  const uint64_t pool_bytes = sc->d_pool.bytes >= pool_need ? sc->d_pool.bytes
                              : (sc->d_dense.ptr ? pool_need + pool_need / 4 : pool_need);
  const uint64_t pdist_bytes = (((pbits_n + 31) / 32) + 3) & ~3ull;
  const uint64_t vbox_bytes = pool_bytes / (vrt::POOL_WORDS * 4) * 4;  // a word per pool slot
  const uint64_t bytes = sc->d_dense.bytes + nb * 4 + pool_bytes + (nb + 63) / 64 * 8 + (nb + 31) / 32 * 4 +
                         (pbits_n + 63) / 64 * 8 + (pbits_n + 31) / 32 * 4 + pdist_bytes + vbox_bytes;
  if (bytes > sc->bytes) {
    int rc = ctx->reserve(bytes - sc->bytes);
    if (rc) return rc;
    sc->bytes = bytes;
  }
  hipError_t e = hipSuccess;
  if (!sc->d_bricks.ptr) e = sc->d_bricks.alloc(nb * 4);
  if (e == hipSuccess && sc->d_pool.bytes < pool_need) e = sc->d_pool.alloc(pool_bytes);
  if (e == hipSuccess && !sc->d_bits.ptr) e = sc->d_bits.alloc((nb + 63) / 64 * 8);
  if (e == hipSuccess && !sc->d_prefix.ptr) e = sc->d_prefix.alloc((nb + 31) / 32 * 4);
  if (e == hipSuccess && !sc->d_pbits.ptr) e = sc->d_pbits.alloc(((pbits_n + 63) / 64) * 8);
  if (e == hipSuccess && !sc->d_pprefix.ptr) e = sc->d_pprefix.alloc((pbits_n + 31) / 32 * 4);
  if (e == hipSuccess && !sc->d_pdist.ptr) e = sc->d_pdist.alloc(pdist_bytes);
  if (e == hipSuccess && sc->d_vbox.bytes < vbox_bytes) e = sc->d_vbox.alloc(vbox_bytes);
  if (e != hipSuccess)
    return ctx->fail(e == hipErrorOutOfMemory ? VRT_ERR_NO_MEMORY : VRT_ERR_DEVICE, "scene allocation failed: %s",
                     hipGetErrorString(e));
  auto err = [&](const char* what, hipError_t er) { return ctx->fail(VRT_ERR_DEVICE, "%s: %s", what, hipGetErrorString(er)); };
  e = vrt::launch_brick_words(ctx->stream, (const uint32_t*)flags.ptr, (const uint32_t*)slots.ptr, nb,
                              (uint32_t*)sc->d_bricks.ptr);
  if (e != hipSuccess) return err("brick words", e);
  e = vrt::launch_brick_bits(ctx->stream, (const uint32_t*)sc->d_bricks.ptr, nb, (uint32_t*)sc->d_bits.ptr);
  if (e != hipSuccess) return err("brick bits", e);
  {
    // prefix[w] = exclusive scan of popcount(bits[w]): the pool slot of the
    // first occupied brick at or after word w (slots follow brick order)
    const uint64_t nw = (nb + 31) / 32;
    e = vrt::launch_popc(ctx->stream, (const uint32_t*)sc->d_bits.ptr, nw, (uint32_t*)flags.ptr);
    if (e != hipSuccess) return err("bit counts", e);
    size_t need = 0;
    e = vrt::exclusive_scan_u32(ctx->stream, nullptr, &need, (const uint32_t*)flags.ptr,
                                (uint32_t*)sc->d_prefix.ptr, nw);
    if (e == hipSuccess && need > tmp.bytes) e = tmp.alloc(need);
    if (e == hipSuccess)
      e = vrt::exclusive_scan_u32(ctx->stream, tmp.ptr, &need, (const uint32_t*)flags.ptr,
                                  (uint32_t*)sc->d_prefix.ptr, nw);
    if (e != hipSuccess) return err("bit prefix", e);
    // the padded layout
    const uint64_t pn = 1ull << sc->psh3, pw = (pn + 31) / 32;
    e = vrt::launch_brick_pbits(ctx->stream, (const uint32_t*)sc->d_bricks.ptr, bgx, bgy, bgz, sc->psh1, sc->psh2,
                                pn, (uint32_t*)sc->d_pbits.ptr);
    if (e == hipSuccess) e = vrt::launch_popc(ctx->stream, (const uint32_t*)sc->d_pbits.ptr, pw, (uint32_t*)flags.ptr);
    if (e != hipSuccess) return err("padded bits", e);
    need = 0;
    e = vrt::exclusive_scan_u32(ctx->stream, nullptr, &need, (const uint32_t*)flags.ptr,
                                (uint32_t*)sc->d_pprefix.ptr, pw);
    if (e == hipSuccess && need > tmp.bytes) e = tmp.alloc(need);
    if (e == hipSuccess)
      e = vrt::exclusive_scan_u32(ctx->stream, tmp.ptr, &need, (const uint32_t*)flags.ptr,
                                  (uint32_t*)sc->d_pprefix.ptr, pw);
    if (e != hipSuccess) return err("padded prefix", e);
    // row distances for the empty-row skip of the single-loop DDA
    vrt::DeviceBuffer rows;
    e = rows.alloc(3ull * (uint64_t)bgy * (uint64_t)bgz);
    if (e == hipSuccess)
      e = vrt::launch_row_dist(ctx->stream, (const uint32_t*)sc->d_bricks.ptr, bgx, bgy, bgz, sc->psh1, sc->psh2, pw,
                               (uint8_t*)rows.ptr, (uint8_t*)sc->d_pdist.ptr);
    // the box of occupied bricks, from each row's first and last occupied bx
    vrt::DeviceBuffer span;
    std::vector<uint32_t> hspan((size_t)bgy * (size_t)bgz);
    if (e == hipSuccess) e = span.alloc(hspan.size() * 4);
    if (e == hipSuccess)
      e = vrt::launch_row_span(ctx->stream, (const uint32_t*)sc->d_bricks.ptr, bgx, bgy, bgz, (uint32_t*)span.ptr);
    if (e == hipSuccess)
      e = hipMemcpyAsync(hspan.data(), span.ptr, hspan.size() * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);  // before rows and span are freed
    if (e != hipSuccess) return err("row distances", e);
    int lo[3] = {bgx, bgy, bgz}, hi[3] = {-1, -1, -1};
    for (int bz = 0; bz < bgz; ++bz)
      for (int by = 0; by < bgy; ++by) {
        const uint32_t sp = hspan[(size_t)by + (size_t)bz * (size_t)bgy];
        if (sp == 0xFFFFu) continue;
        lo[0] = std::min(lo[0], (int)(sp & 0xFFFFu));
        hi[0] = std::max(hi[0], (int)(sp >> 16));
        lo[1] = std::min(lo[1], by);
        hi[1] = std::max(hi[1], by);
        lo[2] = std::min(lo[2], bz);
        hi[2] = std::max(hi[2], bz);
      }
    for (int c = 0; c < 3; ++c) {
      sc->box_lo[c] = lo[c];
      sc->box_hi[c] = hi[c];
    }
  }
  // pass 2: pack occupied bricks into the pool
  for (int bz0 = 0; bz0 < bgz; bz0 += slab) {
    const int nbz = std::min(slab, bgz - bz0);
    if (!dvox && !single) {
      e = hipMemcpyAsync(stage.ptr, hvox + (uint64_t)bz0 * 8 * nx * ny, layer_bytes * nbz, hipMemcpyHostToDevice,
                         ctx->stream);
      if (e != hipSuccess) return err("upload", e);
    }
    e = vrt::launch_pack_pool(ctx->stream, src + (dvox ? (uint64_t)bz0 * 8 * nx * ny : 0), (int)nx, (int)ny, bgx,
                              bgy, bz0, nbz, (const uint32_t*)masks.ptr, (const uint32_t*)sc->d_bricks.ptr,
                              (uint32_t*)sc->d_pool.ptr);
    if (e != hipSuccess) return err("pack", e);
  }
  e = vrt::launch_brick_vbox(ctx->stream, (const uint32_t*)sc->d_pool.ptr, nocc, (uint32_t*)sc->d_vbox.ptr);
  if (e != hipSuccess) return err("voxel boxes", e);
  e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return err("scene build", e);
  return VRT_OK;
}

extern "C" {

int vrt_scene_upload_ex(vrt_ctx* ctx, const uint32_t* vox, uint32_t nx, uint32_t ny, uint32_t nz,
                        uint64_t scene_key, uint32_t flags, vrt_scene** out) {
  if (!ctx || !vox || !out) return ctx ? ctx->fail(VRT_ERR_INVALID, "null argument") : VRT_ERR_INVALID;
  GUARD_BEGIN
  std::lock_guard<std::mutex> lock_(ctx->mu);
  *out = nullptr;
  if (flags & ~(VRT_SCENE_EDITABLE | VRT_SCENE_EVICTABLE))
    return ctx->fail(VRT_ERR_INVALID, "unknown scene flags 0x%x", flags);
  if (nx == 0 || ny == 0 || nz == 0 || nx % 8 || ny % 8 || nz % 8)
    return ctx->fail(VRT_ERR_INVALID, "grid dims %ux%ux%u must be positive multiples of 8", nx, ny, nz);
  if ((uint64_t)nx * ny * nz > 0xFFFFFFFEull)
    return ctx->fail(VRT_ERR_INVALID, "grid has more than 2^32-2 voxels");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  std::unique_ptr<vrt_scene> sc(new vrt_scene());
  sc->nx = nx;
  sc->ny = ny;
  sc->nz = nz;
  sc->key = scene_key;
  sc->nbricks = (uint64_t)(nx / 8) * (ny / 8) * (nz / 8);
  sc->pins = 1;  // building: not evictable (joins ctx->scenes at the end)
  const uint64_t dense = (uint64_t)nx * ny * nz * 4;
  if (flags & VRT_SCENE_EDITABLE) {
    int rc = ctx->reserve(dense);
    if (rc) return rc;
    sc->bytes = dense;
    hipError_t e = sc->d_dense.alloc(dense);
    if (e == hipSuccess) e = hipMemcpyAsync(sc->d_dense.ptr, vox, dense, hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) {
      ctx->unreserve(dense);
      return ctx->fail(e == hipErrorOutOfMemory ? VRT_ERR_NO_MEMORY : VRT_ERR_DEVICE, "dense grid: %s",
                       hipGetErrorString(e));
    }
  }
  const int rc = scene_build(ctx, sc.get(), vox, (const uint32_t*)sc->d_dense.ptr);
  if (rc) {
    ctx->unreserve(sc->bytes);
    return rc;
  }
  sc->lru = ++ctx->clock;
  sc->pins = 0;
  sc->evictable = (flags & VRT_SCENE_EVICTABLE) != 0;
  vrt_scene* raw = sc.release();
  ctx->scenes.push_back(raw);
  *out = raw;
  return VRT_OK;
  GUARD_END(ctx)
}

int vrt_scene_upload(vrt_ctx* ctx, const uint32_t* vox, uint32_t nx, uint32_t ny, uint32_t nz,
                     uint64_t scene_key, vrt_scene** out) {
  return vrt_scene_upload_ex(ctx, vox, nx, ny, nz, scene_key, 0u, out);
}

int vrt_scene_edit(vrt_ctx* ctx, vrt_scene* scene, const uint32_t* vox, uint32_t x0, uint32_t y0, uint32_t z0,
                   uint32_t sx, uint32_t sy, uint32_t sz) {
  if (!ctx || !scene || !vox) return ctx ? ctx->fail(VRT_ERR_INVALID, "null argument") : VRT_ERR_INVALID;
  GUARD_BEGIN
  std::lock_guard<std::mutex> lock_(ctx->mu);
  if (std::find(ctx->scenes.begin(), ctx->scenes.end(), scene) == ctx->scenes.end())
    return ctx->fail(VRT_ERR_INVALID, "unknown scene handle");
  if (!scene->d_dense.ptr && !(scene->evicted && !scene->host[7].empty()))
    return ctx->fail(VRT_ERR_INVALID, "scene was not uploaded with VRT_SCENE_EDITABLE");
  if (sx == 0 || sy == 0 || sz == 0) return VRT_OK;
  if ((uint64_t)x0 + sx > scene->nx || (uint64_t)y0 + sy > scene->ny || (uint64_t)z0 + sz > scene->nz)
    return ctx->fail(VRT_ERR_INVALID, "edit box %u+%u, %u+%u, %u+%u outside the %ux%ux%u grid", x0, sx, y0, sy, z0,
                     sz, scene->nx, scene->ny, scene->nz);
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  {
    // running serve batches may read this scene: the rebuild comes after them
    int rc = ctx->order_after_serve();
    if (rc) return rc;
  }
  ScenePins pins(ctx);
  if (int rc = pins.add(scene)) return rc;
  uint32_t* dense = (uint32_t*)scene->d_dense.ptr;
  for (uint32_t z = 0; z < sz; ++z)
    HIP_TRY(ctx, hipMemcpy2DAsync(dense + ((uint64_t)(z0 + z) * scene->ny + y0) * scene->nx + x0,
                                  (size_t)scene->nx * 4, vox + (uint64_t)z * sx * sy, (size_t)sx * 4,
                                  (size_t)sx * 4, sy, hipMemcpyHostToDevice, ctx->stream));
  // only the bricks the box touches: their pool entries are rewritten in
  // place when none of them changes occupancy (the brick words, bitmaps,
  // prefix counts and row distances then stay as they are); a change of
  // occupancy moves the slots of later bricks, and the scene is rebuilt
  const int bx0 = (int)(x0 / 8), by0 = (int)(y0 / 8), bz0 = (int)(z0 / 8);
  const int ex = (int)((x0 + sx - 1) / 8) - bx0 + 1, ey = (int)((y0 + sy - 1) / 8) - by0 + 1,
            ez = (int)((z0 + sz - 1) / 8) - bz0 + 1;
  uint32_t* flag = (uint32_t*)ctx->d_queue.ptr + 2 * vrt::QUEUE_WORDS;  // one word past the work queues
  HIP_TRY(ctx, vrt::launch_zero_words(ctx->stream, flag, 1u));
  HIP_TRY(ctx, vrt::launch_edit_bricks(ctx->stream, dense, (int)scene->nx, (int)scene->ny, (int)scene->nx / 8,
                                       (int)scene->ny / 8, bx0, by0, bz0, ex, ey, ez,
                                       (const uint32_t*)scene->d_bricks.ptr, (uint32_t*)scene->d_pool.ptr, flag));
  uint32_t changed = 0;
  HIP_TRY(ctx, hipMemcpyAsync(&changed, flag, 4, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  scene->edits_in_place += changed ? 0u : 1u;
  scene->edits_rebuilt += changed ? 1u : 0u;
  if (changed) return scene_build(ctx, scene, nullptr, dense);
  // in place: the touched bricks' masks may have changed, and with them
  // their boxes of solid voxels
  HIP_TRY(ctx, vrt::launch_brick_vbox(ctx->stream, (const uint32_t*)scene->d_pool.ptr, scene->nocc,
                                      (uint32_t*)scene->d_vbox.ptr));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return VRT_OK;
  GUARD_END(ctx)
}

int vrt_scene_release(vrt_ctx* ctx, vrt_scene* scene) {
  if (!ctx || !scene) return VRT_ERR_INVALID;
  GUARD_BEGIN
  std::lock_guard<std::mutex> lock_(ctx->mu);
  auto it = std::find(ctx->scenes.begin(), ctx->scenes.end(), scene);
  if (it == ctx->scenes.end()) return ctx->fail(VRT_ERR_INVALID, "unknown scene handle");
  if (scene->uses > 0)
    return ctx->fail(VRT_ERR_INVALID, "scene is used by %llu queued or running requests",
                     (unsigned long long)scene->uses);
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  if (!scene->evicted) ctx->unreserve(scene->bytes);  // an evicted scene holds no budget
  ctx->scenes.erase(it);
  delete scene;
  return VRT_OK;
  GUARD_END(ctx)
}

int vrt_scene_restore(vrt_ctx* ctx, vrt_scene* scene) {
  if (!ctx || !scene) return ctx ? ctx->fail(VRT_ERR_INVALID, "null argument") : VRT_ERR_INVALID;
  GUARD_BEGIN
  std::lock_guard<std::mutex> lock_(ctx->mu);
  if (std::find(ctx->scenes.begin(), ctx->scenes.end(), scene) == ctx->scenes.end())
    return ctx->fail(VRT_ERR_INVALID, "unknown scene handle");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  ScenePins pins(ctx);
  return pins.add(scene);
  GUARD_END(ctx)
}

int vrt_scene_resident(const vrt_scene* scene, int32_t* resident) {
  if (!scene || !resident) return VRT_ERR_INVALID;
  *resident = scene->evicted ? 0 : 1;
  return VRT_OK;
}

int vrt_scene_edit_counts(const vrt_scene* scene, uint64_t* in_place, uint64_t* rebuilt) {
  if (!scene) return VRT_ERR_INVALID;
  if (in_place) *in_place = scene->edits_in_place;
  if (rebuilt) *rebuilt = scene->edits_rebuilt;
  return VRT_OK;
}

int vrt_scene_dims(const vrt_scene* scene, uint32_t dims[3], uint64_t* bytes) {
  if (!scene) return VRT_ERR_INVALID;
  if (dims) { dims[0] = scene->nx; dims[1] = scene->ny; dims[2] = scene->nz; }
  if (bytes) *bytes = scene->bytes;
  return VRT_OK;
}

int vrt_scene_bricks(vrt_ctx* ctx, const vrt_scene* scene, uint32_t* bricks, uint32_t* masks) {
  if (!ctx || !scene) return VRT_ERR_INVALID;
  GUARD_BEGIN
  std::lock_guard<std::mutex> lock_(ctx->mu);
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  ScenePins pins(ctx);
  if (int rc = pins.add(scene)) return rc;
  const uint64_t nb = scene->nbricks;
  std::vector<uint32_t> words(nb);
  HIP_TRY(ctx, hipMemcpyAsync(words.data(), scene->d_bricks.ptr, nb * 4, hipMemcpyDeviceToHost, ctx->stream));
  std::vector<uint32_t> pool;
  if (masks) {
    pool.resize(std::max<uint64_t>(scene->nocc, 1) * vrt::POOL_WORDS);
    HIP_TRY(ctx, hipMemcpyAsync(pool.data(), scene->d_pool.ptr, pool.size() * 4, hipMemcpyDeviceToHost,
                                ctx->stream));
  }
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  for (uint64_t i = 0; i < nb; ++i) {
    if (bricks) bricks[i] = words[i];
    if (masks) {
      if (words[i] == 0) std::memset(masks + i * 16, 0, 64);
      else std::memcpy(masks + i * 16, pool.data() + (uint64_t)(words[i] - 1) * vrt::POOL_WORDS, 64);
    }
  }
  return VRT_OK;
  GUARD_END(ctx)
}

// ---------------------------------------------------------------------------
// traversal
// ---------------------------------------------------------------------------
// vrt_tuning.trace_kernel -> trace variant of the one-ray-per-lane kernel:
// 0 nested, 4 single loop with both levels in LDS (with the empty-row skip
// when skip is set), 6 the single loop on global memory; the persistent
// kernels 5 and 7 run 4 for ray lists of 2^31 rays and more.
static int trace_variant(uint32_t kernel, uint32_t skip) {
  switch (kernel) {
    case 4:
    case 5:
    case 7: return skip ? vrt::TV_FASTSKIP : vrt::TV_FAST;
    case 6: return vrt::TV_FASTG;
    default: return vrt::TV_NESTED;
  }
}

// vrt_tuning.pt_kernel -> trace variant of the one-pixel-per-lane path
// tracers (the single-view kernel and the batched forms): 0 nested, 4 single
// loop in LDS, 7 (and 8 in the batched forms) the single loop on global
// memory.
static int pt_variant(uint32_t kernel) {
  switch (kernel) {
    case 4: return vrt::TV_FAST;
    case 7:
    case 8: return vrt::TV_FASTG;
    default: return vrt::TV_NESTED;
  }
}

hipError_t vrt_ctx::launch_pathtrace(const vrt::SceneDev& s0, const vrt_uniforms& u, float fclamp,
                                     uint32_t spp, void* radiance, void* albedo, void* gbuf) {
  const vrt::SceneDev s = tuned(s0);
  if (tuning.pt_kernel == 8)
    return vrt::launch_pathtrace_wave(stream, s, u, fclamp, spp, radiance, albedo, gbuf,
                                      (unsigned long long*)d_counters.ptr, (uint32_t*)d_queue.ptr + vrt::QUEUE_WORDS,
                                      tuning.pt_blocks ? (int)tuning.pt_blocks : cus * 16,
                                      (int)tuning.pt_min_refill, (int)tuning.pt_shade_min, (int)tuning.pt_waves,
                                      (int)tuning.pt_tile);
  return vrt::launch_pathtrace(stream, s, u, fclamp, spp, radiance, albedo, gbuf,
                               (unsigned long long*)d_counters.ptr, pt_variant(tuning.pt_kernel),
                               (int)tuning.pt_waves, (int)tuning.pt_tile);
}

hipError_t vrt_ctx::launch_trace(const vrt::SceneDev& s0, const float* rays, uint64_t n, float tmax,
                                 void* hits, unsigned long long* counters, uint32_t* queue,
                                 bool clear_queue) {
  const vrt::SceneDev s = tuned(s0);
  if (!counters) counters = (unsigned long long*)d_counters.ptr;
  if (!queue) queue = (uint32_t*)d_queue.ptr;
  if (tuning.trace_kernel == 7 && n < 0x80000000ull)
    return vrt::launch_trace_wave(stream, s, rays, (uint32_t)n, tmax, hits, counters, queue,
                                  (int)tuning.trace_blocks, (int)tuning.trace_min_refill, (int)tuning.trace_steps,
                                  (int)tuning.trace_waves, clear_queue);
  if (tuning.trace_kernel == 5 && n < 0x80000000ull)
    return vrt::launch_trace_persistent(stream, s, rays, (uint32_t)n, tmax, hits, counters, queue,
                                        (int)tuning.trace_blocks, (int)tuning.trace_steps,
                                        (int)tuning.trace_min_refill,
                                        tuning.empty_skip ? vrt::TV_FASTSKIP : vrt::TV_FAST,
                                        (int)tuning.trace_waves, clear_queue);
  return vrt::launch_trace_rays(stream, s, rays, n, tmax, hits, counters, trace_blocks,
                                trace_variant(tuning.trace_kernel, tuning.empty_skip), (int)tuning.trace_waves);
}

int vrt_trace_rays_device(vrt_ctx* ctx, const vrt_scene* scene, const float* d_rays, uint64_t n,
                          float tmax, vrt_hit* d_hits, float* ms) {
  if (!ctx || !scene || (n && (!d_rays || !d_hits)))
    return ctx ? ctx->fail(VRT_ERR_INVALID, "null argument") : VRT_ERR_INVALID;
  GUARD_BEGIN
  std::lock_guard<std::mutex> lock_(ctx->mu);
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  if (n == 0) { if (ms) *ms = 0.0f; return VRT_OK; }
  ScenePins pins(ctx);
  if (int rc = pins.add(scene)) return rc;
  const vrt::SceneDev s = scene->dev();
  HIP_TRY(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
  HIP_TRY(ctx, ctx->launch_trace(s, d_rays, n, tmax, d_hits));
  HIP_TRY(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
  if (ms) {
    HIP_TRY(ctx, hipEventSynchronize(ctx->ev[1]));
    HIP_TRY(ctx, hipEventElapsedTime(ms, ctx->ev[0], ctx->ev[1]));
  }
  return VRT_OK;
  GUARD_END(ctx)
}

int vrt_trace_rays_device_noalloc(vrt_ctx* ctx, const vrt_scene* scene, const float* d_rays, uint64_t n,
                                  float tmax, vrt_hit* d_hits) {
  if (!ctx || !scene || (n && (!d_rays || !d_hits)))
    return ctx ? ctx->fail(VRT_ERR_INVALID, "null argument") : VRT_ERR_INVALID;
  if (n > VRT_NOALLOC_MAX_RAYS)
    return ctx->fail(VRT_ERR_INVALID,
                     "vrt_trace_rays_device_noalloc: %llu rays exceed its maximum of %llu (the ray-queue head "
                     "in its fixed scratch is 32-bit)",
                     (unsigned long long)n, (unsigned long long)VRT_NOALLOC_MAX_RAYS);
  GUARD_BEGIN
  std::lock_guard<std::mutex> lock_(ctx->mu);
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  if (n == 0) return VRT_OK;
  if (!ctx->noalloc_scratch)
    return ctx->fail(VRT_ERR_DEVICE, "hipGetSymbolAddress gave no scratch address on device %d", ctx->device);
  if (scene->evicted)  // restoring it would allocate
    return ctx->fail(VRT_ERR_INVALID, "scene is evicted to host memory; vrt_scene_restore it first");
  // the counters and the queue head start at zero; the last call's values
  // are still there (every call on this GPU shares the array)
  HIP_TRY(ctx, vrt::launch_zero_words(ctx->stream, ctx->noalloc_scratch, NOALLOC_SCRATCH_WORDS));
  const vrt::SceneDev s = scene->dev();
  HIP_TRY(ctx, ctx->launch_trace(s, d_rays, n, tmax, d_hits, (unsigned long long*)ctx->noalloc_scratch,
                                 (uint32_t*)ctx->noalloc_scratch + NOALLOC_QUEUE_WORD, false));
  return VRT_OK;
  GUARD_END(ctx)
}

int vrt_device_alloc(vrt_ctx* ctx, uint64_t bytes, void** dptr) {
  if (!ctx || !dptr || bytes == 0) return ctx ? ctx->fail(VRT_ERR_INVALID, "bad argument") : VRT_ERR_INVALID;
  GUARD_BEGIN
  std::lock_guard<std::mutex> lock_(ctx->mu);
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  const int rc = ctx->reserve(bytes);
  if (rc) return rc;
  void* p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess || !p) {
    ctx->unreserve(bytes);
    return ctx->fail(VRT_ERR_NO_MEMORY, "hipMalloc of %llu bytes failed", (unsigned long long)bytes);
  }
  ctx->user_buffers[p] = bytes;
  *dptr = p;
  return VRT_OK;
  GUARD_END(ctx)
}

int vrt_device_free(vrt_ctx* ctx, void* dptr) {
  if (!ctx) return VRT_ERR_INVALID;
  if (!dptr) return VRT_OK;
  GUARD_BEGIN
  std::lock_guard<std::mutex> lock_(ctx->mu);
  auto it = ctx->user_buffers.find(dptr);
  if (it == ctx->user_buffers.end()) return ctx->fail(VRT_ERR_INVALID, "not a vrt_device_alloc buffer");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  HIP_TRY(ctx, hipFree(dptr));
  ctx->unreserve(it->second);
  ctx->user_buffers.erase(it);
  return VRT_OK;
  GUARD_END(ctx)
}

int vrt_copy(vrt_ctx* ctx, void* dst, const void* src, uint64_t bytes, uint32_t kind) {
  if (!ctx || (bytes && (!dst || !src)) || kind < VRT_COPY_TO_DEVICE || kind > VRT_COPY_DEVICE)
    return ctx ? ctx->fail(VRT_ERR_INVALID, "bad argument") : VRT_ERR_INVALID;
  GUARD_BEGIN
  std::lock_guard<std::mutex> lock_(ctx->mu);
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  if (bytes == 0) return VRT_OK;
  const hipMemcpyKind k = kind == VRT_COPY_TO_DEVICE ? hipMemcpyHostToDevice
                          : kind == VRT_COPY_TO_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
  HIP_TRY(ctx, hipMemcpyAsync(dst, src, bytes, k, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return VRT_OK;
  GUARD_END(ctx)
}

int vrt_trace_rays(vrt_ctx* ctx, const vrt_scene* scene, const float* rays, uint64_t n, float tmax,
                   vrt_hit* hits) {
  if (!ctx || !scene || (n && (!rays || !hits)))
    return ctx ? ctx->fail(VRT_ERR_INVALID, "null argument") : VRT_ERR_INVALID;
  GUARD_BEGIN
  std::lock_guard<std::mutex> lock_(ctx->mu);
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  if (n == 0) return VRT_OK;
  ScenePins pins(ctx);
  if (int rc = pins.add(scene)) return rc;
  HIP_TRY(ctx, ctx->scratch_rays.ensure(n * 24));
  HIP_TRY(ctx, ctx->scratch_hits.ensure(n * 16));
  HIP_TRY(ctx, hipMemcpyAsync(ctx->scratch_rays.ptr, rays, n * 24, hipMemcpyHostToDevice, ctx->stream));
  const vrt::SceneDev s = scene->dev();
  HIP_TRY(ctx, ctx->launch_trace(s, (const float*)ctx->scratch_rays.ptr, n, tmax, ctx->scratch_hits.ptr));
  HIP_TRY(ctx, hipMemcpyAsync(hits, ctx->scratch_hits.ptr, n * 16, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return VRT_OK;
  GUARD_END(ctx)
}

// ---------------------------------------------------------------------------
// frames
// ---------------------------------------------------------------------------
int vrt_ctx::validate_view(const vrt_view& v) {
  vrt_ctx* ctx = this;
  if (!v.scene) return ctx->fail(VRT_ERR_INVALID, "view without scene");
  if (std::find(ctx->scenes.begin(), ctx->scenes.end(), v.scene) == ctx->scenes.end())
    return ctx->fail(VRT_ERR_INVALID, "view references an unknown scene");
  const float w = v.u.params1[0], h = v.u.params1[1];
  if (!(w >= 1.0f && h >= 1.0f && w <= 16384.0f && h <= 16384.0f) || w != std::floor(w) ||
      h != std::floor(h))
    return ctx->fail(VRT_ERR_INVALID, "bad view size %gx%g", (double)w, (double)h);
  if (v.spp < 1 || v.spp > 4096) return ctx->fail(VRT_ERR_INVALID, "spp must be in [1, 4096]");
  if (v.accum_format > 1) return ctx->fail(VRT_ERR_INVALID, "bad accum_format");
  if (v.u.params0[1] > 64) return ctx->fail(VRT_ERR_INVALID, "bounces must be <= 64");
  return VRT_OK;
}

static int copy_out(vrt_ctx* ctx, ViewState& vs, const vrt_view& v, vrt_frame* fo, int parity,
                    bool denoised) {
  const uint64_t px = (uint64_t)vs.w * vs.h;
  const uint64_t abytes = px * (vs.fmt == VRT_ACCUM_RGBA32F ? 16 : 8);
  if (fo->radiance)
    HIP_TRY(ctx, hipMemcpyAsync(fo->radiance, vs.radiance.ptr, px * 8, hipMemcpyDeviceToHost, ctx->stream));
  if (fo->albedo)
    HIP_TRY(ctx, hipMemcpyAsync(fo->albedo, vs.albedo.ptr, px * 4, hipMemcpyDeviceToHost, ctx->stream));
  if (fo->gbuf)
    HIP_TRY(ctx, hipMemcpyAsync(fo->gbuf, vs.gbuf[parity].ptr, px * 16, hipMemcpyDeviceToHost, ctx->stream));
  if (fo->accum)
    HIP_TRY(ctx, hipMemcpyAsync(fo->accum, vs.accum[parity].ptr, abytes, hipMemcpyDeviceToHost, ctx->stream));
  if (fo->denoised && denoised)
    HIP_TRY(ctx, hipMemcpyAsync(fo->denoised, vs.den[0].ptr, px * 8, hipMemcpyDeviceToHost, ctx->stream));
  if (fo->hdr) {
    // main.js:629-632: the denoised target when denoise is on, else the
    // accumulation (rgba32f only when that is its format)
    const void* illum = denoised ? vs.den[0].ptr : vs.accum[parity].ptr;
    const int fmt = denoised ? 0 : (int)vs.fmt;
    HIP_TRY(ctx, ctx->scratch_hdr.ensure(px * 12));
    HIP_TRY(ctx, vrt::launch_hdr(ctx->stream, (uint32_t)px, illum, fmt, vs.albedo.ptr,
                                 (float*)ctx->scratch_hdr.ptr));
    HIP_TRY(ctx, hipMemcpyAsync(fo->hdr, ctx->scratch_hdr.ptr, px * 12, hipMemcpyDeviceToHost, ctx->stream));
  }
  if (fo->ldr) {
    // present.wgsl:40-55 on the same illumination as the HDR capture
    const void* illum = denoised ? vs.den[0].ptr : vs.accum[parity].ptr;
    const int fmt = denoised ? 0 : (int)vs.fmt;
    HIP_TRY(ctx, ctx->scratch_ldr.ensure(px * 4));
    HIP_TRY(ctx, vrt::launch_present(ctx->stream, (uint32_t)px, illum, fmt, vs.albedo.ptr, v.u.params1[2],
                                     ctx->scratch_ldr.ptr));
    HIP_TRY(ctx, hipMemcpyAsync(fo->ldr, ctx->scratch_ldr.ptr, px * 4, hipMemcpyDeviceToHost, ctx->stream));
  }
  (void)v;
  return VRT_OK;
}

int vrt_render(vrt_ctx* ctx, const vrt_view* views, uint32_t nviews, vrt_frame* outs) {
  if (!ctx || (nviews && !views)) return ctx ? ctx->fail(VRT_ERR_INVALID, "null argument") : VRT_ERR_INVALID;
  GUARD_BEGIN
  std::lock_guard<std::mutex> lock_(ctx->mu);
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  for (uint32_t k = 0; k < nviews; ++k) {
    int rc = ctx->validate_view(views[k]);
    if (rc) return rc;
    for (uint32_t j = 0; j < k; ++j)
      if (views[j].view_id == views[k].view_id)
        return ctx->fail(VRT_ERR_INVALID, "view id %u appears twice in one call", views[k].view_id);
  }
  {
    int rc = ctx->order_after_serve();
    if (rc) return rc;
  }
  ScenePins pins(ctx);
  for (uint32_t k = 0; k < nviews; ++k)
    if (int rc = pins.add(views[k].scene)) return rc;
  std::vector<ViewState*> states(nviews);
  for (uint32_t k = 0; k < nviews; ++k) {
    const vrt_view& v = views[k];
    int rc = ctx->view_state(v.view_id, (uint32_t)v.u.params1[0], (uint32_t)v.u.params1[1],
                             v.accum_format, &states[k]);
    if (rc) return rc;
  }
  uint64_t before[4] = {0, 0, 0, 0};
  const bool want_counts = outs != nullptr;
  if (want_counts) {
    int rc = ctx->counters_to_host(before);
    if (rc) return rc;
  }
  // timing events go into a ring and are read later (vrt_stats_get), so a
  // frame without host outputs never waits for the device
  if (ctx->ring_count == vrt_ctx::RING) {
    int rc = ctx->harvest(1);
    if (rc) return rc;
  }
  hipEvent_t* ev = ctx->ring[(ctx->ring_head) % vrt_ctx::RING].e;
  ctx->ring_head++;
  ctx->ring_count++;
  if (nviews > 1) {
    // batched form: one launch per pass for all views
    int rc = ctx->render_batch(ctx->lane0, views, nviews, states.data(), ev, nullptr, nullptr);
    if (rc) return rc;
  } else {
  HIP_TRY(ctx, hipEventRecord(ev[0], ctx->stream));
  for (uint32_t k = 0; k < nviews; ++k) {
    const vrt_view& v = views[k];
    ViewState& vs = *states[k];
    const int parity = (int)(vs.frames & 1u);
    HIP_TRY(ctx, ctx->launch_pathtrace(v.scene->dev(), v.u, v.firefly_clamp, v.spp, vs.radiance.ptr,
                                       vs.albedo.ptr, vs.gbuf[parity].ptr));
  }
  HIP_TRY(ctx, hipEventRecord(ev[1], ctx->stream));
  for (uint32_t k = 0; k < nviews; ++k) {
    const vrt_view& v = views[k];
    ViewState& vs = *states[k];
    const int parity = (int)(vs.frames & 1u);
    HIP_TRY(ctx, vrt::launch_temporal(ctx->stream, v.u, vs.radiance.ptr, vs.gbuf[parity].ptr,
                                      vs.accum[1 - parity].ptr, vs.gbuf[1 - parity].ptr,
                                      vs.accum[parity].ptr, (int)vs.fmt));
  }
  HIP_TRY(ctx, hipEventRecord(ev[2], ctx->stream));
  for (uint32_t k = 0; k < nviews; ++k) {
    const vrt_view& v = views[k];
    ViewState& vs = *states[k];
    if (!(v.u.params0[2] & VRT_FLAG_DENOISE)) continue;
    const int parity = (int)(vs.frames & 1u);
    // main.js:502-507: accum -> dA (1) -> dB (2) -> dA (4)
    HIP_TRY(ctx, vrt::launch_atrous(ctx->stream, v.u, vs.accum[parity].ptr, (int)vs.fmt, vs.gbuf[parity].ptr, 1,
                                    vs.den[0].ptr));
    HIP_TRY(ctx, vrt::launch_atrous(ctx->stream, v.u, vs.den[0].ptr, 0, vs.gbuf[parity].ptr, 2, vs.den[1].ptr));
    HIP_TRY(ctx, vrt::launch_atrous(ctx->stream, v.u, vs.den[1].ptr, 0, vs.gbuf[parity].ptr, 4, vs.den[0].ptr));
  }
  HIP_TRY(ctx, hipEventRecord(ev[3], ctx->stream));
  }
  for (uint32_t k = 0; k < nviews && outs; ++k) {
    ViewState& vs = *states[k];
    const int parity = (int)(vs.frames & 1u);
    int rc = copy_out(ctx, vs, views[k], &outs[k], parity, (views[k].u.params0[2] & VRT_FLAG_DENOISE) != 0);
    if (rc) return rc;
  }
  if (outs) {
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    int rc = ctx->harvest(vrt_ctx::RING);
    if (rc) return rc;
  }
  for (uint32_t k = 0; k < nviews; ++k) {
    states[k]->frames++;
    states[k]->lru = ++ctx->clock;
  }
  ctx->frames += nviews;
  if (want_counts) {
    uint64_t after[4];
    int rc = ctx->counters_to_host(after);
    if (rc) return rc;
    // per-view split is not tracked by the shared counters: the totals of the
    // call go to the first frame record when several views are rendered
    for (uint32_t k = 0; k < nviews; ++k) {
      outs[k].rays = k == 0 ? after[0] - before[0] : 0;
      outs[k].brick_steps = k == 0 ? after[1] - before[1] : 0;
      outs[k].voxel_steps = k == 0 ? after[2] - before[2] : 0;
    }
  }
  return VRT_OK;
  GUARD_END(ctx)
}

int vrt_view_release(vrt_ctx* ctx, uint32_t view_id) {
  if (!ctx) return VRT_ERR_INVALID;
  GUARD_BEGIN
  std::lock_guard<std::mutex> lock_(ctx->mu);
  auto it = ctx->views.find(view_id);
  if (it == ctx->views.end()) return ctx->fail(VRT_ERR_INVALID, "unknown view id %u", view_id);
  if (ctx->server && ctx->server->vuse.count(view_id))
    return ctx->fail(VRT_ERR_INVALID, "view id %u has queued or running requests", view_id);
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  ctx->unreserve(it->second.bytes);
  ctx->views.erase(it);
  return VRT_OK;
  GUARD_END(ctx)
}

int vrt_host_alloc(vrt_ctx* ctx, uint64_t bytes, void** ptr) {
  if (!ctx || !ptr || bytes == 0) return ctx ? ctx->fail(VRT_ERR_INVALID, "host_alloc arguments") : VRT_ERR_INVALID;
  GUARD_BEGIN
  std::lock_guard<std::mutex> lock_(ctx->mu);
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  *ptr = nullptr;
  const hipError_t e = hipHostMalloc(ptr, bytes, hipHostMallocDefault);
  if (e != hipSuccess) {
    *ptr = nullptr;
    return ctx->fail(VRT_ERR_NO_MEMORY, "host_alloc %llu B: %s", (unsigned long long)bytes, hipGetErrorString(e));
  }
  return VRT_OK;
  GUARD_END(ctx)
}

int vrt_host_free(vrt_ctx* ctx, void* ptr) {
  if (!ptr) return VRT_OK;
  if (!ctx) return hipHostFree(ptr) == hipSuccess ? VRT_OK : VRT_ERR_DEVICE;
  GUARD_BEGIN
  std::lock_guard<std::mutex> lock_(ctx->mu);
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  HIP_TRY(ctx, hipHostFree(ptr));
  return VRT_OK;
  GUARD_END(ctx)
}

int vrt_temporal_pass(vrt_ctx* ctx, const vrt_uniforms* u, const uint16_t* radiance,
                      const float* gbuf_cur, const void* accum_prev, const float* gbuf_prev,
                      uint32_t accum_format, void* accum_out) {
  if (!ctx || !u || !radiance || !gbuf_cur || !accum_prev || !gbuf_prev || !accum_out || accum_format > 1)
    return ctx ? ctx->fail(VRT_ERR_INVALID, "bad argument") : VRT_ERR_INVALID;
  GUARD_BEGIN
  std::lock_guard<std::mutex> lock_(ctx->mu);
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  const uint64_t px = (uint64_t)u->params1[0] * (uint64_t)u->params1[1];
  const uint64_t ab = px * (accum_format ? 16 : 8);
  vrt::DeviceBuffer r, gc, ap, gp, ao;
  HIP_TRY(ctx, r.alloc(px * 8));
  HIP_TRY(ctx, gc.alloc(px * 16));
  HIP_TRY(ctx, ap.alloc(ab));
  HIP_TRY(ctx, gp.alloc(px * 16));
  HIP_TRY(ctx, ao.alloc(ab));
  HIP_TRY(ctx, hipMemcpyAsync(r.ptr, radiance, px * 8, hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(ctx, hipMemcpyAsync(gc.ptr, gbuf_cur, px * 16, hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(ctx, hipMemcpyAsync(ap.ptr, accum_prev, ab, hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(ctx, hipMemcpyAsync(gp.ptr, gbuf_prev, px * 16, hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(ctx, vrt::launch_temporal(ctx->stream, *u, r.ptr, gc.ptr, ap.ptr, gp.ptr, ao.ptr, (int)accum_format));
  HIP_TRY(ctx, hipMemcpyAsync(accum_out, ao.ptr, ab, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return VRT_OK;
  GUARD_END(ctx)
}

int vrt_atrous_pass(vrt_ctx* ctx, const vrt_uniforms* u, const void* src, uint32_t src_format,
                    const float* gbuf, int32_t step, uint16_t* dst) {
  if (!ctx || !u || !src || !gbuf || !dst || src_format > 1 || (step != 1 && step != 2 && step != 4))
    return ctx ? ctx->fail(VRT_ERR_INVALID, "bad argument (step must be 1, 2 or 4)") : VRT_ERR_INVALID;
  GUARD_BEGIN
  std::lock_guard<std::mutex> lock_(ctx->mu);
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  const uint64_t px = (uint64_t)u->params1[0] * (uint64_t)u->params1[1];
  const uint64_t sb = px * (src_format ? 16 : 8);
  vrt::DeviceBuffer s, g, d;
  HIP_TRY(ctx, s.alloc(sb));
  HIP_TRY(ctx, g.alloc(px * 16));
  HIP_TRY(ctx, d.alloc(px * 8));
  HIP_TRY(ctx, hipMemcpyAsync(s.ptr, src, sb, hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(ctx, hipMemcpyAsync(g.ptr, gbuf, px * 16, hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(ctx, vrt::launch_atrous(ctx->stream, *u, s.ptr, (int)src_format, g.ptr, step, d.ptr));
  HIP_TRY(ctx, hipMemcpyAsync(dst, d.ptr, px * 8, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return VRT_OK;
  GUARD_END(ctx)
}

int vrt_tuning_get(vrt_ctx* ctx, vrt_tuning* out) {
  if (!ctx || !out) return VRT_ERR_INVALID;
  std::lock_guard<std::mutex> lock_(ctx->mu);
  *out = ctx->tuning;
  return VRT_OK;
}

int vrt_tuning_set(vrt_ctx* ctx, const vrt_tuning* in) {
  if (!ctx || !in) return VRT_ERR_INVALID;
  std::lock_guard<std::mutex> lock_(ctx->mu);
  const uint32_t tk = in->trace_kernel, pk = in->pt_kernel;
  if (!(tk == 0 || tk == 4 || tk == 5 || tk == 6 || tk == 7) || !(pk == 0 || pk == 4 || pk == 7 || pk == 8) ||
      in->trace_blocks > (1u << 20) || in->pt_blocks > (1u << 20) || in->trace_steps > 1024 ||
      in->trace_min_refill > 64 || in->pt_min_refill > 64 ||
      !(in->pt_waves == 0 || in->pt_waves == 4 || in->pt_waves == 5 || (in->pt_waves == 6 && pk == 8)) ||
      in->batch_mode > 2 ||
      in->batch_blocks > (1u << 16) || in->dda_batch > 64 ||
      in->empty_skip > 1 || in->pt_order > 3 || in->pt_tile > 5 || !(in->trace_waves == 0 || in->trace_waves == 6) ||
      in->pt_shade_min > 64 || in->batch_tile > 5)
    return ctx->fail(VRT_ERR_INVALID, "tuning value out of range");
  ctx->tuning = *in;
  return VRT_OK;
}

int vrt_autotune(vrt_ctx* ctx, const vrt_scene* scene, uint64_t nrays, const vrt_uniforms* u,
                 vrt_tuning* chosen) {
  if (!ctx || !scene) return ctx ? ctx->fail(VRT_ERR_INVALID, "null argument") : VRT_ERR_INVALID;
  GUARD_BEGIN
  std::lock_guard<std::mutex> lock_(ctx->mu);
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  ScenePins pins(ctx);
  if (int rc = pins.add(scene)) return rc;
  const vrt::SceneDev s = scene->dev();
  uint64_t before[4];
  int rc = ctx->counters_to_host(before);
  if (rc) return rc;
  // On every return: the sweep's rays leave the statistics, and the tuning
  // is the one in effect before the call unless the sweep completed.
  struct Restore {
    vrt_ctx* c;
    const uint64_t* cnt;
    vrt_tuning keep;
    ~Restore() {
      c->tuning = keep;
      if (hipMemcpyAsync(c->d_counters.ptr, cnt, 4 * sizeof(uint64_t), hipMemcpyHostToDevice, c->stream) ==
          hipSuccess)
        (void)hipStreamSynchronize(c->stream);
    }
  } restore{ctx, before, ctx->tuning};
  vrt_tuning best = ctx->tuning;
  // a warm-up launch, then the fastest of three rounds of three launches:
  // a transient slowdown of one round (clocks, a neighbour on the device)
  // does not decide the choice
  auto time_it = [&](auto&& launch, float* out_ms) -> int {
    HIP_TRY(ctx, launch());  // warm-up
    float best_round = 1e30f;
    for (int round = 0; round < 3; ++round) {
      HIP_TRY(ctx, hipEventRecord(ctx->ev[4], ctx->stream));
      for (int r = 0; r < 3; ++r) HIP_TRY(ctx, launch());
      HIP_TRY(ctx, hipEventRecord(ctx->ev[5], ctx->stream));
      HIP_TRY(ctx, hipEventSynchronize(ctx->ev[5]));
      float ms = 0.0f;
      HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev[4], ctx->ev[5]));
      best_round = ms < best_round ? ms : best_round;
    }
    *out_ms = best_round;
    return VRT_OK;
  };
  if (nrays > 0) {
    // random rays inside the grid's box, directions spread over the sphere
    const uint64_t n = std::min<uint64_t>(nrays, 1ull << 24);
    std::vector<float> rays(n * 6);
    uint64_t z = 0x9E3779B97F4A7C15ull;
    auto u01 = [&z]() {
      z += 0x9E3779B97F4A7C15ull;
      uint64_t x = z;
      x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
      x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
      return (float)((x ^ (x >> 31)) >> 40) * (1.0f / 16777216.0f);
    };
    for (uint64_t i = 0; i < n; ++i) {
      float* r = rays.data() + i * 6;
      r[0] = u01() * scene->nx * VRT_VOXEL_SIZE;
      r[1] = u01() * scene->ny * VRT_VOXEL_SIZE;
      r[2] = u01() * scene->nz * VRT_VOXEL_SIZE;
      const float cz = u01() * 2.0f - 1.0f, ph = u01() * 6.2831853f;
      const float rr = std::sqrt(std::max(0.0f, 1.0f - cz * cz));
      r[3] = rr * std::cos(ph);
      r[4] = cz;
      r[5] = rr * std::sin(ph);
    }
    HIP_TRY(ctx, ctx->scratch_rays.ensure(n * 24));
    HIP_TRY(ctx, ctx->scratch_hits.ensure(n * 16));
    HIP_TRY(ctx, hipMemcpyAsync(ctx->scratch_rays.ptr, rays.data(), n * 24, hipMemcpyHostToDevice, ctx->stream));
    float best_ms = 1e30f;
    auto try_trace = [&](const vrt_tuning& cand) -> int {
      ctx->tuning = cand;
      float ms = 0;
      const int r = time_it([&] { return ctx->launch_trace(s, (const float*)ctx->scratch_rays.ptr, n, 1e4f,
                                                            ctx->scratch_hits.ptr); }, &ms);
      if (r) return r;
      if (ms < best_ms) {
        best_ms = ms;
        best = cand;
      }
      return VRT_OK;
    };
    vrt_tuning c = best;
    c.trace_kernel = 0;
    c.dda_batch = 1;
    c.trace_waves = 0;
    if ((rc = try_trace(c))) return rc;
    // one-ray-per-lane single-loop variants over their brick entry batch and
    // register budget
    for (uint32_t k : {4u, 6u})
      for (uint32_t db : {1u, 8u, 16u, 32u})
        for (uint32_t wv : {0u, 6u}) {
          c.trace_kernel = k;
          c.dda_batch = db;
          c.trace_waves = wv;
          if ((rc = try_trace(c))) return rc;
        }
    // step-level refill over its launch shape and pass thresholds
    c.trace_kernel = 7;
    c.dda_batch = 1;
    c.trace_waves = 0;
    for (uint32_t b : {1024u, 1280u, 1536u})
      for (uint32_t mw : {12u, 16u, 20u})
        for (uint32_t fm : {24u, 32u}) {
          c.trace_blocks = b;
          c.trace_min_refill = mw;
          c.trace_steps = fm;
          if ((rc = try_trace(c))) return rc;
        }
    // the persistent ray queue over its launch shape
    c.trace_kernel = 5;
    for (uint32_t db : {1u, 16u})
      for (uint32_t wv : {0u, 6u})
        for (uint32_t b : {1024u, 2048u, 4096u, 8192u})
          for (uint32_t st : {1u, 4u, 8u, 16u})
            for (uint32_t rf : {1u, 16u, 32u}) {
              c.dda_batch = db;
              c.trace_waves = wv;
              c.trace_blocks = b;
              c.trace_steps = st;
              c.trace_min_refill = rf;
              if ((rc = try_trace(c))) return rc;
            }
  }
  if (u != nullptr) {
    // path tracer on the given view (its size, pose and bounce count)
    const uint64_t px = (uint64_t)u->params1[0] * (uint64_t)u->params1[1];
    if (px == 0 || px > (1ull << 26)) return ctx->fail(VRT_ERR_INVALID, "autotune view size");
    vrt::DeviceBuffer rad, alb, gb;
    HIP_TRY(ctx, rad.alloc(px * 8));
    HIP_TRY(ctx, alb.alloc(px * 4));
    HIP_TRY(ctx, gb.alloc(px * 16));
    auto launch = [&] { return ctx->launch_pathtrace(s, *u, 48.0f, 1, rad.ptr, alb.ptr, gb.ptr); };
    float best_ms = 1e30f;
    // margin < 1: a candidate must beat the best by that factor (tile orders:
    // the tuning view is one pose, and the default order -- costly bottom
    // rows first -- is the better one over a camera path unless the pose
    // says otherwise by more than the noise)
    auto try_pt = [&](const vrt_tuning& cand, float margin) -> int {
      ctx->tuning = cand;
      float ms = 0;
      const int r = time_it(launch, &ms);
      if (r) return r;
      if (ms < best_ms * margin) {
        best_ms = ms;
        best = cand;
      }
      return VRT_OK;
    };
    vrt_tuning c = best;
    c.pt_tile = 0;
    // one-pixel-per-lane variants x register budgets (waves per SIMD)
    for (uint32_t k : {0u, 4u, 7u})
      for (uint32_t wv : {0u, 4u, 5u}) {
        c.pt_kernel = k;
        c.pt_waves = wv;
        if ((rc = try_pt(c, 1.0f))) return rc;
      }
    // the global-memory variant on 8x8 tiles of one wave; refill tiles of
    // the single-loop variants (32x32, 32x16; 16x16, 16x8 on one-wave blocks)
    for (uint32_t k : {4u, 7u})
      for (uint32_t form : {1u, 2u, 3u, 4u, 5u})
        for (uint32_t wv : {0u, 4u, 5u}) {
          if ((form == 1u || form >= 4u) && k != 7u) continue;
          c.pt_kernel = k;
          c.pt_waves = wv;
          c.pt_tile = form;
          if ((rc = try_pt(c, 1.0f))) return rc;
        }
    // tile order of the best of these
    c = best;
    for (uint32_t o : {0u, 1u, 2u, 3u}) {
      if (o == best.pt_order) continue;
      c.pt_order = o;
      if ((rc = try_pt(c, 0.98f))) return rc;
    }
    // step-level refill: level-switch and shading thresholds, DDA memory
    c = best;
    c.pt_kernel = 8;
    c.pt_blocks = 0;
    c.pt_waves = 4;
    c.pt_order = 2;
    for (uint32_t form : {0u, 1u})
      for (uint32_t mw : {8u, 12u, 16u, 24u})
        for (uint32_t sm : {32u, 40u}) {
          c.pt_tile = form;
          c.pt_min_refill = mw;
          c.pt_shade_min = sm;
          if ((rc = try_pt(c, 1.0f))) return rc;
        }
    if (best.pt_kernel == 8) {
      c = best;
      c.pt_order = 0;
      if ((rc = try_pt(c, 0.98f))) return rc;
    }
  }
  restore.keep = best;
  if (chosen) *chosen = best;
  return VRT_OK;
  GUARD_END(ctx)
}

int vrt_autotune_batch(vrt_ctx* ctx, const vrt_view* views, uint32_t nviews, uint32_t modes,
                       vrt_tuning* chosen) {
  if (!ctx || !views || nviews < 2)
    return ctx ? ctx->fail(VRT_ERR_INVALID, "autotune_batch needs 2 or more views") : VRT_ERR_INVALID;
  if (modes == 0u) modes = 7u;
  if ((modes & 7u) == 0u || modes > 7u)
    return ctx->fail(VRT_ERR_INVALID, "autotune_batch modes: bits 0..2 (batch_mode 0..2)");
  GUARD_BEGIN
  std::lock_guard<std::mutex> lock_(ctx->mu);
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  for (uint32_t k = 0; k < nviews; ++k) {
    int rc = ctx->validate_view(views[k]);
    if (rc) return rc;
  }
  {
    int rc = ctx->order_after_serve();
    if (rc) return rc;
  }
  ScenePins pins(ctx);
  for (uint32_t k = 0; k < nviews; ++k)
    if (int rc = pins.add(views[k].scene)) return rc;
  // the sweep renders into view states of its own: the callers' histories
  // are not touched
  std::vector<ViewState> tmp(nviews);
  std::vector<ViewState*> states(nviews);
  uint64_t reserved = 0;
  struct Release {  // runs before tmp's buffers are freed
    vrt_ctx* c;
    uint64_t* b;
    ~Release() {
      (void)hipStreamSynchronize(c->stream);
      c->unreserve(*b);
    }
  } release_{ctx, &reserved};
  for (uint32_t k = 0; k < nviews; ++k) {
    const vrt_view& v = views[k];
    int rc = ctx->alloc_view_state(tmp[k], (uint32_t)v.u.params1[0], (uint32_t)v.u.params1[1], v.accum_format,
                                   ctx->stream);
    if (rc) return rc;
    reserved += tmp[k].bytes;
    states[k] = &tmp[k];
  }
  uint64_t before[4];
  int rc = ctx->counters_to_host(before);
  if (rc) return rc;
  const vrt_tuning start = ctx->tuning;
  // On every return: the sweep's rays leave the statistics, the graphs it
  // captured are dropped (recaptured on first use), and the tuning is the
  // one in effect before the call unless the sweep completed.
  struct Restore {
    vrt_ctx* c;
    const uint64_t* cnt;
    vrt_tuning keep;
    ~Restore() {
      c->tuning = keep;
      if (hipMemcpyAsync(c->d_counters.ptr, cnt, 4 * sizeof(uint64_t), hipMemcpyHostToDevice, c->stream) ==
          hipSuccess)
        (void)hipStreamSynchronize(c->stream);
      c->lane0.drop_graphs();
    }
  } restore{ctx, before, start};
  vrt_tuning best = start;
  float best_ms = 1e30f;
  auto time_batch = [&](float* out_ms) -> int {
    int r = ctx->render_batch(ctx->lane0, views, nviews, states.data(), ctx->ev, nullptr, nullptr);  // warm-up
    if (r) return r;
    HIP_TRY(ctx, hipEventRecord(ctx->ev[4], ctx->stream));
    for (int i = 0; i < 3; ++i) {
      r = ctx->render_batch(ctx->lane0, views, nviews, states.data(), ctx->ev, nullptr, nullptr);
      if (r) return r;
    }
    HIP_TRY(ctx, hipEventRecord(ctx->ev[5], ctx->stream));
    HIP_TRY(ctx, hipEventSynchronize(ctx->ev[5]));
    HIP_TRY(ctx, hipEventElapsedTime(out_ms, ctx->ev[4], ctx->ev[5]));
    return VRT_OK;
  };
  // launch mode x path-trace tile form x persistent grid (0 = the default:
  // 8 blocks of 256 threads per CU, or 32 one-wave blocks for the 16-wide
  // refill tiles)
  const uint32_t cus = (uint32_t)ctx->cus;
  for (uint32_t mode : {0u, 1u, 2u})
    for (uint32_t form : {0u, 2u, 3u, 4u, 5u})
      for (uint32_t b : {0u, 1u, 2u}) {
        if (!(modes >> mode & 1u) || (mode == 0u && b != 0u)) continue;
        ctx->tuning = start;
        ctx->tuning.batch_mode = mode;
        // pt 8 keeps its own pt_tile (the single-view kernel's DDA memory);
        // its batch_tile 0 is the step-level refill batch kernel
        if (start.pt_kernel == 8) {
          ctx->tuning.batch_tile = form;
        } else {
          ctx->tuning.pt_tile = form;
        }
        // half or twice the default grid
        ctx->tuning.batch_blocks = b == 0u ? 0u : cus * (form >= 4u ? 32u : 8u) * (b == 1u ? 1u : 4u) / 2u;
        float ms = 0;
        rc = time_batch(&ms);
        if (rc) return rc;
        if (ms < best_ms) {
          best_ms = ms;
          best = ctx->tuning;
        }
      }
  restore.keep = best;
  if (chosen) *chosen = best;
  return VRT_OK;
  GUARD_END(ctx)
}

int vrt_stats_get(vrt_ctx* ctx, vrt_stats* out) {
  if (!ctx || !out) return VRT_ERR_INVALID;
  GUARD_BEGIN
  std::lock_guard<std::mutex> lock_(ctx->mu);
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  uint64_t c[4];
  int rc = ctx->counters_to_host(c);
  if (rc) return rc;
  rc = ctx->harvest(vrt_ctx::RING);
  if (rc) return rc;
  std::memset(out, 0, sizeof(*out));
  out->timed_calls = ctx->timed_calls;
  out->sum_ms_pathtrace = ctx->sum_pathtrace;
  out->sum_ms_temporal = ctx->sum_temporal;
  out->sum_ms_atrous = ctx->sum_atrous;
  out->sum_ms_frame = ctx->sum_frame;
  out->scene_evictions = ctx->scene_evictions;
  out->scene_restores = ctx->scene_restores;
  out->rays = c[0];
  out->brick_steps = c[1];
  out->voxel_steps = c[2];
  out->hits = c[3];
  out->frames = ctx->frames;
  out->gpu_ms_pathtrace = ctx->ms_pathtrace;
  out->gpu_ms_temporal = ctx->ms_temporal;
  out->gpu_ms_atrous = ctx->ms_atrous;
  out->gpu_ms_frame = ctx->ms_frame;
  out->hbm_resident = ctx->resident;
  out->hbm_budget = ctx->budget;
  return VRT_OK;
  GUARD_END(ctx)
}

int vrt_stats_reset(vrt_ctx* ctx) {
  if (!ctx) return VRT_ERR_INVALID;
  GUARD_BEGIN
  std::lock_guard<std::mutex> lock_(ctx->mu);
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  {
    int rc = ctx->order_after_serve();
    if (rc) return rc;
  }
  HIP_TRY(ctx, hipMemsetAsync(ctx->d_counters.ptr, 0, 4 * sizeof(uint64_t), ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  ctx->frames = 0;
  {
    int rc = ctx->harvest(vrt_ctx::RING);
    if (rc) return rc;
  }
  ctx->timed_calls = 0;
  ctx->sum_pathtrace = ctx->sum_temporal = ctx->sum_atrous = ctx->sum_frame = 0.0;
  return VRT_OK;
  GUARD_END(ctx)
}

}  // extern "C"

namespace vrt {

void BatchLane::drop_graphs() {
  for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
  graphs.clear();
}

BatchLane::~BatchLane() {
  drop_graphs();
  for (uint32_t k = 0; k < STAGES; ++k) {
    if (desc_host[k]) (void)hipHostFree(desc_host[k]);
    if (desc_done[k]) (void)hipEventDestroy(desc_done[k]);
  }
  if (own_stream && stream) (void)hipStreamDestroy(stream);
}

hipError_t BatchLane::ensure(uint32_t nviews) {
  if (nviews <= desc_cap) return hipSuccess;
  hipError_t e = hipStreamSynchronize(stream);
  if (e != hipSuccess) return e;
  for (uint32_t k = 0; k < STAGES; ++k) {
    if (desc_host[k]) (void)hipHostFree(desc_host[k]);
    desc_host[k] = nullptr;
  }
  desc_cap = 0;
  drop_graphs();  // they hold the old table's address
  const uint32_t cap = std::max<uint32_t>(nviews, 64);
  for (uint32_t k = 0; k < STAGES; ++k) {
    e = hipHostMalloc(&desc_host[k], BATCH_HEAD_BYTES + cap * sizeof(ViewDesc), hipHostMallocDefault);
    if (e != hipSuccess) return e;
    if (!desc_done[k]) {
      e = hipEventCreateWithFlags(&desc_done[k], hipEventDisableTiming);
      if (e != hipSuccess) return e;
    }
  }
  e = desc_dev.alloc(BATCH_HEAD_BYTES + cap * sizeof(ViewDesc));
  if (e != hipSuccess) return e;
  desc_cap = cap;
  return hipSuccess;
}

}  // namespace vrt

// Batched form: fill one ViewDesc per view (vrt_batch.h) in the lane's pinned
// staging buffer, copy the table to the device, and launch each pass once
// over the tiles of every view.
int vrt_ctx::render_batch(vrt::BatchLane& lane, const vrt_view* views, uint32_t nviews, ViewState** states,
                          hipEvent_t* ev, unsigned long long* const* vcount, float* const* hdr) {
  HIP_TRY(this, lane.ensure(nviews));
  const hipStream_t st = lane.stream;
  const uint32_t stage = lane.desc_next++ % vrt::BatchLane::STAGES;
  HIP_TRY(this, hipEventSynchronize(lane.desc_done[stage]));  // its previous copy has run
  vrt::BatchHead* hh = (vrt::BatchHead*)lane.desc_host[stage];
  vrt::ViewDesc* d = (vrt::ViewDesc*)((char*)lane.desc_host[stage] + vrt::BATCH_HEAD_BYTES);
  uint32_t tiles = 0, max_bit_words = 0, max_pwords = 0;
  bool any_denoise = false, any_hdr = false;
  for (uint32_t k = 0; k < nviews; ++k) {
    const vrt_view& v = views[k];
    ViewState& vs = *states[k];
    const int parity = (int)(vs.frames & 1u);
    std::memset(&d[k], 0, sizeof(d[k]));
    d[k].s = tuned(v.scene->dev());
    d[k].u = v.u;
    d[k].fclamp = v.firefly_clamp;
    d[k].spp = v.spp;
    d[k].fmt = vs.fmt;
    d[k].w = vs.w;
    d[k].h = vs.h;
    d[k].tiles_x = (vs.w + 15u) / 16u;
    d[k].tile0 = tiles;
    d[k].denoise = (v.u.params0[2] & VRT_FLAG_DENOISE) ? 1u : 0u;
    d[k].radiance = vs.radiance.ptr;
    d[k].albedo = vs.albedo.ptr;
    d[k].gbuf_cur = vs.gbuf[parity].ptr;
    d[k].gbuf_prev = vs.gbuf[1 - parity].ptr;
    d[k].accum_cur = vs.accum[parity].ptr;
    d[k].accum_prev = vs.accum[1 - parity].ptr;
    d[k].den0 = vs.den[0].ptr;
    d[k].den1 = vs.den[1].ptr;
    d[k].vcount = vcount ? vcount[k] : nullptr;
    d[k].hdr = hdr ? hdr[k] : nullptr;
    max_bit_words = std::max(max_bit_words, d[k].s.bit_words);
    max_pwords = std::max(max_pwords, d[k].s.pwords);
    any_denoise |= d[k].denoise != 0u;
    any_hdr |= d[k].hdr != nullptr;
    tiles += d[k].tiles_x * ((vs.h + 15u) / 16u);
  }
  // refill forms: the path tracer's own 16 x rth tiles, numbered per view.
  // The step-level refill kernel (pt 8) runs its batched form
  // (k_pathtrace_wave_batch) with batch_tile 0, or pt 7 with 4 waves per
  // SIMD on the tile form batch_tile.
  const bool wave_pt = tuning.pt_kernel == 8;
  const bool wave_batch = wave_pt && tuning.batch_tile == 0;
  const int ptv = pt_variant(tuning.pt_kernel);
  const int pt_waves = wave_pt && !wave_batch ? 4 : (int)tuning.pt_waves;
  const int tile = wave_pt ? (int)tuning.batch_tile : (int)tuning.pt_tile;
  const int rform = wave_batch ? 0 : vrt::refill_form(ptv, tile, max_bit_words, max_pwords);
  const uint32_t wave_blocks = tuning.pt_blocks ? tuning.pt_blocks : (uint32_t)cus * 16u;
  uint64_t patches = 0;
  for (uint32_t k = 0; k < nviews; ++k) patches += ((d[k].w + 7u) / 8u) * ((d[k].h + 7u) / 8u);
  const uint32_t wchunk = vrt::wave_chunk(patches, wave_blocks, nviews);
  uint32_t wchunks = 0;
  for (uint32_t k = 0; k < nviews; ++k) {
    d[k].wchunk0 = wchunks;
    const uint32_t np = ((d[k].w + 7u) / 8u) * ((d[k].h + 7u) / 8u);
    wchunks += (np + wchunk - 1u) / wchunk;
  }
  uint32_t rtiles = 0;
  if (rform != 0) {
    uint32_t tw = 0, th = 0;
    vrt::refill_tile_dims(rform, &tw, &th);
    for (uint32_t k = 0; k < nviews; ++k) {
      d[k].rtile0 = rtiles;
      d[k].rtx = (d[k].w + tw - 1u) / tw;
      rtiles += d[k].rtx * ((d[k].h + th - 1u) / th);
    }
  }
  std::memset(hh, 0, sizeof *hh);
  hh->nviews = nviews;
  hh->tiles = tiles;
  hh->queue = 0;
  hh->rtiles = rtiles;
  hh->wchunks = wchunks;
  hh->wchunk = wchunk;
  vrt::BatchHead* dh = (vrt::BatchHead*)lane.desc_dev.ptr;
  const vrt::ViewDesc* dd = (const vrt::ViewDesc*)((char*)lane.desc_dev.ptr + vrt::BATCH_HEAD_BYTES);
  HIP_TRY(this, hipMemcpyAsync(lane.desc_dev.ptr, lane.desc_host[stage],
                               vrt::BATCH_HEAD_BYTES + nviews * sizeof(vrt::ViewDesc), hipMemcpyHostToDevice, st));
  HIP_TRY(this, hipEventRecord(lane.desc_done[stage], st));
  HIP_TRY(this, hipEventRecord(ev[0], st));
  if (tuning.batch_mode == 1 || tuning.batch_mode == 2) {
    const int tv = ptv, wv = pt_waves;
    // one-wave blocks for the 16-wide refill tiles: four times as many for the same waves
    const uint32_t ptb = tuning.batch_blocks ? tuning.batch_blocks : (uint32_t)cus * (rform >= 4 ? 32u : 8u);
    const uint32_t gsb = (uint32_t)cus * 8u;
    auto enqueue = [&] {
      if (wave_batch) {
        const hipError_t e = vrt::launch_pathtrace_wave_batch(st, dh, dd, (unsigned long long*)d_counters.ptr,
                                                              (int)wave_blocks, (int)tuning.pt_min_refill,
                                                              (int)tuning.pt_shade_min, wv);
        return e != hipSuccess ? e : vrt::launch_batch_filters(st, dh, dd, gsb, any_denoise, any_hdr);
      }
      return vrt::launch_batch_persistent(st, dh, dd, (unsigned long long*)d_counters.ptr, tv, wv, max_bit_words,
                                          max_pwords, ptb, gsb, any_denoise, any_hdr, rform);
    };
    if (tuning.batch_mode == 1) {
      HIP_TRY(this, enqueue());
    } else {
      // one graph per launch shape; the table it reads is rewritten per batch
      const uint64_t shape = wave_batch ? ((uint64_t)wave_blocks | (uint64_t)std::min(tuning.pt_min_refill, 64u) << 21 |
                                           (uint64_t)std::min(tuning.pt_shade_min, 64u) << 28)
                                        : (uint64_t)ptb;
      const std::pair<uint64_t, uint64_t> key{
          (uint64_t)tv | (uint64_t)wv << 4 | (uint64_t)rform << 8 | (uint64_t)any_denoise << 12 |
              (uint64_t)any_hdr << 13 | (uint64_t)wave_batch << 14 | shape << 16,
          (uint64_t)max_bit_words | (uint64_t)max_pwords << 32};
      auto it = lane.graphs.find(key);
      if (it == lane.graphs.end()) {
        if (lane.graphs.size() >= 16) lane.drop_graphs();
        hipGraph_t g = nullptr;
        HIP_TRY(this, hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
        const hipError_t le = enqueue();
        const hipError_t ce = hipStreamEndCapture(st, &g);
        if (le != hipSuccess || ce != hipSuccess) {
          if (g) (void)hipGraphDestroy(g);
          return fail(VRT_ERR_DEVICE, "graph capture: %s", hipGetErrorString(le != hipSuccess ? le : ce));
        }
        hipGraphExec_t ex = nullptr;
        const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ie != hipSuccess) return fail(VRT_ERR_DEVICE, "graph instantiate: %s", hipGetErrorString(ie));
        it = lane.graphs.emplace(key, ex).first;
      }
      HIP_TRY(this, hipGraphLaunch(it->second, st));
    }
    HIP_TRY(this, hipEventRecord(ev[1], st));
    HIP_TRY(this, hipEventRecord(ev[2], st));
    HIP_TRY(this, hipEventRecord(ev[3], st));
    return VRT_OK;
  }
  if (wave_batch)
    HIP_TRY(this, vrt::launch_pathtrace_wave_batch(st, dh, dd, (unsigned long long*)d_counters.ptr, (int)wave_blocks,
                                                   (int)tuning.pt_min_refill, (int)tuning.pt_shade_min, pt_waves));
  else
    HIP_TRY(this, vrt::launch_pathtrace_batch(st, dd, nviews, tiles, (unsigned long long*)d_counters.ptr,
                                              ptv, pt_waves, max_bit_words, max_pwords, rform, rtiles));
  HIP_TRY(this, hipEventRecord(ev[1], st));
  HIP_TRY(this, vrt::launch_temporal_batch(st, dd, nviews, tiles));
  HIP_TRY(this, hipEventRecord(ev[2], st));
  if (any_denoise)
    for (int pass = 0; pass < 3; ++pass) HIP_TRY(this, vrt::launch_atrous_batch(st, dd, nviews, tiles, pass));
  HIP_TRY(this, hipEventRecord(ev[3], st));
  if (any_hdr) HIP_TRY(this, vrt::launch_hdr_batch(st, dd, nviews, tiles));
  return VRT_OK;
}

int vrt_ctx::harvest(uint32_t max_slots) {
  while (ring_count > 0 && max_slots-- > 0) {
    const TimingSlot& sl = ring[(ring_head - ring_count) % RING];
    HIP_TRY(this, hipEventSynchronize(sl.e[3]));
    float t01 = 0, t12 = 0, t23 = 0, t03 = 0;
    HIP_TRY(this, hipEventElapsedTime(&t01, sl.e[0], sl.e[1]));
    HIP_TRY(this, hipEventElapsedTime(&t12, sl.e[1], sl.e[2]));
    HIP_TRY(this, hipEventElapsedTime(&t23, sl.e[2], sl.e[3]));
    HIP_TRY(this, hipEventElapsedTime(&t03, sl.e[0], sl.e[3]));
    ms_pathtrace = t01;
    ms_temporal = t12;
    ms_atrous = t23;
    ms_frame = t03;
    sum_pathtrace += t01;
    sum_temporal += t12;
    sum_atrous += t23;
    sum_frame += t03;
    ++timed_calls;
    --ring_count;
  }
  return VRT_OK;
}

vrt_ctx::~vrt_ctx() {
  server.reset();  // joins the serving worker; its streams drain first
  for (vrt_scene* s : scenes) delete s;
  scenes.clear();
  views.clear();
  for (auto& kv : user_buffers) (void)hipFree(kv.first);
  user_buffers.clear();
  for (auto& e : ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& slot : ring)
    for (auto& e : slot.e)
      if (e) (void)hipEventDestroy(e);
  // lane0 borrows the context's stream (own_stream is false)
  if (stream) (void)hipStreamDestroy(stream);
}
