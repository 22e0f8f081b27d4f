// vrt_cpu_stub.cpp -- TEST INFRASTRUCTURE ONLY.
//
// The C ABI calls VoxelBench's Technique adapter makes (include/
// voxelrt_technique.h: vrt_create, vrt_destroy, vrt_last_error,
// vrt_scene_upload_ex, vrt_scene_edit, vrt_scene_release, vrt_view_release,
// vrt_render), answered on the CPU by the oracle's frame chain (pathtrace ->
// temporal -> a-trous 1/2/4 -> hdr, the order of main.js:680-747).  Linked
// with VoxelBench's core, the adapter and the driver under AddressSanitizer
// and UBSan (oracle/Makefile, target asan), so that the adapter and the
// reference's scene code run their init -> frame -> event -> frame -> exit
// sequence with every host access checked, without a GPU.  vrt_uniforms_init
// and vrt_camera_matrices come from the library's host code
// (voxelrt_amd/csrc/vrt_host.cpp), compiled into the same binary.
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../include/voxelrt.h"
#include "vrt_oracle.h"

struct vrt_scene {
  uint32_t nx = 0, ny = 0, nz = 0;
  std::vector<uint32_t> vox, bricks, masks;
  orc_grid g{};
  bool build() {
    const size_t nb = (size_t)(nx / 8) * (ny / 8) * (nz / 8);
    bricks.assign(nb, 0u);
    masks.assign(nb * 16, 0u);
    if (orc_build_bricks(vox.data(), (int)nx, (int)ny, (int)nz, bricks.data(), masks.data()) != 0) return false;
    g = orc_grid{vox.data(), bricks.data(), masks.data(), (int32_t)nx, (int32_t)ny, (int32_t)nz, VRT_VOXEL_SIZE};
    return true;
  }
};

namespace {
struct ViewState {
  uint32_t w = 0, h = 0, fmt = 0, frames = 0;
  std::vector<float> gbuf[2];
  std::vector<uint8_t> accum[2];
};
}  // namespace

struct vrt_ctx {
  std::string err;
  std::map<uint32_t, ViewState> views;
  std::vector<vrt_scene*> scenes;
  int fail(int code, const char* msg) {
    err = msg;
    return code;
  }
};

extern "C" {

int vrt_create(const vrt_create_info*, vrt_ctx** out) {
  if (!out) return VRT_ERR_INVALID;
  *out = new vrt_ctx();
  return VRT_OK;
}

void vrt_destroy(vrt_ctx* ctx) {
  if (!ctx) return;
  for (vrt_scene* s : ctx->scenes) delete s;
  delete ctx;
}

const char* vrt_last_error(const vrt_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int vrt_scene_upload_ex(vrt_ctx* ctx, const uint32_t* vox, uint32_t nx, uint32_t ny, uint32_t nz, uint64_t,
                        uint32_t, vrt_scene** out) {
  if (!ctx || !vox || !out) return VRT_ERR_INVALID;
  if (!nx || !ny || !nz || nx % 8 || ny % 8 || nz % 8) return ctx->fail(VRT_ERR_INVALID, "dims");
  auto* s = new vrt_scene();
  s->nx = nx;
  s->ny = ny;
  s->nz = nz;
  s->vox.assign(vox, vox + (size_t)nx * ny * nz);
  if (!s->build()) {
    delete s;
    return ctx->fail(VRT_ERR_INVALID, "brick build");
  }
  ctx->scenes.push_back(s);
  *out = s;
  return VRT_OK;
}

int vrt_scene_edit(vrt_ctx* ctx, vrt_scene* s, const uint32_t* vox, uint32_t x0, uint32_t y0, uint32_t z0,
                   uint32_t sx, uint32_t sy, uint32_t sz) {
  if (!ctx || !s || !vox) return VRT_ERR_INVALID;
  if ((uint64_t)x0 + sx > s->nx || (uint64_t)y0 + sy > s->ny || (uint64_t)z0 + sz > s->nz)
    return ctx->fail(VRT_ERR_INVALID, "edit box outside the grid");
  for (uint32_t z = 0; z < sz; ++z)
    for (uint32_t y = 0; y < sy; ++y)
      for (uint32_t x = 0; x < sx; ++x)
        s->vox[(size_t)(x0 + x) + (size_t)(y0 + y) * s->nx + (size_t)(z0 + z) * s->nx * s->ny] =
            vox[(size_t)x + (size_t)y * sx + (size_t)z * sx * sy];
  return s->build() ? VRT_OK : ctx->fail(VRT_ERR_INVALID, "brick build");
}

int vrt_scene_release(vrt_ctx* ctx, vrt_scene* s) {
  if (!ctx || !s) return VRT_ERR_INVALID;
  for (size_t i = 0; i < ctx->scenes.size(); ++i)
    if (ctx->scenes[i] == s) {
      ctx->scenes.erase(ctx->scenes.begin() + (long)i);
      delete s;
      return VRT_OK;
    }
  return ctx->fail(VRT_ERR_INVALID, "unknown scene");
}

int vrt_view_release(vrt_ctx* ctx, uint32_t view_id) {
  if (!ctx) return VRT_ERR_INVALID;
  ctx->views.erase(view_id);
  return VRT_OK;
}

int vrt_render(vrt_ctx* ctx, const vrt_view* views, uint32_t nviews, vrt_frame* frames) {
  if (!ctx || !views || !nviews) return VRT_ERR_INVALID;
  for (uint32_t k = 0; k < nviews; ++k) {
    const vrt_view& v = views[k];
    if (!v.scene) return ctx->fail(VRT_ERR_INVALID, "view without scene");
    const uint32_t w = (uint32_t)v.u.params1[0], h = (uint32_t)v.u.params1[1];
    if (!w || !h) return ctx->fail(VRT_ERR_INVALID, "view size");
    ViewState& st = ctx->views[v.view_id];
    const size_t npx = (size_t)w * h;
    const uint32_t abytes = v.accum_format == VRT_ACCUM_RGBA32F ? 16u : 8u;
    if (st.w != w || st.h != h || st.fmt != v.accum_format) {
      st = ViewState();
      st.w = w;
      st.h = h;
      st.fmt = v.accum_format;
      for (int p = 0; p < 2; ++p) {
        st.gbuf[p].assign(npx * 4, 0.0f);
        st.accum[p].assign(npx * abytes, 0u);
      }
    }
    const int p = (int)(st.frames & 1u);
    std::vector<uint16_t> rad(npx * 4);
    std::vector<uint8_t> alb(npx * 4);
    orc_counters cnt{};
    orc_pathtrace(&v.scene->g, &v.u, v.firefly_clamp, v.spp ? v.spp : 1u, 0, (int)h, rad.data(), alb.data(),
                  st.gbuf[p].data(), &cnt, 0);
    orc_temporal(&v.u, rad.data(), st.gbuf[p].data(), st.accum[1 - p].data(), st.gbuf[1 - p].data(), st.fmt,
                 st.accum[p].data(), 0);
    const void* illum = st.accum[p].data();
    uint32_t ifmt = st.fmt;
    std::vector<uint16_t> d0, d1;
    if (v.u.params0[2] & VRT_FLAG_DENOISE) {
      d0.assign(npx * 4, 0u);
      d1.assign(npx * 4, 0u);
      orc_atrous(&v.u, st.accum[p].data(), st.fmt, st.gbuf[p].data(), 1, d0.data(), 0);
      orc_atrous(&v.u, d0.data(), VRT_ACCUM_RGBA16F, st.gbuf[p].data(), 2, d1.data(), 0);
      orc_atrous(&v.u, d1.data(), VRT_ACCUM_RGBA16F, st.gbuf[p].data(), 4, d0.data(), 0);
      illum = d0.data();
      ifmt = VRT_ACCUM_RGBA16F;
    }
    if (frames) {
      vrt_frame& f = frames[k];
      if (f.radiance) std::memcpy(f.radiance, rad.data(), npx * 8);
      if (f.albedo) std::memcpy(f.albedo, alb.data(), npx * 4);
      if (f.gbuf) std::memcpy(f.gbuf, st.gbuf[p].data(), npx * 16);
      if (f.accum) std::memcpy(f.accum, st.accum[p].data(), npx * abytes);
      if (f.denoised && !d0.empty()) std::memcpy(f.denoised, d0.data(), npx * 8);
      if (f.hdr) orc_hdr((int)w, (int)h, illum, ifmt, alb.data(), f.hdr);
      if (f.ldr) orc_present((int)w, (int)h, illum, ifmt, alb.data(), v.u.params1[2], f.ldr);
      f.rays = cnt.rays;
      f.brick_steps = cnt.brick_steps;
      f.voxel_steps = cnt.voxel_steps;
    }
    ++st.frames;
  }
  return VRT_OK;
}

}  // extern "C"
