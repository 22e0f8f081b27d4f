// voxelrt_technique.h -- libvoxelrt_hip as a VoxelBench GI technique.
//
// Drop-in for VoxelBench's plugin interface (VoxelBench/pipeline/
// technique.h:8-15).  Include it in harness.cpp after technique.h and
// register it in makeTech (harness.cpp:39-199):
//     if(m=="HIP") return new MethodHIP();
// then link -lvoxelrt_hip.  The harness owns the world (core.h globals
// `grid`, `MAT`, `SKY`, camera) and calls init() once, onEvent() at the event
// frame and frame(budget) per frame (harness.cpp:305-350).
//
// Per frame, MethodHIP renders the harness's W x H view on the GPU through
// the C ABI -- pathtrace (2 bounces, indirect only: no primary emission, no
// primary sun NEE, no primary sky, since the harness adds its own direct
// pass), temporal accumulation and the 3-pass a-trous filter -- and writes
// `out` = the filtered indirect radiance times albedo (linear, pre-tonemap),
// the quantity VoxelBench techniques estimate (technique.h:2-3).
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <vector>

#include "technique.h"
#include "voxelrt.h"

struct MethodHIP : Technique {
  vrt_ctx* ctx = nullptr;
  vrt_scene* scene = nullptr;
  uint32_t px = 0, py = 0, pz = 0;  // padded dims (multiples of 8)
  uint32_t frameIndex = 0;
  uint32_t bounces = 2;
  float prevVP[16];
  double prevPos[3];
  bool havePrev = false;
  std::vector<float> hdr;
  std::vector<uint32_t> packed;  // the grid as uploaded (padded dims px, py, pz)
  vrt_uniforms last_u{};         // uniforms of the last frame() (for inspection / replay)

  const char* name() override { return "HIP"; }

  // VoxelBench palette entry -> packed word (scene.js:18 layout): albedo
  // byte = round(255 sqrt(alb)) (the device squares byte/255); the emissive
  // byte makes albedo * e/255 * 24 match the emission's luminance.
  static uint32_t packMaterial(uint8_t id) {
    if (id == 0) return 0u;
    const Material& m = MAT[id];
    auto byte = [](float a) {
      return (uint32_t)std::lround(255.0 * std::sqrt(std::max(0.0, std::min(1.0, (double)a))));
    };
    uint32_t r = byte(m.alb.x), g = byte(m.alb.y), b = byte(m.alb.z);
    const double la = 0.2126 * m.alb.x + 0.7152 * m.alb.y + 0.0722 * m.alb.z;
    const double le = 0.2126 * m.emi.x + 0.7152 * m.emi.y + 0.0722 * m.emi.z;
    uint32_t e = 0;
    if (le > 0.0) {
      if (la < 1e-3) r = g = b = 255;  // pure emitter: carry the emission on white
      e = (uint32_t)std::min(255.0, std::max(1.0, std::round(255.0 * le / (24.0 * std::max(la, 1e-3)))));
    }
    uint32_t w = r | (g << 8) | (b << 16) | (e << 24);
    return w ? w : 1u;  // solid voxels must stay non-zero
  }

  void pack(std::vector<uint32_t>& dst) const {
    dst.assign((size_t)px * py * pz, 0u);
    for (int z = 0; z < NZ; ++z)
      for (int y = 0; y < NY; ++y)
        for (int x = 0; x < NX; ++x)
          dst[(size_t)x + (size_t)y * px + (size_t)z * px * py] = packMaterial(grid[gi(x, y, z)]);
  }

  void upload() {
    if (scene) vrt_scene_release(ctx, scene);
    scene = nullptr;
    px = (uint32_t)((NX + 7) / 8 * 8);
    py = (uint32_t)((NY + 7) / 8 * 8);
    pz = (uint32_t)((NZ + 7) / 8 * 8);
    pack(packed);
    // editable: geometry events rebuild on the device from a box edit
    if (vrt_scene_upload_ex(ctx, packed.data(), px, py, pz, 0, VRT_SCENE_EDITABLE, &scene) != VRT_OK)
      std::fprintf(stderr, "MethodHIP: scene upload failed: %s\n", vrt_last_error(ctx));
  }

  // A geometry event (scenes.h:8-12): send the bounding box of the changed
  // voxels to the device, which rebuilds bricks, masks and pool there.
  void edit() {
    std::vector<uint32_t> next;
    pack(next);
    uint32_t lo[3] = {px, py, pz}, hi[3] = {0, 0, 0};
    bool any = false;
    for (uint32_t z = 0; z < pz; ++z)
      for (uint32_t y = 0; y < py; ++y)
        for (uint32_t x = 0; x < px; ++x) {
          const size_t i = (size_t)x + (size_t)y * px + (size_t)z * px * py;
          if (next[i] == packed[i]) continue;
          const uint32_t c[3] = {x, y, z};
          for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], c[k]);
            hi[k] = std::max(hi[k], c[k]);
          }
          any = true;
        }
    packed.swap(next);
    if (!any) return;
    const uint32_t sx = hi[0] - lo[0] + 1, sy = hi[1] - lo[1] + 1, sz = hi[2] - lo[2] + 1;
    std::vector<uint32_t> box((size_t)sx * sy * sz);
    for (uint32_t z = 0; z < sz; ++z)
      for (uint32_t y = 0; y < sy; ++y)
        for (uint32_t x = 0; x < sx; ++x)
          box[(size_t)x + (size_t)y * sx + (size_t)z * sx * sy] =
              packed[(size_t)(lo[0] + x) + (size_t)(lo[1] + y) * px + (size_t)(lo[2] + z) * px * py];
    if (vrt_scene_edit(ctx, scene, box.data(), lo[0], lo[1], lo[2], sx, sy, sz) != VRT_OK)
      std::fprintf(stderr, "MethodHIP: scene edit failed: %s\n", vrt_last_error(ctx));
  }

  void init() override {
    if (!ctx) {
      vrt_create_info ci{0, 0, 0};
      if (vrt_create(&ci, &ctx) != VRT_OK) {
        std::fprintf(stderr, "MethodHIP: no HIP device\n");
        ctx = nullptr;
      }
    }
    out.assign(W * H, V3());
    hdr.assign((size_t)W * H * 3, 0.0f);
    frameIndex = 0;
    havePrev = false;
    if (ctx) {
      vrt_view_release(ctx, 1);
      upload();
    }
  }

  void onEvent(const SceneEvent& e) override {
    if (e.geomEdit && ctx && scene) edit();  // lighting-only events: SKY is read every frame
  }

  long frame(long rayBudget) override {
    (void)rayBudget;  // one 1-spp frame per call; the GPU reports what it traced
    if (!ctx || !scene) return 0;
    // VoxelBench camera (voxel units, core.h:141) -> metres; yaw/pitch of camFwd
    const double pos[3] = {camPos.x * 0.0625, camPos.y * 0.0625, camPos.z * 0.0625};
    const double yaw = std::atan2((double)camFwd.x, (double)camFwd.z);
    const double pitch = std::asin(std::max(-1.0, std::min(1.0, (double)camFwd.y)));
    const double sunAz = std::atan2((double)SKY.sunDir.z, (double)SKY.sunDir.x);
    const double sunEl = std::asin(std::max(-1.0, std::min(1.0, (double)SKY.sunDir.y)));
    vrt_view v{};
    v.scene = scene;
    vrt_uniforms_init(&v.u, pos, yaw, pitch, havePrev ? prevVP : nullptr, havePrev ? prevPos : nullptr,
                      (uint32_t)W, (uint32_t)H, frameIndex, bounces,
                      VRT_FLAG_TEMPORAL | VRT_FLAG_DENOISE | VRT_FLAG_INDIRECT, sunAz, sunEl,
                      std::cos(0.5 * M_PI / 180.0), 64.0f);
    v.u.sun_radiance[0] = SKY.sunE.x;
    v.u.sun_radiance[1] = SKY.sunE.y;
    v.u.sun_radiance[2] = SKY.sunE.z;
    v.u.sun_radiance[3] = SKY.enabled ? 1.0f : 0.0f;
    // fov: VoxelBench's camFovDeg is vertical (core.cpp:180-184)
    double fovRad = camFovDeg * M_PI / 180.0;
    float vp[16], ivp[16];
    vrt_camera_matrices(pos, yaw, pitch, fovRad, (double)W / H, 0.05, 400.0, vp, ivp);
    std::copy(ivp, ivp + 16, v.u.inv_view_proj);
    if (!havePrev) std::copy(vp, vp + 16, v.u.prev_view_proj);
    v.firefly_clamp = 48.0f;
    v.view_id = 1;
    v.spp = 1;
    v.accum_format = VRT_ACCUM_RGBA32F;
    last_u = v.u;
    vrt_frame f{};
    f.hdr = hdr.data();
    if (vrt_render(ctx, &v, 1, &f) != VRT_OK) {
      std::fprintf(stderr, "MethodHIP: render failed: %s\n", vrt_last_error(ctx));
      return 0;
    }
    for (int i = 0; i < W * H; ++i) out[i] = V3(hdr[i * 3 + 0], hdr[i * 3 + 1], hdr[i * 3 + 2]);
    std::copy(vp, vp + 16, prevVP);
    std::copy(pos, pos + 3, prevPos);
    havePrev = true;
    ++frameIndex;
    return (long)f.rays;
  }

  ~MethodHIP() override {
    if (ctx) {
      if (scene) vrt_scene_release(ctx, scene);
      vrt_destroy(ctx);
    }
  }
};
