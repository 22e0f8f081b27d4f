// vb_technique_driver.cpp -- test infrastructure.  Runs MethodHIP
// (include/voxelrt_technique.h) through VoxelBench's Technique interface the
// way the harness's run loop does (VoxelBench/pipeline/harness.cpp:30-37
// prepState, :305-350 run): build a VoxelBench scene (scenes.h), emitters,
// camera, G-buffer; init(); frame() per frame with the scene's event applied
// and onEvent() called at the event frame.  Writes what the GPU test compares
// against the oracle:
//   <out>/grid<k>.bin   u32 px, py, pz, then the packed grid MethodHIP uploaded
//                       (k = 0 before the event, 1 after a geometry event)
//   <out>/frame<f>.bin  the frame's 304-byte uniforms, i64 rays, then out[]
//                       (W*H rgb f32, row-major)
//   <out>/orient.txt    primary-visibility agreement between VoxelBench's
//                       G-buffer (pixDir, core.cpp:180-184) and a G-buffer
//                       rendered through the C ABI with MethodHIP's camera
// Usage: vb_technique_driver <bunker|courtyard|cavern> <frames> <event_frame> <out_dir> [light_frame]
// light_frame: also a light change at that frame -- the sun moves to
// norm(-0.3, 0.9, 0.25) with radiance (14, 9, 4.5) and onEvent() gets an event
// with only lightChange set (as courtyard's own event, scenes.h:75-77,121)
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include "technique.h"
#include "voxelrt_technique.h"

static std::unique_ptr<Scene> make_scene(const std::string& n) {
  if (n == "bunker") return std::make_unique<SceneBunker>();
  if (n == "courtyard") return std::make_unique<SceneCourtyard>();
  if (n == "cavern") return std::make_unique<SceneCavern>();
  return nullptr;
}

static bool write_file(const std::string& path, const std::vector<const void*>& ptrs,
                       const std::vector<size_t>& sizes) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) return false;
  for (size_t i = 0; i < ptrs.size(); ++i) std::fwrite(ptrs[i], 1, sizes[i], f);
  return std::fclose(f) == 0;
}

static bool write_grid(const std::string& path, const MethodHIP& m) {
  const uint32_t dims[3] = {m.px, m.py, m.pz};
  return write_file(path, {dims, m.packed.data()}, {sizeof dims, m.packed.size() * 4});
}

int main(int argc, char** argv) {
  if (argc != 5 && argc != 6) {
    std::fprintf(stderr, "usage: %s <scene> <frames> <event_frame> <out_dir> [light_frame]\n", argv[0]);
    return 2;
  }
  setvbuf(stdout, nullptr, _IOLBF, 0);  // progress lines survive an abnormal end
  auto s = make_scene(argv[1]);
  const int frames = std::atoi(argv[2]), event_frame = std::atoi(argv[3]);
  const std::string out = argv[4];
  const int light_frame = argc == 6 ? std::atoi(argv[5]) : -1;
  if (!s || frames < 1) return 2;
  s->build();
  collectEmitters();
  s->camera();
  buildGBuffer();
  std::printf("scene built\n");
  MethodHIP m;
  m.init();
  std::printf("technique initialised\n");
  if (!m.ctx || !m.scene) return 3;
  if (!write_grid(out + "/grid0.bin", m)) return 4;
  for (int f = 0; f < frames; ++f) {
    if (f == event_frame) {
      s->applyEvent();
      collectEmitters();
      buildGBuffer();
      m.onEvent(s->event);
      if (s->event.geomEdit && !write_grid(out + "/grid1.bin", m)) return 4;
    }
    if (f == light_frame) {
      SKY.enabled = true;
      SKY.sunDir = norm(V3(-0.3f, 0.9f, 0.25f));
      SKY.sunE = V3(14, 9, 4.5f);
      SceneEvent le;
      le.lightChange = true;
      m.onEvent(le);
      std::printf("light change before render %d\n", f);
    }
    const int64_t rays = m.frame(4L * W * H);
    if (rays <= 0) return 5;
    char name[64];
    std::snprintf(name, sizeof name, "/frame%03d.bin", f);
    if (!write_file(out + name, {&m.last_u, &rays, m.out.data()}, {sizeof m.last_u, sizeof rays, m.out.size() * 12}))
      return 4;
    std::printf("frame %d rays %lld\n", f, (long long)rays);
  }
  // primary visibility through the C ABI with the technique's camera (no
  // indirect-only flag, no history) against VoxelBench's own G-buffer
  vrt_view v{};
  v.scene = m.scene;
  v.u = m.last_u;
  v.u.params0[2] = 0u;
  v.firefly_clamp = 48.0f;
  v.view_id = 2;
  v.spp = 1;
  v.accum_format = VRT_ACCUM_RGBA32F;
  std::vector<float> gb((size_t)W * H * 4);
  vrt_frame fr{};
  fr.gbuf = gb.data();
  if (vrt_render(m.ctx, &v, 1, &fr) != VRT_OK) return 6;
  int agree = 0, both = 0;
  std::vector<float> rel;
  for (int i = 0; i < W * H; ++i) {
    const bool ours = gb[i * 4 + 3] >= 0.0f, theirs = gbuf[i].hit;
    agree += ours == theirs;
    if (ours && theirs) {
      ++both;
      const float t = gb[i * 4 + 3] / 0.0625f;  // metres -> voxels
      rel.push_back(std::abs(t - gbuf[i].depth) / std::max(gbuf[i].depth, 1e-3f));
    }
  }
  std::sort(rel.begin(), rel.end());
  FILE* o = std::fopen((out + "/orient.txt").c_str(), "w");
  if (!o) return 4;
  std::fprintf(o, "pixels %d agree %d both_hit %d median_rel_depth %.6f\n", W * H, agree, both,
               rel.empty() ? -1.0 : (double)rel[rel.size() / 2]);
  std::fclose(o);
  std::printf("done\n");
  return 0;
}
