// vb_shim.cpp -- TEST INFRASTRUCTURE ONLY.
//
// C entry points around VoxelBench's own CPU traversal, compiled together with
// the reference's VoxelBench/pipeline/core.cpp (read in place, never copied)
// into oracle/_ref/libvoxelbench_core.so by oracle/Makefile.  Used to
// cross-check the oracle's DDA topology (voxel + face) against
// traceNaive / trace (core.cpp:23-130) and as the "reference CPU DDA" timing.
#include <cstring>
#include <omp.h>
#include "core.h"

extern "C" {

// Load an occupancy grid (1 byte per voxel, x fastest: core.h:57 gi()).
// Dims must be multiples of VoxelBench's 4^3 brick (core.h:52-53).
int vb_set_grid(const uint8_t* occ, int nx, int ny, int nz) {
  if (nx % 4 || ny % 4 || nz % 4 || nx <= 0 || ny <= 0 || nz <= 0) return -1;
  setDims(nx, ny, nz);
  std::memcpy(grid.data(), occ, (size_t)nx * ny * nz);
  rebuildAllBricks();
  for (int i = 0; i < MAXMAT; ++i) { MAT[i].alb = V3(0.5f, 0.5f, 0.5f); MAT[i].emi = V3(); }
  return 0;
}

// Trace n rays (o.xyz, d.xyz in voxel units); naive != 0 selects traceNaive.
// out: vx, vy, vz, face per ray (-1s on miss); t: hit distance (voxels).
void vb_trace_rays(const float* rays, long n, float tmax, int naive, int threads,
                   int* out, float* t) {
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 1024)
  for (long i = 0; i < n; ++i) {
    const float* r = rays + i * 6;
    Hit h;
    bool hit = naive ? traceNaive(V3(r[0], r[1], r[2]), V3(r[3], r[4], r[5]), tmax, h)
                     : trace(V3(r[0], r[1], r[2]), V3(r[3], r[4], r[5]), tmax, h);
    out[i * 4 + 0] = hit ? h.vx : -1;
    out[i * 4 + 1] = hit ? h.vy : -1;
    out[i * 4 + 2] = hit ? h.vz : -1;
    out[i * 4 + 3] = hit ? h.face : -1;
    t[i] = hit ? h.t : -1.0f;
  }
}

}  // extern "C"
