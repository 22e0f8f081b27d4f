// Exhaustive check of the device safe_inv (voxelrt_amd/csrc/vrt_trace.h)
// against the correctly rounded quotient s / max(|v|, 1e-8) by IEEE division
// (voxel.wgsl:65-71), over every one of the 2^32 float bit patterns v.
// Built by voxelrt_amd/csrc/Makefile into tests/_build/rcp_exhaustive; run by
// tests/test_gpu_rcp.py.  Prints the mismatch count and up to 16 of them;
// exit status 0 only when there are none.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>

#include "../voxelrt_amd/csrc/vrt_trace.h"

constexpr uint32_t NFIRST = 16;

__global__ void k_check(unsigned long long* nbad, uint32_t* first) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (1ull << 32); i += stride) {
    const float v = __builtin_bit_cast(float, (uint32_t)i);
    const float s = v >= 0.0f ? 1.0f : -1.0f;
    const float want = s / vrt::vmax(vrt::vabs(v), 1e-8f);
    const float got = vrt::safe_inv(v);
    if (__builtin_bit_cast(uint32_t, want) != __builtin_bit_cast(uint32_t, got)) {
      const unsigned long long k = atomicAdd(nbad, 1ull);
      if (k < NFIRST) first[k] = (uint32_t)i;
    }
  }
}

#define CHECK(x)                                                               \
  do {                                                                         \
    const hipError_t e_ = (x);                                                 \
    if (e_ != hipSuccess) {                                                    \
      std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));             \
      return 2;                                                                \
    }                                                                          \
  } while (0)

int main() {
  unsigned long long* d_n = nullptr;
  uint32_t* d_first = nullptr;
  CHECK(hipMalloc(&d_n, sizeof(unsigned long long)));
  CHECK(hipMalloc(&d_first, NFIRST * sizeof(uint32_t)));
  CHECK(hipMemset(d_n, 0, sizeof(unsigned long long)));
  CHECK(hipMemset(d_first, 0, NFIRST * sizeof(uint32_t)));
  hipLaunchKernelGGL(k_check, dim3(16384), dim3(256), 0, 0, d_n, d_first);
  CHECK(hipGetLastError());
  CHECK(hipDeviceSynchronize());
  unsigned long long n = 0;
  uint32_t first[NFIRST];
  CHECK(hipMemcpy(&n, d_n, sizeof(n), hipMemcpyDeviceToHost));
  CHECK(hipMemcpy(first, d_first, sizeof(first), hipMemcpyDeviceToHost));
  std::printf("safe_inv vs IEEE division over 2^32 floats: mismatches %llu\n", n);
  for (uint32_t k = 0; k < NFIRST && k < n; ++k) {
    float v;
    std::memcpy(&v, &first[k], 4);
    std::printf("  v = %a (0x%08x)\n", v, first[k]);
  }
  CHECK(hipFree(d_n));
  CHECK(hipFree(d_first));
  return n == 0 ? 0 : 1;
}
