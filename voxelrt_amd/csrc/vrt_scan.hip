// vrt_scan.hip -- device exclusive prefix sum over the brick occupancy flags,
// giving each occupied brick its slot in the compacted brick pool.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <stdint.h>

namespace vrt {

// Query (d_tmp == nullptr) or run the scan.  tmp_bytes in/out as in hipCUB.
hipError_t exclusive_scan_u32(hipStream_t st, void* d_tmp, size_t* tmp_bytes,
                              const uint32_t* in, uint32_t* out, uint64_t n) {
  return hipcub::DeviceScan::ExclusiveSum(d_tmp, *tmp_bytes, in, out, (int)n, st);
}

}  // namespace vrt
