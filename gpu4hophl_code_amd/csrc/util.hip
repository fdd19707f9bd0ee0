// util.hip -- device copy benchmark: the achievable HBM bandwidth on this
// GPU (STREAM-style copy, 16 B vector loads and stores) and the reference
// kernel of known traffic that PMC byte counters are calibrated against.
#include <hip/hip_runtime.h>

#include "device.hpp"

namespace hsdl {

__global__ void __launch_bounds__(256) k_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst,
                                              uint64_t n16) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n16;
       i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

}  // namespace hsdl

using namespace hsdl;

extern "C" int hsdl_device_copy_bench(uint64_t bytes, uint32_t iters, double *gbps) {
  if (!gbps || bytes < 16 || iters == 0) return fail(HSDL_EINVAL, "bad argument");
  int dev = 0;
  int rc = use_device(-1, &dev);
  if (rc) return rc;
  const uint64_t n16 = bytes / 16;
  DevBuf a, b;
  if ((rc = a.alloc(n16 * 16)) || (rc = b.alloc(n16 * 16))) return rc;
  HSDL_HIP(hipMemset(a.p, 1, n16 * 16));
  int cus = 256;
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  const dim3 grid((uint32_t)cus * 8), block(256);
  hipLaunchKernelGGL(k_copy, grid, block, 0, nullptr, a.as<uint4>(), b.as<uint4>(), n16);
  HSDL_HIP(hipDeviceSynchronize());
  hipEvent_t e0, e1;
  HSDL_HIP(hipEventCreate(&e0));
  HSDL_HIP(hipEventCreate(&e1));
  HSDL_HIP(hipEventRecord(e0, nullptr));
  for (uint32_t i = 0; i < iters; ++i)
    hipLaunchKernelGGL(k_copy, grid, block, 0, nullptr, a.as<uint4>(), b.as<uint4>(), n16);
  HSDL_HIP(hipEventRecord(e1, nullptr));
  HSDL_HIP(hipEventSynchronize(e1));
  float ms = 0;
  HSDL_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *gbps = 2.0 * (double)n16 * 16 * iters / (ms * 1e-3) / 1e9;  // read + write
  return HSDL_OK;
}
