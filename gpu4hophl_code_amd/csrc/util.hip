// util.hip -- device microbenchmarks for the rooflines:
//  * copy: the achievable HBM bandwidth on this GPU (STREAM-style copy, 16 B
//    vector loads and stores), also the reference kernel of known traffic
//    that PMC byte counters are calibrated against;
//  * random atomicMin: the rate of returning 32-bit atomicMin on uniformly
//    random words of a large array -- the operation the generation's dense
//    dedupe slots are built on.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device.hpp"

namespace hsdl {

__global__ void __launch_bounds__(256) k_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst,
                                              uint64_t n16) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n16;
       i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// splitmix64 finaliser: a well-mixed index stream without state
__device__ inline uint64_t mix64(uint64_t x) {
  x ^= x >> 30;
  x *= 0xbf58476d1ce4e5b9ull;
  x ^= x >> 27;
  x *= 0x94d049bb133111ebull;
  x ^= x >> 31;
  return x;
}

__global__ void __launch_bounds__(256) k_atomic_min(uint32_t *a, uint64_t words, uint64_t ops,
                                                    uint64_t salt, uint32_t *sink) {
  uint32_t acc = 0;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < ops;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t r = mix64(i ^ salt);
    const uint64_t idx = (uint64_t)(((unsigned __int128)r * words) >> 64);  // < words
    acc += atomicMin(&a[idx], (uint32_t)(r >> 40));
  }
  if (acc == 0x9E3779B9u) sink[0] = acc;  // keeps the returned values live
}

// Random 4 B gathers from a `words`-word table: `reads` loads at uniformly
// random word indices, summed so none is dropped.  Known algorithmic bytes
// (4 B per read) for calibrating the PMC byte counters on random access,
// where each load moves a whole memory sector.
__global__ void __launch_bounds__(256) k_gather_bench(const uint32_t *__restrict__ a, uint64_t words,
                                                      uint64_t reads, uint64_t salt,
                                                      uint32_t *sink) {
  uint32_t acc = 0;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < reads;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t r = mix64(i ^ salt);
    acc += a[(uint64_t)(((unsigned __int128)r * words) >> 64)];  // index < words
  }
  if (acc == 0x9E3779B9u) sink[0] = acc;  // keeps the loads live
}

}  // namespace hsdl

using namespace hsdl;

extern "C" int hsdl_device_gather_bench(uint64_t bytes, uint64_t reads, double *gbps) {
  if (!gbps || bytes < 4 || reads == 0) return fail(HSDL_EINVAL, "bad argument");
  int dev = 0;
  int rc = use_device(-1, &dev);
  if (rc) return rc;
  const uint64_t words = bytes / 4;
  DevBuf a, sink;
  if ((rc = a.alloc(words * 4)) || (rc = sink.alloc(4))) return rc;
  HSDL_HIP(hipMemset(a.p, 1, words * 4));
  int cus = 256;
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  const dim3 grid((uint32_t)cus * 8), block(256);
  hipEvent_t e0, e1;
  HSDL_HIP(hipEventCreate(&e0));
  HSDL_HIP(hipEventCreate(&e1));
  HSDL_HIP(hipEventRecord(e0, nullptr));
  hipLaunchKernelGGL(k_gather_bench, grid, block, 0, nullptr, a.as<uint32_t>(), words, reads, 3ull,
                     sink.as<uint32_t>());
  HSDL_HIP(hipEventRecord(e1, nullptr));
  HSDL_HIP(hipEventSynchronize(e1));
  float ms = 0;
  HSDL_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *gbps = 4.0 * (double)reads / (ms * 1e-3) / 1e9;  // algorithmic: 4 B per read
  return HSDL_OK;
}

extern "C" int hsdl_device_atomic_bench(uint64_t bytes, uint64_t ops, double *gops) {
  if (!gops || bytes < 4 || ops == 0) return fail(HSDL_EINVAL, "bad argument");
  int dev = 0;
  int rc = use_device(-1, &dev);
  if (rc) return rc;
  const uint64_t words = bytes / 4;
  DevBuf a, sink;
  if ((rc = a.alloc(words * 4)) || (rc = sink.alloc(4))) return rc;
  HSDL_HIP(hipMemset(a.p, 0xFF, words * 4));
  int cus = 256;
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  const dim3 grid((uint32_t)cus * 8), block(256);
  hipLaunchKernelGGL(k_atomic_min, grid, block, 0, nullptr, a.as<uint32_t>(), words,
                     std::min<uint64_t>(ops, 1u << 24), 1ull, sink.as<uint32_t>());
  HSDL_HIP(hipDeviceSynchronize());
  hipEvent_t e0, e1;
  HSDL_HIP(hipEventCreate(&e0));
  HSDL_HIP(hipEventCreate(&e1));
  HSDL_HIP(hipEventRecord(e0, nullptr));
  hipLaunchKernelGGL(k_atomic_min, grid, block, 0, nullptr, a.as<uint32_t>(), words, ops, 2ull,
                     sink.as<uint32_t>());
  HSDL_HIP(hipEventRecord(e1, nullptr));
  HSDL_HIP(hipEventSynchronize(e1));
  float ms = 0;
  HSDL_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *gops = (double)ops / (ms * 1e-3) / 1e9;
  return HSDL_OK;
}

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
