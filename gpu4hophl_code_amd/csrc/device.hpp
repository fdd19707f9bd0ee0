// device.hpp -- device-side types and helpers shared by the HSDL kernels.
//
// Label encoding in HBM (8 B, one u64 per label):
//     bits 63..32  hub (u32 vertex id)
//     bits 31..27  hop (0..31)
//     bits 26..0   dist (0..2^27-1)
// Sorting the u64 keys orders a vertex's labels by (hub, hop, dist), the
// reference's per-vertex order (CPU_HopHL/include/HBPLL/
// hop_constrained_two_hop_labels_generation.h:373-387).  The logical label of
// the C ABI is (hub u32, hop u8, dist i32) = 9 B.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <string>

#include "common.hpp"

namespace hsdl {

constexpr uint32_t kDistBits = 27;
constexpr uint32_t kMaxDist = (1u << kDistBits) - 1;
constexpr uint32_t kMaxK = 15;
constexpr uint32_t kInf32 = 0xFFFFFFFFu;
constexpr int kWave = 64;

__host__ __device__ inline uint64_t pack_label(uint32_t hub, uint32_t hop, uint32_t dist) {
  return ((uint64_t)hub << 32) | ((uint64_t)hop << kDistBits) | (uint64_t)dist;
}
__host__ __device__ inline uint32_t label_hub(uint64_t l) { return (uint32_t)(l >> 32); }
__host__ __device__ inline uint32_t label_hop(uint64_t l) { return ((uint32_t)l) >> kDistBits; }
__host__ __device__ inline uint32_t label_dist(uint64_t l) { return ((uint32_t)l) & kMaxDist; }

// HIP error to status code with detail.
int hip_fail(hipError_t e, const char *what, const char *file, int line);

#define HSDL_HIP(call)                                                  \
  do {                                                                  \
    hipError_t _e = (call);                                             \
    if (_e != hipSuccess) return ::hsdl::hip_fail(_e, #call, __FILE__, __LINE__); \
  } while (0)

// Selects `device` (or keeps the current one for -1); fails with HSDL_ENODEV
// when no GPU is visible.
int use_device(int32_t device, int *out_device);

// Integer tuning knob from the environment, clamped to [lo, hi].
inline uint32_t env_u32(const char *name, uint32_t dflt, uint32_t lo, uint32_t hi) {
  const char *v = getenv(name);
  if (!v || !*v) return dflt;
  const unsigned long x = strtoul(v, nullptr, 10);
  return (uint32_t)(x < lo ? lo : (x > hi ? hi : x));
}

// Hash tile of a label list (the prune keeps L(s) in one, the clean L(v)): an
// open-addressing hash with keys[c] = hub and vals[c*(K+1)+x] = min dist over
// the list's labels (hub, hb, db) with hb <= x.
// V is u16 when every label distance fits (K * max weight < 65535), halving
// the tile; the empty value is V's maximum.  Capacities are powers of two.
template <class V> struct TileVal;
template <> struct TileVal<uint16_t> { static constexpr uint32_t kEmpty = 0xFFFFu; };
template <> struct TileVal<uint32_t> { static constexpr uint32_t kEmpty = 0xFFFFFFFFu; };

__host__ __device__ inline uint32_t pow2_at_least(uint64_t x) {
  uint32_t p = 2;
  while (p < x) p <<= 1;
  return p;
}
__host__ __device__ inline uint32_t pow2_at_most(uint32_t x) {
  uint32_t p = 2;
  while ((uint64_t)p * 2 <= x) p <<= 1;
  return p;
}

// Fibonacci hashing: the top log2(cap) bits of key * G mod 2^32, with G =
// 2^24 / golden ratio (odd), so consecutive hub ids land about 8 / 1.618
// slots apart per 2^11 slots and never cluster.  The product is a 24-bit
// multiply of the key's low 24 bits (full rate, where a 32-bit multiply runs
// at a quarter); keys that differ only above bit 23 share a home slot, which
// costs probes, never correctness.
// (HIP's __umul24 returns int: the product is taken as unsigned before the
// shift, so the shift is logical and the slot stays below cap.)
constexpr uint32_t kGold24 = 0x9E3779u;
__device__ inline uint32_t mul_u24(uint32_t x, uint32_t y) { return (uint32_t)__umul24(x, y); }
__device__ inline uint32_t hash_slot(uint32_t key, uint32_t cap) {
  return mul_u24(key, kGold24) >> (__builtin_clz(cap) + 1);
}

// Tile capacity for `len` distinct hubs: a power of two, load factor <= 2/3.
__host__ __device__ inline uint32_t tile_cap_for(uint32_t len) {
  return pow2_at_least((uint64_t)len + (len >> 1) + 1);
}

// Capacity of a tile that fits in `lds_slots` LDS slots: the load factor is
// lowered to 1/8 where the slots allow (a lookup of an absent hub then ends
// after about 1.1 probes on average instead of 5 at 2/3, and a wave waits for
// its longest lane).
__host__ __device__ inline uint32_t lds_tile_cap(uint32_t len, uint32_t lds_slots) {
  const uint32_t roomy = pow2_at_least(8ull * len + 1);
  const uint32_t most = pow2_at_most(lds_slots);
  return roomy < most ? roomy : most;
}

// u32 words a tile of `cap` slots takes (keys + values, value rows padded)
template <class V>
__device__ __host__ inline uint64_t tile_words(uint32_t cap, uint32_t K1) {
  return (uint64_t)cap + ((uint64_t)cap * K1 * sizeof(V) + 3) / 4;
}

// Pointers in a named address space.  A tile read through a generic pointer
// compiles to flat loads, and the wait for a flat load also waits for every
// global load in flight (the list loads a wave issues ahead); tiles are read
// through one of these instead, so LDS tiles use ds_read and arena tiles
// global loads.
template <class T> using lds_ptr = __attribute__((address_space(3))) T *;
template <class T> using glb_ptr = __attribute__((address_space(1))) T *;

// Device buffer owned by RAII.
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  int alloc(size_t n) {
    release();
    if (n == 0) n = 16;
    hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess) {
      p = nullptr;
      (void)hipGetLastError();
      return fail(HSDL_ENOMEM, "hipMalloc of " + std::to_string(n) + " bytes failed");
    }
    bytes = n;
    return HSDL_OK;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <class T> T *as() const { return static_cast<T *>(p); }
};

// Wave-level helpers (wave64).
__device__ inline int lane_id() { return (int)(threadIdx.x & (kWave - 1)); }

__device__ inline uint32_t wave_min_u32(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    uint32_t t = __shfl_xor(v, o, kWave);
    v = t < v ? t : v;
  }
  return v;
}

__device__ inline uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    uint32_t t = __shfl_xor(v, o, kWave);
    v = t > v ? t : v;
  }
  return v;
}

__device__ inline uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

__device__ inline unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

// Inclusive prefix sum across the wave.
__device__ inline uint32_t wave_incl_scan_u32(uint32_t v) {
  const int l = lane_id();
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    uint32_t t = __shfl_up(v, o, kWave);
    if (l >= o) v += t;
  }
  return v;
}

__device__ inline unsigned long long wave_incl_scan_u64(unsigned long long v) {
  const int l = lane_id();
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    unsigned long long t = __shfl_up(v, o, kWave);
    if (l >= o) v += t;
  }
  return v;
}

}  // namespace hsdl

// ---------------------------------------------------------------- index
// A label index resident in HBM: per-vertex CSR of packed labels, sorted by
// (hub, hop, dist) within each vertex.
struct hsdl_index {
  int device = 0;
  uint32_t n = 0;
  uint8_t K = 0;
  uint64_t total = 0;
  hsdl::DevBuf off;  // u64[n+1]
  hsdl::DevBuf lab;  // u64[total]
  hsdl::DevBuf par;  // u32[total]: parent vertex of each label (has_parents)
  bool has_parents = false;
  bool cleaned = false;
  hsdl_stats stats{};
  uint32_t dist_max = hsdl::kMaxDist;  // bound on every label's dist
  // Query copy of `lab`: hub << (hop_bits + dist_bits) | hop << dist_bits |
  // dist in 32 bits (same per-vertex order), made by the first query call that
  // may allocate when the fields fit; released whenever `lab` changes.
  // q32: 0 not made yet, 1 made, 2 does not fit or switched off.
  mutable hsdl::DevBuf lab32;
  mutable std::atomic<uint32_t> q32{0};  // set with release after the fields below
  mutable uint32_t q32_hop_bits = 0, q32_dist_bits = 0;
};

namespace hsdl {

// Internal entry points shared between translation units.
// Sorts each vertex segment of `lab` (u64 keys) given u64 offsets; in place
// through a temporary buffer.
int segmented_sort_labels(uint64_t *d_lab, uint64_t total, const uint64_t *d_off, uint32_t n,
                          hipStream_t st);
// Sorts each vertex segment of 32-bit keys (bits [0, end_bit)) from `in` into
// `out` with the caller's temporary storage; tmp == nullptr: only sets
// tmp_bytes.
int segmented_sort_u32(void *tmp, size_t &tmp_bytes, const uint32_t *in, uint32_t *out,
                       uint64_t total, const uint64_t *d_off, uint32_t n, uint32_t end_bit,
                       hipStream_t st);
// Exclusive scan of u32 counts into u64 offsets (n+1 entries, last = total).
int scan_counts_u32(const uint32_t *d_cnt, uint64_t *d_off, uint32_t n, hipStream_t st);
int scan_counts_u64(const uint64_t *d_cnt, uint64_t *d_off, uint32_t n, hipStream_t st);

// Parent vertices of an uncleaned K-round index (parents.hip).
int derive_parents(hsdl_index *idx, const hsdl_graph *g, hipStream_t st);

// Query launch used by hsdl_query_device and the server.
// may_alloc: the call may make the index's 32-bit query copy (device memory
// of 4 B per label, kept with the index); without it the query reads the
// copy only when an earlier call made it.
int launch_query(const hsdl_index *idx, size_t nq, const uint32_t *d_s, const uint32_t *d_t,
                 const int8_t *d_k, int32_t *d_out, hipStream_t st, bool may_alloc = true);
// Makes the index's 32-bit query copy now, when it fits (synchronous).
int prepare_query(const hsdl_index *idx);

}  // namespace hsdl
