// common.hpp -- shared host-side definitions of the HSDL engine.
#pragma once

#include <cstdint>
#include <exception>
#include <new>
#include <string>
#include <vector>

#include "hsdl.h"

struct hsdl_graph {
  uint32_t n = 0;
  std::vector<uint64_t> off;  // n+1
  std::vector<uint32_t> col;  // neighbour ids, sorted per row
  std::vector<int32_t> w;     // weights >= 1
  int32_t max_w = 0;
  // device copy made by hsdl_graph_upload (off u64[n+1], adj (col,w) u32x2)
  void *d_off = nullptr;
  void *d_adj = nullptr;
  // the generation's dense dedupe slots, kept between builds on this graph
  // (tens of GB: freeing and reallocating them per build stalled the
  // allocations that followed)
  void *d_slots = nullptr;
  size_t d_slots_bytes = 0;
  // what d_slots hold after the last build that ended normally: the slot
  // encoding they were written with (0: unknown, clear them first), the next
  // attempt tag of the count-down, the hub rows cleared at the last full
  // clear (rows above it may hold values of any tag) and the rows used since
  // the last clear
  uint64_t slot_key = 0, slot_tag_next = 0;
  uint32_t slot_rows_valid = 0, slot_rows_used = 0;
  // the last build's label pool and its 32-bit copy, kept for the next build
  // on this graph for the same reason (and so that build starts at the size
  // the pool grew to)
  void *d_pool = nullptr, *d_pool32 = nullptr;
  size_t d_pool_bytes = 0, d_pool32_bytes = 0;
  // list and tile-arena sizes the last build on this graph grew to: the next
  // build starts there instead of rolling batches back to grow them again
  uint32_t hint_ccap = 0, hint_fcap = 0;
  uint64_t hint_arena = 0;
  // the generation's other working buffers (per-vertex state, candidate and
  // frontier lists, tile arena, re-sort slices), kept for the next build on
  // this graph: freeing and allocating gigabytes inside every build now and
  // then cost up to a second there
  static constexpr int kScratch = 32;
  void *scratch_p[kScratch] = {};
  size_t scratch_bytes[kScratch] = {};
  int d_device = -1;
  void (*d_free)(hsdl_graph *) = nullptr;
  ~hsdl_graph() {
    if (d_free) d_free(this);
  }
  void drop_device_copy() {
    if (d_free) d_free(this);
    d_free = nullptr;
  }
};

namespace hsdl {

// Thread-local error detail for hsdl_last_error().
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

// Runs the body of a C ABI entry point: a host allocation failure
// (std::bad_alloc, e.g. a vector sized from a corrupt file) or any other C++
// exception becomes a status code, since nothing may unwind through the ABI
// into a C, ctypes or cgo caller.
template <class F>
int guarded(F &&body) noexcept {
  try {
    return body();
  } catch (const std::bad_alloc &) {
    return fail(HSDL_ENOMEM, "host allocation failed");
  } catch (const std::exception &e) {
    return fail(HSDL_EINVAL, std::string("host error: ") + e.what());
  } catch (...) {
    return fail(HSDL_EINVAL, "unexpected host error");
  }
}

// Wall clock in ns (CLOCK_MONOTONIC).
uint64_t now_ns();

// ---------------------------------------------------------------- PRNG
// splitmix64 (Steele, Lea, Flood 2014) seeds xoshiro256** (Blackman, Vigna
// 2018); bounded integers use Lemire's nearly-divisionless method (2019).
struct SplitMix64 {
  uint64_t s;
  explicit SplitMix64(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
};

struct Xoshiro256ss {
  uint64_t s[4];
  explicit Xoshiro256ss(uint64_t seed) {
    SplitMix64 sm(seed);
    for (auto &x : s) x = sm.next();
  }
  static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  // uniform in [0, range), range >= 1
  uint32_t bounded(uint32_t range) {
    uint64_t m = (uint64_t)(uint32_t)(next() >> 32) * range;
    uint32_t l = (uint32_t)m;
    if (l < range) {
      uint32_t t = (uint32_t)(-range) % range;
      while (l < t) {
        m = (uint64_t)(uint32_t)(next() >> 32) * range;
        l = (uint32_t)m;
      }
    }
    return (uint32_t)(m >> 32);
  }
};

// Graph construction helpers (graph.cpp).
// Builds symmetric CSR from canonical unique edges (a<=b, sorted by (a,b)).
void csr_from_unique_edges(uint32_t n, const std::vector<uint64_t> &keys,
                           const std::vector<int32_t> &w, hsdl_graph &g);
// (degree desc, id asc) order; fills new_to_old and rebuilds g in place.
void reorder_by_degree(hsdl_graph &g, std::vector<uint32_t> &new_to_old);

}  // namespace hsdl
