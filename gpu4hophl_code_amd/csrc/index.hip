// index.hip -- the resident label index: device selection, scans and the
// per-vertex label sort, import/export and the label checkpoint format.
#include <cstring>  // rocPRIM's texture iterator uses memset on the host

#include <hip/hip_runtime.h>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_segmented_radix_sort.hpp>

#include <algorithm>
#include <cstdio>
#include <vector>

#include "device.hpp"

namespace hsdl {

int hip_fail(hipError_t e, const char *what, const char *file, int line) {
  char buf[512];
  snprintf(buf, sizeof(buf), "%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
  (void)hipGetLastError();
  return fail(e == hipErrorOutOfMemory ? HSDL_ENOMEM : HSDL_EDEVICE, buf);
}

int use_device(int32_t device, int *out_device) {
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count <= 0) {
    (void)hipGetLastError();
    return fail(HSDL_ENODEV, "no HIP device visible");
  }
  if (device >= 0) {
    if (device >= count) return fail(HSDL_ENODEV, "device ordinal out of range");
    HSDL_HIP(hipSetDevice(device));
  }
  int cur = 0;
  HSDL_HIP(hipGetDevice(&cur));
  if (out_device) *out_device = cur;
  return HSDL_OK;
}

template <class T>
static int scan_counts_impl(const T *d_cnt, uint64_t *d_off, uint32_t n, hipStream_t st) {
  // off[0] = 0; off[i+1] = sum(cnt[0..i]) via an inclusive scan into off+1.
  HSDL_HIP(hipMemsetAsync(d_off, 0, sizeof(uint64_t), st));
  if (n == 0) return HSDL_OK;
  size_t tmp_bytes = 0;
  HSDL_HIP(rocprim::inclusive_scan(nullptr, tmp_bytes, d_cnt, d_off + 1, (size_t)n,
                                   rocprim::plus<uint64_t>(), st));
  DevBuf tmp;
  int rc = tmp.alloc(tmp_bytes);
  if (rc) return rc;
  HSDL_HIP(rocprim::inclusive_scan(tmp.p, tmp_bytes, d_cnt, d_off + 1, (size_t)n,
                                   rocprim::plus<uint64_t>(), st));
  HSDL_HIP(hipStreamSynchronize(st));
  return HSDL_OK;
}

int scan_counts_u32(const uint32_t *d_cnt, uint64_t *d_off, uint32_t n, hipStream_t st) {
  return scan_counts_impl(d_cnt, d_off, n, st);
}
int scan_counts_u64(const uint64_t *d_cnt, uint64_t *d_off, uint32_t n, hipStream_t st) {
  return scan_counts_impl(d_cnt, d_off, n, st);
}

int segmented_sort_u32(void *tmp, size_t &tmp_bytes, const uint32_t *in, uint32_t *out,
                       uint64_t total, const uint64_t *d_off, uint32_t n, uint32_t end_bit,
                       hipStream_t st) {
  HSDL_HIP(rocprim::segmented_radix_sort_keys(tmp, tmp_bytes, in, out, (size_t)total, n, d_off,
                                              d_off + 1, 0, (int)end_bit, st));
  return HSDL_OK;
}

int segmented_sort_labels(uint64_t *d_lab, uint64_t total, const uint64_t *d_off, uint32_t n,
                          hipStream_t st) {
  if (total == 0 || n == 0) return HSDL_OK;
  DevBuf out;
  int rc = out.alloc(total * sizeof(uint64_t));
  if (rc) return rc;
  size_t tmp_bytes = 0;
  HSDL_HIP(rocprim::segmented_radix_sort_keys(nullptr, tmp_bytes, d_lab, out.as<uint64_t>(),
                                              (size_t)total, n, d_off, d_off + 1, 0, 64, st));
  DevBuf tmp;
  rc = tmp.alloc(tmp_bytes);
  if (rc) return rc;
  HSDL_HIP(rocprim::segmented_radix_sort_keys(tmp.p, tmp_bytes, d_lab, out.as<uint64_t>(),
                                              (size_t)total, n, d_off, d_off + 1, 0, 64, st));
  HSDL_HIP(hipMemcpyAsync(d_lab, out.p, total * sizeof(uint64_t), hipMemcpyDeviceToDevice, st));
  HSDL_HIP(hipStreamSynchronize(st));
  return HSDL_OK;
}

}  // namespace hsdl

using namespace hsdl;

extern "C" {

int hsdl_label_count(const hsdl_index *idx, uint64_t *total) {
  if (!idx || !total) return fail(HSDL_EINVAL, "null argument");
  *total = idx->total;
  return HSDL_OK;
}

int hsdl_index_info(const hsdl_index *idx, uint32_t *n, uint8_t *K) {
  if (!idx) return fail(HSDL_EINVAL, "null index");
  if (n) *n = idx->n;
  if (K) *K = idx->K;
  return HSDL_OK;
}

int hsdl_index_stats(const hsdl_index *idx, hsdl_stats *st) {
  if (!idx || !st) return fail(HSDL_EINVAL, "null argument");
  *st = idx->stats;
  return HSDL_OK;
}

static int export_labels_body(const hsdl_index *idx, uint64_t *off, uint32_t *hub, uint8_t *hop,
                       int32_t *dist) {
  if (!idx || !off) return fail(HSDL_EINVAL, "null argument");
  int rc = use_device(idx->device, nullptr);
  if (rc) return rc;
  HSDL_HIP(hipMemcpy(off, idx->off.p, sizeof(uint64_t) * (idx->n + 1ull), hipMemcpyDeviceToHost));
  if (idx->total == 0 || (!hub && !hop && !dist)) return HSDL_OK;
  std::vector<uint64_t> lab(idx->total);
  HSDL_HIP(hipMemcpy(lab.data(), idx->lab.p, sizeof(uint64_t) * idx->total, hipMemcpyDeviceToHost));
  for (uint64_t i = 0; i < idx->total; ++i) {
    if (hub) hub[i] = label_hub(lab[i]);
    if (hop) hop[i] = (uint8_t)label_hop(lab[i]);
    if (dist) dist[i] = (int32_t)label_dist(lab[i]);
  }
  return HSDL_OK;
}

int hsdl_export_labels(const hsdl_index *idx, uint64_t *off, uint32_t *hub, uint8_t *hop,
                       int32_t *dist) {
  return hsdl::guarded([&] { return export_labels_body(idx, off, hub, hop, dist); });
}

static int index_from_labels_body(uint32_t n, uint8_t K, const uint64_t *off, const uint32_t *hub,
                           const uint8_t *hop, const int32_t *dist, int32_t device,
                           hsdl_index **out) {
  if (!out || !off || K < 1 || K > kMaxK) return fail(HSDL_EINVAL, "bad argument");
  const uint64_t total = off[n];
  if (total && (!hub || !hop || !dist)) return fail(HSDL_EINVAL, "null label arrays");
  if (off[0] != 0) return fail(HSDL_EINVAL, "off[0] must be 0");
  for (uint32_t v = 0; v < n; ++v)
    if (off[v + 1] < off[v]) return fail(HSDL_EINVAL, "offsets must be non-decreasing");
  std::vector<uint64_t> lab(total);
  uint32_t dist_max = 0;
  for (uint64_t i = 0; i < total; ++i) {
    if (hub[i] >= n) return fail(HSDL_EINVAL, "hub id out of range");
    if (hop[i] > K) return fail(HSDL_ERANGE, "hop above K");
    if (dist[i] < 0 || (uint32_t)dist[i] > kMaxDist) return fail(HSDL_ERANGE, "dist outside 0..2^27-1");
    lab[i] = pack_label(hub[i], hop[i], (uint32_t)dist[i]);
    dist_max = std::max(dist_max, (uint32_t)dist[i]);
  }
  // per-vertex (hub, hop, dist) order; one label per (hub, hop)
  for (uint32_t v = 0; v < n; ++v) {
    std::sort(lab.begin() + off[v], lab.begin() + off[v + 1]);
    for (uint64_t i = off[v] + 1; i < off[v + 1]; ++i)
      if ((lab[i] >> kDistBits) == (lab[i - 1] >> kDistBits))
        return fail(HSDL_EINVAL, "two labels with the same (hub, hop) at one vertex");
    for (uint64_t i = off[v]; i < off[v + 1]; ++i)
      if (label_hop(lab[i]) == 0 && (label_hub(lab[i]) != v || label_dist(lab[i]) != 0))
        return fail(HSDL_EINVAL, "a hop-0 label must be the vertex's self label (v,0,0)");
  }
  int dev = 0;
  int rc = use_device(device, &dev);
  if (rc) return rc;
  hsdl_index *idx = new hsdl_index();
  idx->device = dev;
  idx->n = n;
  idx->K = K;
  idx->total = total;
  idx->dist_max = dist_max;
  if ((rc = idx->off.alloc(sizeof(uint64_t) * (n + 1ull))) ||
      (rc = idx->lab.alloc(sizeof(uint64_t) * std::max<uint64_t>(total, 1)))) {
    delete idx;
    return rc;
  }
  hipError_t e = hipMemcpy(idx->off.p, off, sizeof(uint64_t) * (n + 1ull), hipMemcpyHostToDevice);
  if (e == hipSuccess && total)
    e = hipMemcpy(idx->lab.p, lab.data(), sizeof(uint64_t) * total, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    delete idx;
    return hip_fail(e, "hipMemcpy", __FILE__, __LINE__);
  }
  idx->stats.labels_uncleaned = total;
  *out = idx;
  return HSDL_OK;
}

int hsdl_index_from_labels(uint32_t n, uint8_t K, const uint64_t *off, const uint32_t *hub,
                           const uint8_t *hop, const int32_t *dist, int32_t device,
                           hsdl_index **out) {
  return hsdl::guarded([&] { return index_from_labels_body(n, K, off, hub, hop, dist, device, out); });
}

// Checkpoint layout of Clean_pc/main.cu:11-57 (save_labels / load_labels):
// size_t n; per vertex: size_t count, then count x {int hub, parent, hop, dist}.
static int save_labels_body(const hsdl_index *idx, const char *path) {
  if (!idx || !path) return fail(HSDL_EINVAL, "null argument");
  std::vector<uint64_t> off(idx->n + 1ull);
  std::vector<uint32_t> hub(idx->total);
  std::vector<uint8_t> hop(idx->total);
  std::vector<int32_t> dist(idx->total);
  int rc = hsdl_export_labels(idx, off.data(), hub.data(), hop.data(), dist.data());
  if (rc) return rc;
  std::vector<uint32_t> par;
  if (idx->has_parents) {
    par.resize(idx->total);
    if ((rc = hsdl_export_parents(idx, par.data()))) return rc;
  }
  FILE *f = fopen(path, "wb");
  if (!f) return fail(HSDL_EIO, std::string("cannot open ") + path);
  size_t n = idx->n;
  bool ok = fwrite(&n, sizeof(size_t), 1, f) == 1;
  std::vector<int32_t> rec;
  for (uint32_t v = 0; ok && v < idx->n; ++v) {
    size_t c = off[v + 1] - off[v];
    ok = fwrite(&c, sizeof(size_t), 1, f) == 1;
    rec.assign(4 * c, 0);
    for (size_t i = 0; i < c; ++i) {
      rec[4 * i + 0] = (int32_t)hub[off[v] + i];
      rec[4 * i + 1] = par.empty() ? 0 : (int32_t)par[off[v] + i];
      rec[4 * i + 2] = hop[off[v] + i];
      rec[4 * i + 3] = dist[off[v] + i];
    }
    if (ok && c) ok = fwrite(rec.data(), sizeof(int32_t), 4 * c, f) == 4 * c;
  }
  if (fclose(f) != 0) ok = false;
  return ok ? HSDL_OK : fail(HSDL_EIO, "write failed");
}

int hsdl_save_labels(const hsdl_index *idx, const char *path) {
  return hsdl::guarded([&] { return save_labels_body(idx, path); });
}

static int load_labels_body(const char *path, uint8_t K, int32_t device, hsdl_index **out) {
  if (!path || !out) return fail(HSDL_EINVAL, "null argument");
  if (K < 1 || K > kMaxK) return fail(HSDL_EINVAL, "K must be in 1..15");
  FILE *f = fopen(path, "rb");
  if (!f) return fail(HSDL_EIO, std::string("cannot open ") + path);
  // the file's size bounds every count read from it, so a corrupt count is
  // refused before anything is sized from it
  uint64_t left = 0;
  if (fseek(f, 0, SEEK_END) == 0) {
    const long sz = ftell(f);
    left = sz > 0 ? (uint64_t)sz : 0;
  }
  if (fseek(f, 0, SEEK_SET) != 0) left = 0;
  size_t n = 0;
  bool ok = left >= sizeof(size_t) && fread(&n, sizeof(size_t), 1, f) == 1;
  if (ok) left -= sizeof(size_t);
  // each vertex holds at least its count word
  ok = ok && n <= 0xFFFFFFFFull && (uint64_t)n <= left / sizeof(size_t);
  int bad = HSDL_EIO;
  std::vector<uint64_t> off(1, 0);
  std::vector<uint32_t> hub, par;
  std::vector<uint8_t> hop;
  std::vector<int32_t> dist;
  std::vector<int32_t> rec;
  // a vertex's records sorted by (hub, hop, dist) with their parents, the
  // order the index keeps
  std::vector<std::pair<uint64_t, int32_t>> row;
  bool any_parent = false;
  for (size_t v = 0; ok && v < n; ++v) {
    size_t c = 0;
    ok = left >= sizeof(size_t) && fread(&c, sizeof(size_t), 1, f) == 1;
    if (ok) left -= sizeof(size_t);
    ok = ok && (uint64_t)c <= left / 16;  // 16 B per record
    if (!ok) break;
    left -= (uint64_t)c * 16;
    rec.resize(4 * c);
    if (c) ok = fread(rec.data(), sizeof(int32_t), 4 * c, f) == 4 * c;
    row.clear();
    for (size_t i = 0; ok && i < c; ++i) {
      if (rec[4 * i] < 0 || rec[4 * i + 2] < 0 || rec[4 * i + 3] < 0 ||
          (uint32_t)rec[4 * i + 3] > kMaxDist) {
        ok = false;
        break;
      }
      // the hop field of a packed label holds 0..31: a hop above K is
      // refused here, before packing could carry it into the hub bits
      if (rec[4 * i + 2] > (int32_t)K) {
        ok = false;
        bad = HSDL_ERANGE;
        break;
      }
      row.push_back({pack_label((uint32_t)rec[4 * i], (uint32_t)rec[4 * i + 2],
                                (uint32_t)rec[4 * i + 3]), rec[4 * i + 1]});
      any_parent = any_parent || rec[4 * i + 1] != 0;
    }
    std::sort(row.begin(), row.end());
    for (const auto &r : row) {
      hub.push_back(label_hub(r.first));
      hop.push_back((uint8_t)label_hop(r.first));
      dist.push_back((int32_t)label_dist(r.first));
      par.push_back((uint32_t)r.second);
    }
    off.push_back(hub.size());
  }
  fclose(f);
  if (!ok && bad == HSDL_ERANGE) return fail(HSDL_ERANGE, std::string("hop above K in ") + path);
  if (!ok) return fail(HSDL_EIO, std::string("malformed label file ") + path);
  int rc = hsdl_index_from_labels((uint32_t)n, K, off.data(), hub.data(), hop.data(), dist.data(),
                                  device, out);
  if (rc || !any_parent) return rc;
  if ((rc = hsdl_index_set_parents(*out, par.data()))) {
    hsdl_free_index(*out);
    *out = nullptr;
  }
  return rc;
}

int hsdl_load_labels(const char *path, uint8_t K, int32_t device, hsdl_index **out) {
  return hsdl::guarded([&] { return load_labels_body(path, K, device, out); });
}

void hsdl_free_index(hsdl_index *idx) {
  if (!idx) return;
  const int dev = idx->device;
  int cur = -1;
  if (hipGetDevice(&cur) == hipSuccess && cur != dev) (void)hipSetDevice(dev);
  delete idx;
  if (cur >= 0 && cur != dev) (void)hipSetDevice(cur);
}

}  // extern "C"
