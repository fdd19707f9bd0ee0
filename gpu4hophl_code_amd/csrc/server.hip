// server.hip -- continuous-batching (s,t,k) query server over a resident index.
//
// The reference answers one query per launch triple with a host sync in
// between (Query_mdnd/src/gpu_query.cu:139-155).  Here requests go into a host
// queue; a server thread closes a batch when it holds max_batch requests or
// its oldest request has waited max_wait_us, and launches it into one of
// `inflight` slots.  A slot owns pinned host staging (device-mapped, read and
// written by the query kernel directly; or, with copy staging, device buffers
// too) and its own stream (default priority), so batches that arrive while
// others run are answered concurrently; a batch waits only for a free slot,
// and the slot count is capped by the staging-memory budget.  Answers are the same
// kernel's as hsdl_query: a request's result does not depend on its batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <climits>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <string>
#include <cstdio>
#include <thread>
#include <vector>

#include "device.hpp"

namespace hsdl {

// A slot's requests travel as one packed block [s: 4n B][t: 4n B][k: n B]
// (one host-to-device copy per batch).
struct Slot {
  hipStream_t st = nullptr;
  hipEvent_t done = nullptr;
  uint8_t *h_req = nullptr;
  int32_t *h_out = nullptr;
  DevBuf d_req, d_out;
  std::vector<uint64_t> ids, arrival;
  size_t count = 0;
  bool busy = false;
  uint64_t t_close = 0, t_h2d = 0, t_kernel = 0, t_d2h = 0, t_launched = 0;
};

struct Result {
  uint64_t id;
  int32_t ans;
  uint64_t lat;
};

// Optional event trace (HSDL_SERVER_TRACE=<path>, appended at destroy), times
// in CLOCK_MONOTONIC ns:
//   B slot n oldest_arrival closed after_h2d after_launch after_d2h launched retired
//   G at gap      (worker loop iteration gap over 500 us)
//   R at lag      (replay submission later than its arrival by over 500 us)
struct TraceEvent {
  char kind;
  uint32_t slot;
  uint64_t v[8];
};

}  // namespace hsdl

struct hsdl_server {
  const hsdl_index *idx = nullptr;
  hsdl_server_opts o{};
  std::vector<hsdl::Slot> slots;
  // pending requests (ring of plain arrays guarded by mu)
  std::mutex mu;
  std::condition_variable cv;
  std::vector<uint32_t> q_s, q_t;
  std::vector<int8_t> q_k;
  std::vector<uint64_t> q_arr;
  uint64_t q_head = 0, q_tail = 0;  // absolute positions; id == position
  // finished results: a ring of queue_capacity entries, preallocated (no
  // allocation on the serving path).  The server thread is the only writer
  // and publishes r_tail; readers (poll, replay) take rmu and advance r_head.
  // Submission keeps submitted - polled <= queue_capacity, so the ring
  // cannot overrun.
  std::mutex rmu;
  std::vector<hsdl::Result> ring;
  std::atomic<uint64_t> r_head{0}, r_tail{0};
  std::atomic<uint64_t> completed{0}, batches{0}, batched_items{0};
  std::atomic<bool> stop{false};
  std::thread worker;
  int device = 0;
  int error = 0;
  bool tracing = false;
  std::string trace_path;
  std::mutex tmu;
  std::vector<hsdl::TraceEvent> trace;
  void note(const hsdl::TraceEvent &ev) {
    if (!tracing) return;
    std::lock_guard<std::mutex> g(tmu);
    if (trace.size() < (1u << 22)) trace.push_back(ev);
  }
};

namespace hsdl {

// Mapped staging: the kernel reads the packed requests from and writes the
// answers to pinned fine-grained host memory, so a batch costs one launch and
// one event record.  Copy staging moves them through device buffers.
static void launch_slot(hsdl_server *srv, Slot &sl) {
  const size_t n = sl.count;
  const bool copy = srv->o.staging == HSDL_STAGE_COPY;
  if (copy) {
    if (hipMemcpyAsync(sl.d_req.p, sl.h_req, n * 9, hipMemcpyHostToDevice, sl.st) != hipSuccess)
      srv->error = HSDL_EDEVICE;
  }
  sl.t_h2d = srv->tracing ? now_ns() : 0;
  const uint8_t *d = copy ? sl.d_req.as<uint8_t>() : sl.h_req;
  int32_t *out = copy ? sl.d_out.as<int32_t>() : sl.h_out;
  if (launch_query(srv->idx, n, (const uint32_t *)d, (const uint32_t *)(d + 4 * n),
                   (const int8_t *)(d + 8 * n), out, sl.st) != HSDL_OK)
    srv->error = HSDL_EDEVICE;
  sl.t_kernel = srv->tracing ? now_ns() : 0;
  if (copy) {
    if (hipMemcpyAsync(sl.h_out, sl.d_out.p, n * 4, hipMemcpyDeviceToHost, sl.st) != hipSuccess)
      srv->error = HSDL_EDEVICE;
  }
  sl.t_d2h = srv->tracing ? now_ns() : 0;
  if (hipEventRecord(sl.done, sl.st) != hipSuccess) srv->error = HSDL_EDEVICE;
  sl.busy = true;
}

static void worker_loop(hsdl_server *srv) {
  (void)hipSetDevice(srv->device);
  const uint64_t cap = srv->o.queue_capacity;
  const uint64_t wait_ns = (uint64_t)srv->o.max_wait_us * 1000ull;
  uint64_t last_iter = now_ns();
  while (true) {
    bool progressed = false;
    if (srv->tracing) {
      const uint64_t t = now_ns();
      if (t - last_iter > 500000) srv->note(TraceEvent{'G', 0, {t, t - last_iter}});
      last_iter = t;
    }
    // 1. retire finished slots
    for (Slot &sl : srv->slots) {
      if (!sl.busy) continue;
      const hipError_t qe = hipEventQuery(sl.done);
      if (qe == hipErrorNotReady) continue;
      if (qe != hipSuccess) {
        srv->error = HSDL_EDEVICE;
        for (size_t i = 0; i < sl.count; ++i) sl.h_out[i] = INT32_MAX;
      }
      const uint64_t t = now_ns();
      srv->note(TraceEvent{'B', (uint32_t)(&sl - srv->slots.data()),
                           {sl.count, sl.count ? sl.arrival[0] : 0, sl.t_close, sl.t_h2d,
                            sl.t_kernel, sl.t_d2h, sl.t_launched, t}});
      const uint64_t rt = srv->r_tail.load(std::memory_order_relaxed);
      for (size_t i = 0; i < sl.count; ++i)
        srv->ring[(rt + i) % cap] =
            Result{sl.ids[i], sl.h_out[i], t > sl.arrival[i] ? t - sl.arrival[i] : 0};
      srv->r_tail.store(rt + sl.count, std::memory_order_release);
      srv->completed += sl.count;
      sl.busy = false;
      sl.count = 0;
      progressed = true;
    }
    // 2. admit a batch into a free slot.  A batch closes when it is full,
    // when its oldest request has waited max_wait, or at once while fewer
    // than two batches are on the device (an idle GPU does not wait).
    Slot *free_slot = nullptr;
    uint32_t busy = 0;
    for (Slot &sl : srv->slots) {
      if (sl.busy) ++busy;
      else if (!free_slot) free_slot = &sl;
    }
    if (free_slot) {
      std::unique_lock<std::mutex> g(srv->mu);
      const uint64_t pending = srv->q_tail - srv->q_head;
      bool close = false;
      if (pending >= srv->o.max_batch) close = true;
      else if (pending > 0) {
        const uint64_t oldest = srv->q_arr[srv->q_head % cap];
        const uint64_t now = now_ns();
        if (busy < 2 || now >= oldest + wait_ns || srv->stop.load()) close = true;
      }
      if (close) {
        const size_t n = (size_t)std::min<uint64_t>(pending, srv->o.max_batch);
        Slot &sl = *free_slot;
        uint32_t *hs = (uint32_t *)sl.h_req;
        uint32_t *ht = (uint32_t *)(sl.h_req + 4 * n);
        int8_t *hk = (int8_t *)(sl.h_req + 8 * n);
        for (size_t i = 0; i < n; ++i) {
          const uint64_t p = (srv->q_head + i) % cap;
          hs[i] = srv->q_s[p];
          ht[i] = srv->q_t[p];
          hk[i] = srv->q_k[p];
          sl.ids[i] = srv->q_head + i;
          sl.arrival[i] = srv->q_arr[p];
        }
        srv->q_head += n;
        g.unlock();
        srv->cv.notify_all();
        sl.count = n;
        sl.t_close = now_ns();
        launch_slot(srv, sl);
        sl.t_launched = now_ns();
        srv->batches += 1;
        srv->batched_items += n;
        progressed = true;
      }
    }
    if (srv->stop.load()) {
      bool idle = true;
      for (Slot &sl : srv->slots) idle = idle && !sl.busy;
      std::lock_guard<std::mutex> g(srv->mu);
      if (idle && srv->q_tail == srv->q_head) break;
    }
    if (!progressed) std::this_thread::yield();
  }
}

}  // namespace hsdl

using namespace hsdl;

extern "C" {

static int server_create_body(const hsdl_index *idx, const hsdl_server_opts *opts, hsdl_server **out) {
  if (!idx || !out) return fail(HSDL_EINVAL, "null argument");
  hsdl_server_opts o;
  if (opts) o = *opts; else hsdl_server_opts_init(&o);
  if (o.max_batch == 0 || o.inflight == 0 || o.queue_capacity == 0)
    return fail(HSDL_EINVAL, "max_batch, inflight and queue_capacity must be positive");
  if (o.staging != HSDL_STAGE_MAPPED && o.staging != HSDL_STAGE_COPY)
    return fail(HSDL_EINVAL, "unknown staging mode");
  int rc = use_device(idx->device, nullptr);
  if (rc) return rc;
  // the index's query copy is made before serving starts, so no batch
  // allocates
  if ((rc = prepare_query(idx))) return rc;
  // admission by memory: each slot stages 13 B per request (device buffers
  // with copy staging; the same budget caps mapped host staging)
  const uint64_t per_slot = (uint64_t)o.max_batch * 13 + 4096;
  uint64_t budget = o.hbm_budget;
  if (!budget) {
    size_t fr = 0, tot = 0;
    HSDL_HIP(hipMemGetInfo(&fr, &tot));
    budget = fr / 4;
  }
  const uint64_t fit = budget / per_slot;
  if (fit == 0) return fail(HSDL_ENOMEM, "device budget below one batch slot");
  o.inflight = (uint32_t)std::min<uint64_t>(o.inflight, fit);
  hsdl_server *srv = new hsdl_server();
  srv->idx = idx;
  srv->o = o;
  srv->device = idx->device;
  srv->q_s.resize(o.queue_capacity);
  srv->q_t.resize(o.queue_capacity);
  srv->q_k.resize(o.queue_capacity);
  srv->q_arr.resize(o.queue_capacity);
  srv->ring.resize(o.queue_capacity);
  srv->slots.resize(o.inflight);
  for (Slot &sl : srv->slots) {
    hipError_t e = hipStreamCreateWithFlags(&sl.st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming);
    // mapped staging needs fine-grained (coherent) host memory: the kernel's
    // reads and writes go straight to host memory, never to a stale GPU cache
    const unsigned hflags = o.staging == HSDL_STAGE_MAPPED
                                ? (hipHostMallocMapped | hipHostMallocCoherent)
                                : hipHostMallocDefault;
    if (e == hipSuccess) e = hipHostMalloc((void **)&sl.h_req, o.max_batch * 9ull, hflags);
    if (e == hipSuccess) e = hipHostMalloc((void **)&sl.h_out, o.max_batch * 4ull, hflags);
    if (e != hipSuccess) {
      hsdl_server_destroy(srv);
      return hip_fail(e, "server slot setup", __FILE__, __LINE__);
    }
    if (o.staging == HSDL_STAGE_COPY &&
        ((rc = sl.d_req.alloc(o.max_batch * 9ull)) || (rc = sl.d_out.alloc(o.max_batch * 4ull)))) {
      hsdl_server_destroy(srv);
      return rc;
    }
    sl.ids.resize(o.max_batch);
    sl.arrival.resize(o.max_batch);
    // warm the slot: the first copy and launch on a fresh stream pay one-time
    // costs that would otherwise land on the first requests' latency
    std::memset(sl.h_req, 0, 9);
    sl.count = 1;
    launch_slot(srv, sl);
    e = hipStreamSynchronize(sl.st);
    sl.busy = false;
    sl.count = 0;
    if (e != hipSuccess || srv->error) {
      hsdl_server_destroy(srv);
      return hip_fail(e != hipSuccess ? e : hipErrorUnknown, "server slot warm-up", __FILE__,
                      __LINE__);
    }
  }
  if (const char *tp = getenv("HSDL_SERVER_TRACE")) {
    srv->tracing = true;
    srv->trace_path = tp;
    srv->trace.reserve(1u << 16);
  }
  srv->worker = std::thread(worker_loop, srv);
  *out = srv;
  return HSDL_OK;
}

int hsdl_server_create(const hsdl_index *idx, const hsdl_server_opts *opts, hsdl_server **out) {
  return hsdl::guarded([&] { return server_create_body(idx, opts, out); });
}

// Enqueues requests [0, n) when `partial` is false (all or nothing:
// HSDL_EBUSY when they do not all fit), or the longest prefix that fits when
// it is true; *taken receives the count enqueued.
static int enqueue_requests(hsdl_server *srv, size_t n, const uint32_t *s, const uint32_t *t,
                            const int8_t *k, const uint64_t *arrival_ns, uint64_t *first_id,
                            bool partial, size_t *taken) {
  if (!srv) return fail(HSDL_EINVAL, "null server");
  if (n && (!s || !t || !k)) return fail(HSDL_EINVAL, "null request arrays");
  for (size_t i = 0; i < n; ++i)
    if (s[i] >= srv->idx->n || t[i] >= srv->idx->n) return fail(HSDL_EINVAL, "vertex out of range");
  const uint64_t cap = srv->o.queue_capacity;
  const uint64_t now = now_ns();
  std::lock_guard<std::mutex> g(srv->mu);
  // pending, in flight and unpolled requests share the capacity
  const uint64_t used = srv->q_tail - srv->r_head.load(std::memory_order_acquire);
  const uint64_t room = used < cap ? cap - used : 0;
  if (!partial && n > room) return fail(HSDL_EBUSY, "request queue full");
  const size_t m = (size_t)std::min<uint64_t>(n, room);
  if (first_id) *first_id = srv->q_tail;
  for (size_t i = 0; i < m; ++i) {
    const uint64_t p = (srv->q_tail + i) % cap;
    srv->q_s[p] = s[i];
    srv->q_t[p] = t[i];
    srv->q_k[p] = k[i];
    srv->q_arr[p] = arrival_ns ? arrival_ns[i] : now;
  }
  srv->q_tail += m;
  if (taken) *taken = m;
  return HSDL_OK;
}

static int server_submit_body(hsdl_server *srv, size_t n, const uint32_t *s, const uint32_t *t,
                              const int8_t *k, const uint64_t *arrival_ns, uint64_t *first_id) {
  return enqueue_requests(srv, n, s, t, k, arrival_ns, first_id, false, nullptr);
}

int hsdl_server_submit(hsdl_server *srv, size_t n, const uint32_t *s, const uint32_t *t,
                       const int8_t *k, const uint64_t *arrival_ns, uint64_t *first_id) {
  return hsdl::guarded([&] { return server_submit_body(srv, n, s, t, k, arrival_ns, first_id); });
}

int hsdl_server_poll(hsdl_server *srv, size_t cap, uint64_t *ids, int32_t *answers,
                     uint64_t *latency_ns, size_t *n_out) {
  if (!srv || !n_out) return fail(HSDL_EINVAL, "null argument");
  std::lock_guard<std::mutex> g(srv->rmu);
  const uint64_t rcap = srv->o.queue_capacity;
  const uint64_t h = srv->r_head.load(std::memory_order_relaxed);
  const uint64_t t = srv->r_tail.load(std::memory_order_acquire);
  const size_t n = (size_t)std::min<uint64_t>(cap, t - h);
  for (size_t i = 0; i < n; ++i) {
    const Result &r = srv->ring[(h + i) % rcap];
    if (ids) ids[i] = r.id;
    if (answers) answers[i] = r.ans;
    if (latency_ns) latency_ns[i] = r.lat;
  }
  srv->r_head.store(h + n, std::memory_order_release);
  *n_out = n;
  return srv->error ? fail(srv->error, "a batch launch failed") : HSDL_OK;
}

int hsdl_server_drain(hsdl_server *srv) {
  if (!srv) return fail(HSDL_EINVAL, "null server");
  uint64_t target;
  {
    std::lock_guard<std::mutex> g(srv->mu);
    target = srv->q_tail;
  }
  while (srv->completed.load() < target) {
    if (srv->error) return fail(srv->error, "a batch launch failed");
    std::this_thread::yield();
  }
  return HSDL_OK;
}

int hsdl_server_stats_get(const hsdl_server *srv, hsdl_server_stats *st) {
  if (!srv || !st) return fail(HSDL_EINVAL, "null argument");
  st->submitted = srv->q_tail;
  st->completed = srv->completed.load();
  st->batches = srv->batches.load();
  st->mean_batch = st->batches ? (double)srv->batched_items.load() / st->batches : 0.0;
  return HSDL_OK;
}

void hsdl_server_destroy(hsdl_server *srv) {
  if (!srv) return;
  srv->stop = true;
  if (srv->worker.joinable()) srv->worker.join();
  for (Slot &sl : srv->slots) {
    if (sl.st) (void)hipStreamSynchronize(sl.st);
    if (sl.done) (void)hipEventDestroy(sl.done);
    if (sl.st) (void)hipStreamDestroy(sl.st);
    if (sl.h_req) (void)hipHostFree(sl.h_req);
    if (sl.h_out) (void)hipHostFree(sl.h_out);
  }
  if (srv->tracing && !srv->trace.empty()) {
    if (FILE *f = fopen(srv->trace_path.c_str(), "a")) {
      for (const TraceEvent &ev : srv->trace) {
        if (ev.kind == 'B') {
          fprintf(f, "B %u", ev.slot);
          for (uint64_t x : ev.v) fprintf(f, " %llu", (unsigned long long)x);
          fputc('\n', f);
        } else {
          fprintf(f, "%c %llu %llu\n", ev.kind, (unsigned long long)ev.v[0],
                  (unsigned long long)ev.v[1]);
        }
      }
      fclose(f);
    }
  }
  delete srv;
}

static int server_replay_body(hsdl_server *srv, size_t n, const uint32_t *s, const uint32_t *t,
                       const int8_t *k, const uint64_t *arrival_offset_ns, int32_t *answers,
                       uint64_t *latency_ns) {
  if (!srv || (n && (!s || !t || !k || !arrival_offset_ns))) return fail(HSDL_EINVAL, "null argument");
  for (size_t i = 1; i < n; ++i)
    if (arrival_offset_ns[i] < arrival_offset_ns[i - 1])
      return fail(HSDL_EINVAL, "arrival offsets must be non-decreasing");
  uint64_t base_id = 0;
  {
    std::lock_guard<std::mutex> g(srv->mu);
    if (srv->q_tail != srv->q_head) return fail(HSDL_ESTATE, "server has pending requests");
    base_id = srv->q_tail;
  }
  {
    // earlier unpolled results are dropped
    std::lock_guard<std::mutex> g(srv->rmu);
    srv->r_head.store(srv->r_tail.load(std::memory_order_acquire), std::memory_order_release);
  }
  const uint64_t start = now_ns();
  const uint64_t rcap = srv->o.queue_capacity;
  std::vector<uint64_t> arr;
  size_t sent = 0, got = 0;
  while (got < n) {
    const uint64_t now = now_ns();
    size_t upto = sent;
    while (upto < n && start + arrival_offset_ns[upto] <= now) ++upto;
    if (upto > sent) {
      if (srv->tracing && now > start + arrival_offset_ns[sent] + 500000)
        srv->note(TraceEvent{'R', 0, {now, now - start - arrival_offset_ns[sent]}});
      arr.resize(upto - sent);
      for (size_t i = sent; i < upto; ++i) arr[i - sent] = start + arrival_offset_ns[i];
      // as many of the due requests as the queue has room for: a backlog
      // larger than the queue (a burst, or a server that fell behind) goes in
      // as results are drained below, instead of being retried all or nothing
      size_t taken = 0;
      int rc = enqueue_requests(srv, upto - sent, s + sent, t + sent, k + sent, arr.data(), nullptr,
                                true, &taken);
      if (rc != HSDL_OK) return rc;
      sent += taken;
    }
    size_t fresh = 0;
    {
      std::lock_guard<std::mutex> g(srv->rmu);
      const uint64_t h = srv->r_head.load(std::memory_order_relaxed);
      const uint64_t t = srv->r_tail.load(std::memory_order_acquire);
      for (uint64_t p = h; p < t; ++p) {
        const Result &r = srv->ring[p % rcap];
        const uint64_t i = r.id - base_id;
        if (i < n) {
          if (answers) answers[i] = r.ans;
          if (latency_ns) latency_ns[i] = r.lat;
          ++got;
        }
      }
      fresh = (size_t)(t - h);
      srv->r_head.store(t, std::memory_order_release);
    }
    if (srv->error) return fail(srv->error, "a batch launch failed");
    if (fresh == 0 && upto == sent) std::this_thread::yield();
  }
  return HSDL_OK;
}

int hsdl_server_replay(hsdl_server *srv, size_t n, const uint32_t *s, const uint32_t *t,
                       const int8_t *k, const uint64_t *arrival_offset_ns, int32_t *answers,
                       uint64_t *latency_ns) {
  return hsdl::guarded([&] { return server_replay_body(srv, n, s, t, k, arrival_offset_ns, answers, latency_ns); });
}

}  // extern "C"
