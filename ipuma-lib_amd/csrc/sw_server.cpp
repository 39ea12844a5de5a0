// sw_server.cpp -- single-GPU request server over the engine.
//
// Replaces the reference's offline batch loop and multi-device driver
// (ipusw/ipumultidriver.hpp:10-145, ipusw/run_comparison.hpp:67-172) and its
// asynchronous submit / join (src/ipu_batch_affine.h:105-108): requests of any
// size arrive from any number of host threads, their sequences are encoded and
// copied once into an HBM arena (admitted against a byte budget), a dispatcher
// thread merges pending requests into one device batch, and a completer thread
// retires batches in order, copying results back and freeing arena space.
//
//   submit -> [pending FIFO] -> dispatcher: batch = oldest requests up to
//   max_batch_pairs, launched when the GPU is idle, the batch is full, or the
//   oldest request has waited max_wait_us -> [in-flight FIFO, <= max_inflight]
//   -> completer: wait for the batch's event, scatter results, free the arena
//   -> wait(ticket) returns the request's results and latency.
//
// Requests leave the arena in submission order (batches take the oldest
// pending requests and complete in stream order), so the arena is a ring.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/ipuma_hip.h"
#include "sw_device.h"
#include "sw_engine_internal.h"

using namespace ipuma;

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
  int64_t id = 0;
  int32_t n = 0;
  int64_t arena_off = 0, arena_bytes = 0;  // ring slice holding the codes
  std::vector<int64_t> a_off, b_off;       // absolute offsets into the arena
  std::vector<int32_t> a_len, b_len;
  int64_t cells = 0;
  Clock::time_point t_submit, t_done;
  std::vector<int32_t> scores, a_range, b_range;
  int status = 0;  // 0 pending, 1 in a batch, 2 done, < 0 error code
};

struct Slot {  // descriptor / result buffers of one in-flight batch
  int64_t *d_aoff = nullptr, *d_boff = nullptr;
  int32_t *d_alen = nullptr, *d_blen = nullptr, *d_res = nullptr;
  int64_t *h_off = nullptr;  // pinned: a_off | b_off
  int32_t *h_len = nullptr;  // pinned: a_len | b_len
  int32_t *h_res = nullptr;  // pinned: scores | a_range | b_range
  int32_t *h_err = nullptr;  // pinned: the batch's kernel error words (forward, reverse)
  hipEvent_t done = nullptr;
  bool busy = false;
};

struct Batch {
  std::vector<Request *> reqs;
  int32_t n = 0;
  int slot = 0;
  int64_t cells = 0;
  Clock::time_point t_launch;
};

}  // namespace

struct ipuma_server {
  ipuma_server_config cfg{};
  ipuma_engine *e = nullptr;
  hipStream_t stream = nullptr;       // the engine's stream: batches run in order on it
  uint8_t *arena = nullptr;
  int64_t arena_size = 0;
  struct Span { int64_t off, bytes; bool freed; };
  std::deque<Span> ring_live;         // allocated slices, oldest first
  int64_t ring_used = 0;
  struct Stage { uint8_t *p; int64_t cap; };
  std::vector<Stage> stage_pool;      // pinned staging buffers not in use (under stage_mu)
  std::mutex stage_mu;
  std::vector<Slot> slots;
  std::mutex mu;
  std::condition_variable cv_dispatch, cv_complete, cv_done, cv_space;
  std::deque<Request *> pending;
  int64_t pending_pairs = 0;
  std::deque<Batch *> inflight;
  std::unordered_map<int64_t, Request *> reqs;
  std::thread dispatcher, completer;
  bool closing = false;
  bool dispatcher_done = false;
  int64_t next_id = 1;
  ipuma_server_stats st{};
  Clock::time_point t_first_launch, t_last_done;
  bool launched = false;
};

namespace {

void free_slot(Slot &s) {
  if (s.d_aoff) (void)hipFree(s.d_aoff);
  if (s.d_boff) (void)hipFree(s.d_boff);
  if (s.d_alen) (void)hipFree(s.d_alen);
  if (s.d_blen) (void)hipFree(s.d_blen);
  if (s.d_res) (void)hipFree(s.d_res);
  if (s.h_off) (void)hipHostFree(s.h_off);
  if (s.h_len) (void)hipHostFree(s.h_len);
  if (s.h_res) (void)hipHostFree(s.h_res);
  if (s.h_err) (void)hipHostFree(s.h_err);
  if (s.done) (void)hipEventDestroy(s.done);
  s = Slot{};
}

int alloc_slot(Slot &s, int32_t cap) {
  const size_t n = (size_t)std::max(cap, 1);
  if (hipMalloc(&s.d_aoff, n * 8) != hipSuccess || hipMalloc(&s.d_boff, n * 8) != hipSuccess ||
      hipMalloc(&s.d_alen, n * 4) != hipSuccess || hipMalloc(&s.d_blen, n * 4) != hipSuccess ||
      hipMalloc(&s.d_res, n * 12) != hipSuccess ||
      hipHostMalloc(&s.h_off, n * 16, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc(&s.h_len, n * 8, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc(&s.h_res, n * 12, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc(&s.h_err, 16, hipHostMallocDefault) != hipSuccess ||
      hipEventCreateWithFlags(&s.done, hipEventDisableTiming) != hipSuccess)
    return set_last_error(IPUMA_ENOMEM, "server: batch buffers for %d pairs", cap);
  return IPUMA_OK;
}

// Ring allocation of `bytes` (multiple of 4); -1 when the arena cannot hold it now.
// Live slices occupy [tail, head) when the newest starts at or after the oldest,
// else [tail, end) and [0, head).
int64_t ring_alloc(ipuma_server *s, int64_t bytes) {
  if (bytes > s->arena_size) return -1;
  int64_t off = 0;
  if (!s->ring_live.empty()) {
    const int64_t tail = s->ring_live.front().off;
    const auto &nw = s->ring_live.back();
    const int64_t head = nw.off + nw.bytes;
    if (nw.off >= tail) {  // no wrap: free [head, end) and [0, tail)
      if (head + bytes <= s->arena_size) off = head;
      else if (bytes <= tail) off = 0;
      else return -1;
    } else {  // wrapped: free [head, tail)
      if (head + bytes <= tail) off = head;
      else return -1;
    }
  }
  s->ring_live.push_back({off, bytes, false});
  s->ring_used += bytes;
  return off;
}

// Marks the slice at `off` free and releases the freed prefix of the ring.
void ring_free(ipuma_server *s, int64_t off) {
  for (auto &sp : s->ring_live)
    if (sp.off == off && !sp.freed) {
      sp.freed = true;
      break;
    }
  while (!s->ring_live.empty() && s->ring_live.front().freed) {
    s->ring_used -= s->ring_live.front().bytes;
    s->ring_live.pop_front();
  }
}

void dispatch_loop(ipuma_server *s) {
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    s->cv_dispatch.wait(lk, [&] { return s->closing || !s->pending.empty(); });
    if (s->pending.empty()) {  // closing with nothing left
      s->dispatcher_done = true;
      s->cv_complete.notify_all();
      return;
    }
    const bool idle = s->inflight.empty();
    const auto age = std::chrono::duration_cast<std::chrono::microseconds>(
                         Clock::now() - s->pending.front()->t_submit).count();
    const bool full = s->pending_pairs >= s->cfg.max_batch_pairs;
    if (!(idle || full || age >= s->cfg.max_wait_us || s->closing)) {
      // sleep until the oldest request is due, or a completion / submission
      s->cv_dispatch.wait_for(lk, std::chrono::microseconds(s->cfg.max_wait_us - age));
      continue;
    }
    if ((int)s->inflight.size() >= s->cfg.max_inflight) {
      s->cv_dispatch.wait(lk, [&] { return (int)s->inflight.size() < s->cfg.max_inflight; });
    }
    int slot = -1;
    for (int i = 0; i < (int)s->slots.size(); ++i)
      if (!s->slots[i].busy) { slot = i; break; }
    if (slot < 0) continue;
    auto *b = new Batch();
    b->slot = slot;
    while (!s->pending.empty()) {
      Request *r = s->pending.front();
      if (!b->reqs.empty() && (!s->cfg.merge_requests || b->n + r->n > s->cfg.max_batch_pairs))
        break;
      s->pending.pop_front();
      s->pending_pairs -= r->n;
      r->status = 1;
      b->reqs.push_back(r);
      b->n += r->n;
      b->cells += r->cells;
    }
    Slot &sl = s->slots[slot];
    sl.busy = true;
    // descriptors of the merged batch
    int64_t k = 0;
    for (Request *r : b->reqs) {
      std::memcpy(sl.h_off + k, r->a_off.data(), r->n * 8);
      std::memcpy(sl.h_off + b->n + k, r->b_off.data(), r->n * 8);
      std::memcpy(sl.h_len + k, r->a_len.data(), r->n * 4);
      std::memcpy(sl.h_len + b->n + k, r->b_len.data(), r->n * 4);
      k += r->n;
    }
    lk.unlock();
    const int32_t n = b->n;
    int rc = IPUMA_OK;
    if (hipMemcpyAsync(sl.d_aoff, sl.h_off, n * 8, hipMemcpyHostToDevice, s->stream) != hipSuccess ||
        hipMemcpyAsync(sl.d_boff, sl.h_off + n, n * 8, hipMemcpyHostToDevice, s->stream) != hipSuccess ||
        hipMemcpyAsync(sl.d_alen, sl.h_len, n * 4, hipMemcpyHostToDevice, s->stream) != hipSuccess ||
        hipMemcpyAsync(sl.d_blen, sl.h_len + n, n * 4, hipMemcpyHostToDevice, s->stream) != hipSuccess)
      rc = IPUMA_EDEVICE;
    if (rc == IPUMA_OK)
      rc = ipuma_compare_device(s->e, n, s->arena, s->arena_size, sl.d_aoff, sl.d_alen, sl.d_boff,
                                sl.d_blen, sl.d_res, sl.d_res + n, sl.d_res + 2 * (size_t)n,
                                s->stream);
    // the batch's error words, copied in stream order before the next batch
    // on the stream clears them
    const int32_t *w0 = nullptr, *w1 = nullptr;
    sl.h_err[0] = sl.h_err[1] = 0;
    if (rc == IPUMA_OK) engine_error_words(s->e, &w0, &w1);
    if (rc == IPUMA_OK &&
        ((w0 && hipMemcpyAsync(sl.h_err, w0, 4, hipMemcpyDeviceToHost, s->stream) != hipSuccess) ||
         (w1 && hipMemcpyAsync(sl.h_err + 1, w1, 4, hipMemcpyDeviceToHost, s->stream) != hipSuccess) ||
         hipMemcpyAsync(sl.h_res, sl.d_res, (size_t)n * 12, hipMemcpyDeviceToHost, s->stream) !=
             hipSuccess ||
         hipEventRecord(sl.done, s->stream) != hipSuccess))
      rc = IPUMA_EDEVICE;
    lk.lock();
    b->t_launch = Clock::now();
    if (!s->launched) {
      s->launched = true;
      s->t_first_launch = b->t_launch;
    }
    if (rc != IPUMA_OK) {
      // the batch never ran: fail its requests at once
      for (Request *r : b->reqs) {
        r->status = rc;
        r->t_done = Clock::now();
        ring_free(s, r->arena_off);
      }
      sl.busy = false;
      delete b;
      s->cv_done.notify_all();
      s->cv_space.notify_all();
      continue;
    }
    s->inflight.push_back(b);
    s->cv_complete.notify_one();
  }
}

void complete_loop(ipuma_server *s) {
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    s->cv_complete.wait(lk, [&] { return !s->inflight.empty() || s->dispatcher_done; });
    if (s->inflight.empty()) return;  // the dispatcher has exited and everything retired
    Batch *b = s->inflight.front();
    Slot &sl = s->slots[b->slot];
    lk.unlock();
    int rc = hipEventSynchronize(sl.done) == hipSuccess ? IPUMA_OK : IPUMA_EDEVICE;
    if (rc == IPUMA_OK) rc = engine_error_status(sl.h_err[0] | sl.h_err[1]);
    const auto now = Clock::now();
    int64_t k = 0;
    for (Request *r : b->reqs) {
      if (rc == IPUMA_OK) {
        r->scores.assign(sl.h_res + k, sl.h_res + k + r->n);
        r->a_range.assign(sl.h_res + b->n + k, sl.h_res + b->n + k + r->n);
        r->b_range.assign(sl.h_res + 2 * (size_t)b->n + k, sl.h_res + 2 * (size_t)b->n + k + r->n);
      }
      k += r->n;
    }
    lk.lock();
    for (Request *r : b->reqs) {
      r->status = rc == IPUMA_OK ? 2 : rc;
      r->t_done = now;
      ring_free(s, r->arena_off);
    }
    s->st.batches += 1;
    s->st.pairs += b->n;
    s->st.cells += b->cells;
    s->st.max_batch_pairs_seen = std::max<int64_t>(s->st.max_batch_pairs_seen, b->n);
    s->st.busy_ms += std::chrono::duration<double, std::milli>(now - b->t_launch).count();
    s->t_last_done = now;
    sl.busy = false;
    s->inflight.pop_front();
    delete b;
    s->cv_done.notify_all();
    s->cv_space.notify_all();
    s->cv_dispatch.notify_one();
  }
}

}  // namespace

extern "C" {

void ipuma_default_server_config(ipuma_server_config *c) {
  if (!c) return;
  c->hbm_budget_bytes = (int64_t)64 << 30;
  c->max_batch_pairs = 1 << 16;
  c->max_wait_us = 2000;
  c->max_inflight = 2;
  c->block_on_full = 1;
  c->merge_requests = 1;
}

static int ipuma_server_create_body(const ipuma_sw_config *sw, const ipuma_algo_config *algo,
                        const ipuma_server_config *cfg, ipuma_server **out) {
  if (!out) return set_last_error(IPUMA_EINVAL, "out is NULL");
  *out = nullptr;
  auto *s = new (std::nothrow) ipuma_server();
  if (!s) return set_last_error(IPUMA_ENOMEM, "host allocation failed");
  if (cfg) s->cfg = *cfg; else ipuma_default_server_config(&s->cfg);
  if (s->cfg.hbm_budget_bytes < 1024 || s->cfg.max_batch_pairs < 1 ||
      s->cfg.max_inflight < 1 || s->cfg.max_inflight > 8 || s->cfg.max_wait_us < 0) {
    delete s;
    return set_last_error(IPUMA_EINVAL, "invalid server configuration");
  }
  ipuma_algo_config al;
  if (algo) al = *algo; else ipuma_default_algo_config(&al);
  al.tiles_used = 1 << 30;  // capacity comes from the server's own limits
  al.max_batches = 1;
  int rc = ipuma_engine_create(sw, &al, &s->e);
  if (rc) {
    delete s;
    return rc;
  }
  s->stream = engine_stream(s->e);
  s->arena_size = s->cfg.hbm_budget_bytes / 4 * 4;
  if (hipMalloc(&s->arena, (size_t)s->arena_size + 64) != hipSuccess) {
    ipuma_server_destroy(s);
    return set_last_error(IPUMA_ENOMEM, "server: HBM arena of %lld bytes",
                          (long long)s->arena_size);
  }
  s->slots.resize(s->cfg.max_inflight);
  for (Slot &sl : s->slots)
    if ((rc = alloc_slot(sl, s->cfg.max_batch_pairs))) {
      ipuma_server_destroy(s);
      return rc;
    }
  // a thread that cannot start: the one already running is stopped and joined
  // and the server freed before the error is reported
  try {
    s->dispatcher = start_thread([s] { dispatch_loop(s); });
    s->completer = start_thread([s] { complete_loop(s); });
  } catch (const std::system_error &x) {
    ipuma_server_destroy(s);
    return set_last_error(IPUMA_EDEVICE, "server: thread start failed: %s", x.what());
  }
  *out = s;
  return IPUMA_OK;
}

void ipuma_server_destroy(ipuma_server *s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->closing = true;
  }
  s->cv_dispatch.notify_all();
  s->cv_complete.notify_all();
  s->cv_space.notify_all();
  if (s->dispatcher.joinable()) s->dispatcher.join();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->dispatcher_done = true;  // also when the dispatcher never started
  }
  s->cv_complete.notify_all();
  if (s->completer.joinable()) s->completer.join();
  for (auto &kv : s->reqs) delete kv.second;
  for (Slot &sl : s->slots) free_slot(sl);
  if (s->arena) (void)hipFree(s->arena);
  for (auto &stg : s->stage_pool) (void)hipHostFree(stg.p);
  if (s->e) ipuma_engine_destroy(s->e);
  delete s;
}

static int ipuma_server_submit_body(ipuma_server *s, int32_t n, const char *a_chars, const int64_t *a_offsets,
                        const char *b_chars, const int64_t *b_offsets, int64_t *ticket) {
  if (!s || !ticket) return set_last_error(IPUMA_EINVAL, "server or ticket is NULL");
  *ticket = 0;
  if (n < 1 || n > s->cfg.max_batch_pairs)
    return set_last_error(IPUMA_EINVAL, "a request holds 1..%d comparisons, got %d",
                          s->cfg.max_batch_pairs, n);
  if (!a_chars || !a_offsets || !b_chars || !b_offsets)
    return set_last_error(IPUMA_EINVAL, "NULL buffer");
  auto *r = new Request();
  r->n = n;
  r->a_off.resize(n); r->b_off.resize(n); r->a_len.resize(n); r->b_len.resize(n);
  int64_t bytes = 16;
  for (int32_t k = 0; k < n; ++k) {
    const int64_t la = a_offsets[k + 1] - a_offsets[k], lb = b_offsets[k + 1] - b_offsets[k];
    if (la < 0 || lb < 0 || la > kMaxSeqLen || lb > kMaxSeqLen) {
      delete r;
      return set_last_error(la < 0 || lb < 0 ? IPUMA_EINVAL : IPUMA_ETOOLONG,
                            "comparison %d: lengths %lld / %lld outside 0..%d", k, (long long)la,
                            (long long)lb, kMaxSeqLen);
    }
    r->a_len[k] = (int32_t)la;
    r->b_len[k] = (int32_t)lb;
    r->cells += la * lb;
    bytes += (la + 3) / 4 * 4 + (lb + 3) / 4 * 4;
  }
  r->arena_bytes = bytes;
  // admission against the HBM budget
  {
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->closing) {
      delete r;
      return set_last_error(IPUMA_ECLOSED, "server is shutting down");
    }
    int64_t off;
    while ((off = ring_alloc(s, bytes)) < 0) {
      if (!s->cfg.block_on_full || bytes > s->arena_size) {
        delete r;
        return set_last_error(IPUMA_EBUSY, "request of %lld bytes does not fit the HBM budget now",
                              (long long)bytes);
      }
      s->cv_space.wait(lk);
      if (s->closing) {
        delete r;
        return set_last_error(IPUMA_ECLOSED, "server is shutting down");
      }
    }
    r->arena_off = off;
    r->id = s->next_id++;
    s->reqs[r->id] = r;
  }
  // encode on the host into a pinned staging buffer of this call (submitting
  // threads encode and copy in parallel), one copy to the arena
  {
    ipuma_server::Stage stg{nullptr, 0};
    {
      std::lock_guard<std::mutex> sl(s->stage_mu);
      if (!s->stage_pool.empty()) {
        stg = s->stage_pool.back();
        s->stage_pool.pop_back();
      }
    }
    if (stg.cap < bytes) {
      if (stg.p) (void)hipHostFree(stg.p);
      stg.cap = std::max<int64_t>(bytes, (int64_t)4 << 20);
      if (hipHostMalloc(&stg.p, (size_t)stg.cap, hipHostMallocDefault) != hipSuccess) {
        stg.p = nullptr;
        stg.cap = 0;
      }
    }
    int rc = IPUMA_OK;
    if (!stg.p) {
      rc = set_last_error(IPUMA_ENOMEM, "server: pinned staging of %lld bytes", (long long)bytes);
    } else {
      const uint8_t *tab = engine_code_table(s->e);
      int64_t pos = 0;
      uint8_t *dst = stg.p;
      for (int32_t k = 0; k < n; ++k) {
        for (int side = 0; side < 2; ++side) {
          const char *src = side == 0 ? a_chars + a_offsets[k] : b_chars + b_offsets[k];
          const int64_t len = side == 0 ? r->a_len[k] : r->b_len[k];
          (side == 0 ? r->a_off : r->b_off)[k] = r->arena_off + pos;
          for (int64_t i = 0; i < len; ++i) dst[pos + i] = tab[(uint8_t)src[i]];
          const int64_t padded = (len + 3) / 4 * 4;
          for (int64_t i = len; i < padded; ++i) dst[pos + i] = 0;
          pos += padded;
        }
      }
      std::memset(dst + pos, 0, (size_t)(bytes - pos));
      if (hipMemcpy(s->arena + r->arena_off, dst, (size_t)bytes, hipMemcpyHostToDevice) !=
          hipSuccess)
        rc = set_last_error(IPUMA_EDEVICE, "server: upload of a request failed");
    }
    if (stg.p) {
      std::lock_guard<std::mutex> sl(s->stage_mu);
      s->stage_pool.push_back(stg);
    }
    if (rc) {
      std::lock_guard<std::mutex> lk(s->mu);
      r->status = rc;
      r->t_done = Clock::now();
      ring_free(s, r->arena_off);  // never queued
      *ticket = r->id;
      return rc;
    }
  }
  {
    std::lock_guard<std::mutex> lk(s->mu);
    r->t_submit = Clock::now();
    s->pending.push_back(r);
    s->pending_pairs += n;
    s->st.requests += 1;
    *ticket = r->id;
  }
  s->cv_dispatch.notify_one();
  return IPUMA_OK;
}

static int ipuma_server_wait_body(ipuma_server *s, int64_t ticket, int32_t *scores, int32_t *a_range,
                      int32_t *b_range, double *latency_ms) {
  if (!s) return set_last_error(IPUMA_EINVAL, "server is NULL");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return set_last_error(IPUMA_EINVAL, "unknown ticket %lld", (long long)ticket);
  Request *r = it->second;
  s->cv_done.wait(lk, [&] { return r->status == 2 || r->status < 0; });
  const int st = r->status;
  if (latency_ms)
    *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
  if (st == 2) {
    if (scores) std::memcpy(scores, r->scores.data(), r->n * 4);
    if (a_range) std::memcpy(a_range, r->a_range.data(), r->n * 4);
    if (b_range) std::memcpy(b_range, r->b_range.data(), r->n * 4);
  }
  s->reqs.erase(it);
  delete r;
  return st == 2 ? IPUMA_OK : set_last_error(st, "request %lld failed", (long long)ticket);
}

static int ipuma_server_get_stats_body(ipuma_server *s, ipuma_server_stats *out) {
  if (!s || !out) return set_last_error(IPUMA_EINVAL, "bad args");
  std::lock_guard<std::mutex> lk(s->mu);
  *out = s->st;
  out->arena_used_bytes = s->ring_used;
  out->pending_pairs = s->pending_pairs;
  out->span_ms = s->launched
      ? std::chrono::duration<double, std::milli>(s->t_last_done - s->t_first_launch).count()
      : 0.0;
  return IPUMA_OK;
}

int ipuma_server_engine(ipuma_server *s, ipuma_engine **out) {
  if (!s || !out) return set_last_error(IPUMA_EINVAL, "bad args");
  *out = s->e;
  return IPUMA_OK;
}

// Exported entry points: the bodies above, with no exception crossing the ABI.
int ipuma_server_create(const ipuma_sw_config *sw, const ipuma_algo_config *algo,
                        const ipuma_server_config *cfg, ipuma_server **out) {
  return ipuma::no_throw([&] { return ipuma_server_create_body(sw, algo, cfg, out); });
}

int ipuma_server_submit(ipuma_server *s, int32_t n, const char *a_chars, const int64_t *a_offsets,
                        const char *b_chars, const int64_t *b_offsets, int64_t *ticket) {
  return ipuma::no_throw([&] { return ipuma_server_submit_body(s, n, a_chars, a_offsets, b_chars, b_offsets, ticket); });
}

int ipuma_server_wait(ipuma_server *s, int64_t ticket, int32_t *scores, int32_t *a_range,
                      int32_t *b_range, double *latency_ms) {
  return ipuma::no_throw([&] { return ipuma_server_wait_body(s, ticket, scores, a_range, b_range, latency_ms); });
}

int ipuma_server_get_stats(ipuma_server *s, ipuma_server_stats *out) {
  return ipuma::no_throw([&] { return ipuma_server_get_stats_body(s, out); });
}

}  // extern "C"
