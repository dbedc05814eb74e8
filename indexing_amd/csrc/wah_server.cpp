// wah_server.cpp -- continuous-batching request loop over the batched build
// (wahgpu_server_* of include/wahgpu.h).
//
// The reference serves index builds through actors: the indexer slices the
// column and keeps `jobs` worker requests in flight, each a separate OpenCL
// launch set with its own transfers (src/gpu.cpp:194-261).  Here requests
// from any caller thread join one queue.  The server owns one device arena
// (allocated once) that a host-side pool allocator carves into per-batch
// blocks.  Each slot (a thread with its own wahgpu_ctx and stream) takes the
// queued requests, in arrival order, whose combined device footprint (values,
// outputs and the batched build's scratch) fits the largest free block, and
// builds them as one batch in that block; when even the first queued request
// does not fit, the slot waits until a running batch ends and frees its block.
// While slots copy or build, arriving requests queue, so batches grow with
// the load (continuous batching); the transfers of a slot overlap the other
// slots' transfers and kernels (the library runs the kernels of one build
// at a time per device).
#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "wah_internal.h"
#include "wahgpu.h"

namespace {

using Clock = std::chrono::steady_clock;

constexpr uint64_t kAlignValues = 64;  // columns start on 256-byte boundaries
constexpr uint64_t kAlignBytes = 256;

uint64_t align_values(uint64_t n) { return (n + kAlignValues - 1) / kAlignValues * kAlignValues; }
uint64_t align_bytes(uint64_t b) { return (b + kAlignBytes - 1) / kAlignBytes * kAlignBytes; }

// Device bytes of a batch's inputs and outputs: values, keys, offsets and
// lengths (V each), words (2V) and one counts record per column.
uint64_t io_bytes(uint64_t values, uint64_t cols) {
  const uint64_t v = values ? values : 1;  // an all-empty batch still carves one value
  return 4 * align_bytes(v * 4) + align_bytes(2 * v * 4) +
         align_bytes(cols * sizeof(wahgpu_counts));
}

// The whole device footprint of a batch: inputs, outputs, scratch.
uint64_t footprint(const wah_batch_acc& acc) {
  return io_bytes(acc.values, acc.cols) + wah_batch_acc_scratch(&acc);
}

// First-fit allocator over [0, cap) with coalescing of freed blocks.
struct Pool {
  uint64_t cap = 0, used = 0, peak = 0;
  std::map<uint64_t, uint64_t> free_;  // offset -> size
  void init(uint64_t c) {
    cap = c;
    free_.clear();
    free_[0] = c;
  }
  uint64_t largest() const {
    uint64_t m = 0;
    for (const auto& f : free_) m = std::max(m, f.second);
    return m;
  }
  bool alloc(uint64_t size, uint64_t* off) {
    size = align_bytes(size);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < size) continue;
      *off = it->first;
      const uint64_t rest = it->second - size, at = it->first + size;
      free_.erase(it);
      if (rest) free_[at] = rest;
      used += size;
      peak = std::max(peak, used);
      return true;
    }
    return false;
  }
  void release(uint64_t off, uint64_t size) {
    size = align_bytes(size);
    used -= size;
    auto it = free_.emplace(off, size).first;
    auto next = std::next(it);
    if (next != free_.end() && it->first + it->second == next->first) {
      it->second += next->second;
      free_.erase(next);
    }
    if (it != free_.begin()) {
      auto prev = std::prev(it);
      if (prev->first + prev->second == it->first) {
        prev->second += it->second;
        free_.erase(it);
      }
    }
  }
};

struct Request {
  uint64_t id = 0;
  const uint32_t* values = nullptr;
  uint64_t n = 0;
  uint32_t rid_base = 0;
  wahgpu_result* result = nullptr;
  Clock::time_point submitted;
  int status = WAHGPU_EBUSY;
};

struct Slot {
  wahgpu_ctx* ctx = nullptr;
  std::vector<wahgpu_counts> counts;  // host landing (pinned) for a batch's counts
  wahgpu_counts* hcounts = nullptr;
  std::vector<hipEvent_t> done;  // one per batch column: its result copies ended
  std::thread th;
};

}  // namespace

struct wahgpu_server {
  wahgpu_server_config cfg{};
  char* arena = nullptr;  // the pooled device arena (batched mode)
  Pool pool;
  std::mutex mu;
  std::condition_variable work_cv;  // queue non-empty or stop
  std::condition_variable done_cv;  // a request completed
  std::condition_variable mem_cv;   // a batch freed its block
  std::deque<Request*> queue;
  std::unordered_map<uint64_t, std::shared_ptr<Request>> reqs;
  uint64_t next_id = 1;
  uint64_t pending = 0;
  bool stop = false;
  std::vector<std::unique_ptr<Slot>> slots;
  // statistics
  std::vector<double> lat_us;
  uint64_t batches = 0, values = 0, max_cols = 0, memory_waits = 0;
  std::string last_error;  // message of the last failed request
  bool have_first = false;
  Clock::time_point first_submit, last_done;
};

namespace {

int run_plain(Slot* sl, Request* r) {
  wahgpu_result* res = r->result;
  return wahgpu_build(sl->ctx, r->values, r->n, r->rid_base, res->keys, res->offsets,
                      res->lengths, res->keys_cap, res->words, res->words_cap, &res->counts,
                      nullptr);
}

// Completes one request: status and latency under the lock, then wakes the
// waiters.  Called as each request's result copies end, so the small
// requests of a batch do not wait for the copies of its large ones.
void publish(wahgpu_server* s, Request* r, int status, const char* why = nullptr) {
  const Clock::time_point now = Clock::now();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    r->status = status;
    if (status != WAHGPU_OK)
      s->last_error = "request " + std::to_string(r->id) + ": " + (why ? why : "failed");
    s->lat_us.push_back(std::chrono::duration<double, std::micro>(now - r->submitted).count());
    s->values += r->n;
    --s->pending;
    s->last_done = now;
  }
  s->done_cv.notify_all();
}

// One batch in the arena block [off, off + bytes): copy in, batched build
// on the block's scratch, counts, then each request's result copies followed
// by its event; requests complete in batch order as their events pass; the
// block is freed once every copy out of it has ended.
void run_batch(wahgpu_server* s, Slot* sl, const std::vector<Request*>& batch, uint64_t off,
               uint64_t bytes) {
  hipStream_t st = (hipStream_t)wahgpu_stream(sl->ctx);
  auto release = [&]() {
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->pool.release(off, bytes);
    }
    s->mem_cv.notify_all();
  };
  auto fail_all = [&](size_t from, int rc) {
    (void)hipStreamSynchronize(st);  // no copy may still target a caller's buffer
    const std::string why = wahgpu_last_error(sl->ctx);
    for (size_t i = from; i < batch.size(); ++i) publish(s, batch[i], rc, why.c_str());
    release();
  };
  std::vector<wahgpu_column> cols(batch.size());
  uint64_t V = 0;
  for (size_t i = 0; i < batch.size(); ++i) {
    cols[i] = wahgpu_column{V, batch[i]->n, batch[i]->rid_base, 0u};
    V += align_values(batch[i]->n);
  }
  const uint64_t nv = V ? V : 1;
  char* p = s->arena + off;
  auto take = [&](uint64_t b) {
    char* q = p;
    p += align_bytes(b);
    return q;
  };
  uint32_t* d_values = reinterpret_cast<uint32_t*>(take(nv * 4));
  uint32_t* d_keys = reinterpret_cast<uint32_t*>(take(nv * 4));
  uint32_t* d_offsets = reinterpret_cast<uint32_t*>(take(nv * 4));
  uint32_t* d_lengths = reinterpret_cast<uint32_t*>(take(nv * 4));
  uint32_t* d_words = reinterpret_cast<uint32_t*>(take(2 * nv * 4));
  wahgpu_counts* d_counts = reinterpret_cast<wahgpu_counts*>(take(batch.size() * sizeof(wahgpu_counts)));
  const uint64_t io = (uint64_t)(p - (s->arena + off));
  for (size_t i = 0; i < batch.size(); ++i) {
    const Request* r = batch[i];
    if (r->n && hipMemcpyAsync(d_values + cols[i].offset, r->values, r->n * 4,
                               hipMemcpyHostToDevice, st) != hipSuccess)
      return fail_all(0, WAHGPU_EHIP);
  }
  int rc = wah_build_batch_ws(sl->ctx, d_values, V, cols.data(), (uint32_t)cols.size(), d_keys,
                              d_offsets, d_lengths, d_words, d_counts, p, bytes - io);
  if (rc != WAHGPU_OK) return fail_all(0, rc);
  if (hipMemcpyAsync(sl->hcounts, d_counts, batch.size() * sizeof(wahgpu_counts),
                     hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return fail_all(0, WAHGPU_EHIP);
  std::vector<int> status(batch.size(), WAHGPU_OK);
  std::string why;  // message of a column the batch could not return
  size_t issued = 0;
  for (; issued < batch.size(); ++issued) {
    const size_t i = issued;
    Request* r = batch[i];
    const wahgpu_counts& k = sl->hcounts[i];
    wahgpu_result* res = r->result;
    res->counts = k;
    bool ok = true;
    if (k.fault) {
      status[i] = WAHGPU_EHIP;
      why = "kernel reported fault " + std::to_string(k.fault) +
            " (look-back wait gave up or index out of range) in a batch of " +
            std::to_string(batch.size());
    } else if (k.keycnt > res->keys_cap || k.length > res->words_cap ||
               (k.length && !res->words) ||
               (k.keycnt && (!res->keys || !res->offsets || !res->lengths))) {
      status[i] = WAHGPU_EINVAL;
      why = "result capacity too small";
    } else {
      const uint64_t o = cols[i].offset;
      if (k.length)
        ok &= hipMemcpyAsync(res->words, d_words + 2 * o, k.length * 4, hipMemcpyDeviceToHost,
                             st) == hipSuccess;
      if (k.keycnt) {
        ok &= hipMemcpyAsync(res->keys, d_keys + o, k.keycnt * 4, hipMemcpyDeviceToHost,
                             st) == hipSuccess;
        ok &= hipMemcpyAsync(res->offsets, d_offsets + o, k.keycnt * 4,
                             hipMemcpyDeviceToHost, st) == hipSuccess;
        ok &= hipMemcpyAsync(res->lengths, d_lengths + o, k.keycnt * 4,
                             hipMemcpyDeviceToHost, st) == hipSuccess;
      }
    }
    if (!ok || hipEventRecord(sl->done[i], st) != hipSuccess) break;
  }
  for (size_t i = 0; i < issued; ++i) {
    if (hipEventSynchronize(sl->done[i]) != hipSuccess) return fail_all(i, WAHGPU_EHIP);
    publish(s, batch[i], status[i], status[i] ? why.c_str() : nullptr);
  }
  if (issued < batch.size()) return fail_all(issued, WAHGPU_EHIP);
  release();
}

void slot_loop(wahgpu_server* s, Slot* sl) {
  (void)hipSetDevice(s->cfg.device);
  std::vector<Request*> batch;
  for (;;) {
    batch.clear();
    uint64_t off = 0, bytes = 0;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      for (;;) {
        s->work_cv.wait(lk, [&] { return s->stop || !s->queue.empty(); });
        if (s->queue.empty()) return;  // stopping and nothing left
        if (s->cfg.plain) {
          batch.push_back(s->queue.front());
          s->queue.pop_front();
          break;
        }
        // the queued requests, in order, whose footprint fits the largest
        // free block (and the batch limits)
        const uint64_t avail = s->pool.largest();
        wah_batch_acc acc{};
        uint64_t fp = 0;
        size_t take = 0;
        for (size_t i = 0; i < s->queue.size() && take < s->cfg.max_batch_cols; ++i) {
          wah_batch_acc next = acc;
          wah_batch_acc_add(&next, s->queue[i]->n);
          if (take && s->cfg.max_batch_values && next.values > s->cfg.max_batch_values) break;
          const uint64_t f = footprint(next);
          if (f > avail) break;
          acc = next;
          fp = f;
          ++take;
        }
        if (take == 0) {
          // the first queued request needs more than is free: wait for a
          // running batch to end (submit refused requests the empty arena
          // cannot hold, so this wait ends)
          ++s->memory_waits;
          s->mem_cv.wait(lk);
          continue;
        }
        if (!s->pool.alloc(fp, &off)) continue;  // cannot happen: fp <= largest
        bytes = fp;
        for (size_t i = 0; i < take; ++i) {
          batch.push_back(s->queue.front());
          s->queue.pop_front();
        }
        break;
      }
      ++s->batches;
      s->max_cols = std::max<uint64_t>(s->max_cols, batch.size());
    }
    if (s->cfg.plain) {
      const int rc = run_plain(sl, batch[0]);
      publish(s, batch[0], rc, rc ? wahgpu_last_error(sl->ctx) : nullptr);
    }
    else run_batch(s, sl, batch, off, bytes);
  }
}

void free_slot(Slot* sl) {
  for (hipEvent_t e : sl->done) (void)hipEventDestroy(e);
  sl->done.clear();
  if (sl->hcounts) (void)hipHostFree(sl->hcounts);
  sl->hcounts = nullptr;
  if (sl->ctx) wahgpu_destroy(sl->ctx);
  sl->ctx = nullptr;
}

// Footprint of a batch of one column of n values.
uint64_t single_footprint(uint64_t n) {
  wah_batch_acc acc{};
  wah_batch_acc_add(&acc, n);
  return footprint(acc);
}

}  // namespace

extern "C" {

int wahgpu_server_create(wahgpu_server** out, const wahgpu_server_config* cfg) {
  if (!out || !cfg || cfg->slots == 0 || cfg->max_batch_cols == 0 || cfg->arena_bytes == 0)
    return WAHGPU_EINVAL;
  *out = nullptr;
  auto s = std::make_unique<wahgpu_server>();
  s->cfg = *cfg;
  if (s->cfg.plain) s->cfg.max_batch_cols = 1;
  if (hipSetDevice(cfg->device) != hipSuccess) return WAHGPU_ENODEV;
  if (!s->cfg.plain) {
    if (hipMalloc(&s->arena, cfg->arena_bytes) != hipSuccess) {
      (void)hipGetLastError();
      return WAHGPU_ENOMEM;
    }
    s->pool.init(cfg->arena_bytes);
  }
  for (uint32_t i = 0; i < cfg->slots; ++i) {
    auto sl = std::make_unique<Slot>();
    int rc = wahgpu_create(&sl->ctx, cfg->device, 0);
    bool ok = rc == WAHGPU_OK;
    if (ok && !s->cfg.plain) {
      ok = hipHostMalloc(reinterpret_cast<void**>(&sl->hcounts),
                         s->cfg.max_batch_cols * sizeof(wahgpu_counts),
                         hipHostMallocDefault) == hipSuccess;
      for (uint32_t e = 0; ok && e < s->cfg.max_batch_cols; ++e) {
        hipEvent_t ev = nullptr;
        ok = hipEventCreateWithFlags(&ev, hipEventDisableTiming) == hipSuccess;
        if (ok) sl->done.push_back(ev);
      }
    }
    if (!ok) {
      (void)hipGetLastError();
      free_slot(sl.get());
      for (auto& o : s->slots) free_slot(o.get());
      if (s->arena) (void)hipFree(s->arena);
      return rc != WAHGPU_OK ? rc : WAHGPU_ENOMEM;
    }
    s->slots.push_back(std::move(sl));
  }
  wahgpu_server* raw = s.release();
  for (auto& sl : raw->slots) sl->th = std::thread(slot_loop, raw, sl.get());
  *out = raw;
  return WAHGPU_OK;
}

void wahgpu_server_destroy(wahgpu_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->work_cv.notify_all();
  s->mem_cv.notify_all();
  for (auto& sl : s->slots)
    if (sl->th.joinable()) sl->th.join();
  for (auto& sl : s->slots) free_slot(sl.get());
  if (s->arena) {
    (void)hipSetDevice(s->cfg.device);
    (void)hipFree(s->arena);
  }
  delete s;
}

int wahgpu_server_submit(wahgpu_server* s, const uint32_t* values, uint64_t n, uint32_t rid_base,
                         wahgpu_result* result, uint64_t* id) {
  if (!s || !result || !id || (n && !values)) return WAHGPU_EINVAL;
  if (n > WAHGPU_MAX_VALUES) return WAHGPU_ERANGE;
  if ((uint64_t)rid_base + n > ((uint64_t)1 << 32)) return WAHGPU_ERANGE;
  if (!s->cfg.plain && (single_footprint(n) > s->cfg.arena_bytes ||
                        (s->cfg.max_batch_values && align_values(n) > s->cfg.max_batch_values)))
    return WAHGPU_ERANGE;  // the empty arena could not hold it
  auto r = std::make_shared<Request>();
  r->values = values;
  r->n = n;
  r->rid_base = rid_base;
  r->result = result;
  r->submitted = Clock::now();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return WAHGPU_EINVAL;
    r->id = s->next_id++;
    *id = r->id;
    if (!s->have_first) {
      s->have_first = true;
      s->first_submit = r->submitted;
    }
    s->queue.push_back(r.get());
    ++s->pending;
    s->reqs.emplace(r->id, std::move(r));
  }
  s->work_cv.notify_one();
  return WAHGPU_OK;
}

int wahgpu_server_poll(wahgpu_server* s, uint64_t id) {
  if (!s) return WAHGPU_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return WAHGPU_EINVAL;
  const int st = it->second->status;
  if (st != WAHGPU_EBUSY) s->reqs.erase(it);
  return st;
}

int wahgpu_server_wait(wahgpu_server* s, uint64_t id) {
  if (!s) return WAHGPU_EINVAL;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return WAHGPU_EINVAL;
  // Another caller's poll or wait on the same id may complete and forget the
  // request while this one sleeps: the shared_ptr keeps it alive, and only
  // the entry still in the table is erased.
  std::shared_ptr<Request> r = it->second;
  s->done_cv.wait(lk, [&] { return r->status != WAHGPU_EBUSY; });
  const int st = r->status;
  auto again = s->reqs.find(id);
  if (again != s->reqs.end() && again->second == r) s->reqs.erase(again);
  return st;
}

const char* wahgpu_server_last_error(wahgpu_server* s) {
  if (!s) return "";
  std::lock_guard<std::mutex> lk(s->mu);
  thread_local std::string copy;
  copy = s->last_error;
  return copy.c_str();
}

int wahgpu_server_drain(wahgpu_server* s) {
  if (!s) return WAHGPU_EINVAL;
  std::unique_lock<std::mutex> lk(s->mu);
  s->done_cv.wait(lk, [&] { return s->pending == 0; });
  return WAHGPU_OK;
}

int wahgpu_server_get_stats(wahgpu_server* s, wahgpu_server_stats* out) {
  if (!s || !out) return WAHGPU_EINVAL;
  std::vector<double> lat;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    lat = s->lat_us;
    out->requests = lat.size();
    out->batches = s->batches;
    out->values = s->values;
    out->max_batch_cols = s->max_cols;
    out->slot_values = s->cfg.max_batch_values;
    out->busy_s = s->have_first && !lat.empty()
                      ? std::chrono::duration<double>(s->last_done - s->first_submit).count()
                      : 0.0;
    out->memory_waits = s->memory_waits;
    out->arena_bytes = s->pool.cap;
    out->arena_peak = s->pool.peak;
  }
  std::sort(lat.begin(), lat.end());
  auto pct = [&](double q) {
    if (lat.empty()) return 0.0;
    size_t i = (size_t)(q * (double)(lat.size() - 1) + 0.5);
    return lat[std::min(i, lat.size() - 1)];
  };
  out->p50_us = pct(0.50);
  out->p99_us = pct(0.99);
  double sum = 0;
  for (double x : lat) sum += x;
  out->mean_us = lat.empty() ? 0.0 : sum / (double)lat.size();
  out->max_us = lat.empty() ? 0.0 : lat.back();
  return WAHGPU_OK;
}

int wahgpu_host_alloc(uint64_t bytes, void** out) {
  if (!out) return WAHGPU_EINVAL;
  *out = nullptr;
  if (hipHostMalloc(out, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    *out = nullptr;
    return WAHGPU_ENOMEM;
  }
  return WAHGPU_OK;
}

void wahgpu_host_free(void* p) {
  if (p) (void)hipHostFree(p);
}

int wahgpu_server_reset_stats(wahgpu_server* s) {
  if (!s) return WAHGPU_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  s->lat_us.clear();
  s->batches = s->values = s->max_cols = s->memory_waits = 0;
  s->pool.peak = s->pool.used;
  s->have_first = false;
  return WAHGPU_OK;
}

}  // extern "C"
