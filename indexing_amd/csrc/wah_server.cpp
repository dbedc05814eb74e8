// wah_server.cpp -- continuous-batching request loop over the batched build
// (wahgpu_server_* of include/wahgpu.h).
//
// The reference serves index builds through actors: the indexer slices the
// column and keeps `jobs` worker requests in flight, each a separate OpenCL
// launch set with its own transfers (src/gpu.cpp:194-261).  Here requests
// from any caller thread join one queue; each slot (a thread with its own
// wahgpu_ctx, stream, device buffers and captured graph) takes every queued
// request that fits its HBM budget and builds them as one batch.  While a
// slot copies or builds, arriving requests wait for the next free slot, so
// batches grow with the load (continuous batching) and the transfers and
// kernels of different slots overlap on the device.
#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "wahgpu.h"

namespace {

using Clock = std::chrono::steady_clock;

// Device bytes a batch needs per value: the values, keys, offsets, lengths
// and words (2 per value) outputs, the two radix ping-pong buffers of
// (key, rid, literal) items and the look-back state; and per column: its
// state block, descriptor and tile rounding.
constexpr uint64_t kBytesPerValue = 4 + 12 + 8 + 24 + 1;
constexpr uint64_t kBytesPerColumn = 12544 + 32 + 3 * 8192;
constexpr uint64_t kSlotMargin = 64ull << 20;
constexpr uint64_t kAlignValues = 64;  // columns start on 256-byte boundaries

uint64_t align_values(uint64_t n) { return (n + kAlignValues - 1) / kAlignValues * kAlignValues; }

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
  uint32_t* d_values = nullptr;
  uint32_t* d_keys = nullptr;
  uint32_t* d_offsets = nullptr;
  uint32_t* d_lengths = nullptr;
  uint32_t* d_words = nullptr;
  wahgpu_counts* d_counts = nullptr;
  std::vector<wahgpu_counts> counts;
  std::vector<hipEvent_t> done;  // one per batch column: its result copies ended
  std::thread th;
};

}  // namespace

struct wahgpu_server {
  wahgpu_server_config cfg{};
  uint64_t slot_values = 0;
  std::mutex mu;
  std::condition_variable work_cv;  // queue non-empty or stop
  std::condition_variable done_cv;  // a request completed
  std::deque<Request*> queue;
  std::unordered_map<uint64_t, std::shared_ptr<Request>> reqs;
  uint64_t next_id = 1;
  uint64_t pending = 0;
  bool stop = false;
  std::vector<std::unique_ptr<Slot>> slots;
  // statistics
  std::vector<double> lat_us;
  uint64_t batches = 0, values = 0, max_cols = 0;
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
void publish(wahgpu_server* s, Request* r, int status) {
  const Clock::time_point now = Clock::now();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    r->status = status;
    s->lat_us.push_back(std::chrono::duration<double, std::micro>(now - r->submitted).count());
    s->values += r->n;
    --s->pending;
    s->last_done = now;
  }
  s->done_cv.notify_all();
}

// One batch: copy in, batched build (graph replay), counts, then each
// request's result copies followed by its event; requests complete in
// batch order as their events pass.
void run_batch(wahgpu_server* s, Slot* sl, const std::vector<Request*>& batch) {
  hipStream_t st = (hipStream_t)wahgpu_stream(sl->ctx);
  auto fail_all = [&](size_t from, int rc) {
    (void)hipStreamSynchronize(st);  // no copy may still target a caller's buffer
    for (size_t i = from; i < batch.size(); ++i) publish(s, batch[i], rc);
  };
  std::vector<wahgpu_column> cols(batch.size());
  uint64_t off = 0;
  for (size_t i = 0; i < batch.size(); ++i) {
    Request* r = batch[i];
    cols[i] = wahgpu_column{off, r->n, r->rid_base, 0u};
    if (r->n &&
        hipMemcpyAsync(sl->d_values + off, r->values, r->n * 4, hipMemcpyHostToDevice, st) !=
            hipSuccess)
      return fail_all(0, WAHGPU_EHIP);
    off += align_values(r->n);
  }
  // values_len is the slot capacity, so every batch replays the same graph
  int rc = wahgpu_build_batch_device(sl->ctx, sl->d_values, s->slot_values, cols.data(),
                                     (uint32_t)cols.size(), sl->d_keys, sl->d_offsets,
                                     sl->d_lengths, sl->d_words, sl->d_counts);
  if (rc != WAHGPU_OK) return fail_all(0, rc);
  if (hipMemcpyAsync(sl->counts.data(), sl->d_counts, batch.size() * sizeof(wahgpu_counts),
                     hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return fail_all(0, WAHGPU_EHIP);
  std::vector<int> status(batch.size(), WAHGPU_OK);
  size_t issued = 0;
  for (; issued < batch.size(); ++issued) {
    const size_t i = issued;
    Request* r = batch[i];
    const wahgpu_counts& k = sl->counts[i];
    wahgpu_result* res = r->result;
    res->counts = k;
    bool ok = true;
    if (k.fault) {
      status[i] = WAHGPU_EHIP;
    } else if (k.keycnt > res->keys_cap || k.length > res->words_cap ||
               (k.length && !res->words) ||
               (k.keycnt && (!res->keys || !res->offsets || !res->lengths))) {
      status[i] = WAHGPU_EINVAL;
    } else {
      const uint64_t o = cols[i].offset;
      if (k.length)
        ok &= hipMemcpyAsync(res->words, sl->d_words + 2 * o, k.length * 4,
                             hipMemcpyDeviceToHost, st) == hipSuccess;
      if (k.keycnt) {
        ok &= hipMemcpyAsync(res->keys, sl->d_keys + o, k.keycnt * 4, hipMemcpyDeviceToHost,
                             st) == hipSuccess;
        ok &= hipMemcpyAsync(res->offsets, sl->d_offsets + o, k.keycnt * 4,
                             hipMemcpyDeviceToHost, st) == hipSuccess;
        ok &= hipMemcpyAsync(res->lengths, sl->d_lengths + o, k.keycnt * 4,
                             hipMemcpyDeviceToHost, st) == hipSuccess;
      }
    }
    if (!ok || hipEventRecord(sl->done[i], st) != hipSuccess) break;
  }
  for (size_t i = 0; i < issued; ++i) {
    if (hipEventSynchronize(sl->done[i]) != hipSuccess) return fail_all(i, WAHGPU_EHIP);
    publish(s, batch[i], status[i]);
  }
  if (issued < batch.size()) fail_all(issued, WAHGPU_EHIP);
}

void slot_loop(wahgpu_server* s, Slot* sl) {
  (void)hipSetDevice(s->cfg.device);
  std::vector<Request*> batch;
  for (;;) {
    batch.clear();
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->work_cv.wait(lk, [&] { return s->stop || !s->queue.empty(); });
      if (s->queue.empty()) return;  // stopping and nothing left
      uint64_t used = 0;
      while (!s->queue.empty() && batch.size() < s->cfg.max_batch_cols) {
        Request* r = s->queue.front();
        const uint64_t need = align_values(r->n);
        if (!batch.empty() && used + need > s->slot_values) break;
        batch.push_back(r);
        used += need;
        s->queue.pop_front();
        if (s->cfg.plain) break;
      }
      ++s->batches;
      s->max_cols = std::max<uint64_t>(s->max_cols, batch.size());
    }
    if (s->cfg.plain) publish(s, batch[0], run_plain(sl, batch[0]));
    else run_batch(s, sl, batch);
  }
}

void free_slot(Slot* sl) {
  if (sl->d_values) (void)hipFree(sl->d_values);
  if (sl->d_keys) (void)hipFree(sl->d_keys);
  if (sl->d_offsets) (void)hipFree(sl->d_offsets);
  if (sl->d_lengths) (void)hipFree(sl->d_lengths);
  if (sl->d_words) (void)hipFree(sl->d_words);
  if (sl->d_counts) (void)hipFree(sl->d_counts);
  for (hipEvent_t e : sl->done) (void)hipEventDestroy(e);
  sl->done.clear();
  if (sl->ctx) wahgpu_destroy(sl->ctx);
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
  const uint64_t per_slot = cfg->arena_bytes / cfg->slots;
  const uint64_t fixed = kSlotMargin + (uint64_t)s->cfg.max_batch_cols * kBytesPerColumn;
  if (per_slot <= fixed) return WAHGPU_ENOMEM;
  uint64_t v = (per_slot - fixed) / kBytesPerValue;
  if (cfg->max_batch_values && cfg->max_batch_values < v) v = cfg->max_batch_values;
  v = v / kAlignValues * kAlignValues;
  if (v > WAHGPU_MAX_VALUES) v = (WAHGPU_MAX_VALUES / kAlignValues) * kAlignValues;
  if (v == 0) return WAHGPU_ENOMEM;
  s->slot_values = v;
  if (hipSetDevice(cfg->device) != hipSuccess) return WAHGPU_ENODEV;
  for (uint32_t i = 0; i < cfg->slots; ++i) {
    auto sl = std::make_unique<Slot>();
    int rc = wahgpu_create(&sl->ctx, cfg->device, 0);
    bool ok = rc == WAHGPU_OK;
    if (ok && !s->cfg.plain) {
      ok = wahgpu_batch_reserve(sl->ctx, v, s->cfg.max_batch_cols) == WAHGPU_OK &&
           hipMalloc(&sl->d_values, v * 4) == hipSuccess &&
           hipMalloc(&sl->d_keys, v * 4) == hipSuccess &&
           hipMalloc(&sl->d_offsets, v * 4) == hipSuccess &&
           hipMalloc(&sl->d_lengths, v * 4) == hipSuccess &&
           hipMalloc(&sl->d_words, 2 * v * 4) == hipSuccess &&
           hipMalloc(&sl->d_counts, s->cfg.max_batch_cols * sizeof(wahgpu_counts)) == hipSuccess;
      sl->counts.resize(s->cfg.max_batch_cols);
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
  for (auto& sl : s->slots)
    if (sl->th.joinable()) sl->th.join();
  for (auto& sl : s->slots) free_slot(sl.get());
  delete s;
}

int wahgpu_server_submit(wahgpu_server* s, const uint32_t* values, uint64_t n, uint32_t rid_base,
                         wahgpu_result* result, uint64_t* id) {
  if (!s || !result || !id || (n && !values)) return WAHGPU_EINVAL;
  if (n > s->slot_values || n > WAHGPU_MAX_VALUES) return WAHGPU_ERANGE;
  if ((uint64_t)rid_base + n > ((uint64_t)1 << 32)) return WAHGPU_ERANGE;
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
    out->slot_values = s->slot_values;
    out->busy_s = s->have_first && !lat.empty()
                      ? std::chrono::duration<double>(s->last_done - s->first_submit).count()
                      : 0.0;
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
  s->batches = s->values = s->max_cols = 0;
  s->have_first = false;
  return WAHGPU_OK;
}

}  // extern "C"
