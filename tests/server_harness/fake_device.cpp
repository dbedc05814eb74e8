// fake_device.cpp -- the CPU stand-ins the host-side server harness links
// wah_server.cpp against: the HIP calls of hip/hip_runtime_api.h (streams
// are worker threads, events count completions) and the libwahgpu entry
// points the server uses (contexts, the plain build, the batch footprint
// accounting and the batched build on a caller workspace).
//
// The fake build does not produce WAH: per column it sorts the values in the
// workspace and writes, for each distinct value v with c occurrences,
// key v, offset j, length 1 and word fake_word(c, rid_base).  That is enough
// to tell which request's values and rid_base reached which result buffers,
// and every byte of the batch's block is written from the stream thread, so
// two batches handed overlapping blocks race where the sanitizers see it.
#include "fake_device.h"

#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "hip/hip_runtime_api.h"
#include "wah_internal.h"
#include "wahgpu.h"

// ---- streams and events ----------------------------------------------------

struct FakeStream {
  std::mutex mu;
  std::condition_variable cv;
  std::deque<std::function<void()>> ops;
  uint64_t queued = 0, finished = 0;
  bool stop = false;
  std::thread th;
};

struct FakeEvent {
  std::mutex mu;
  std::condition_variable cv;
  uint64_t recorded = 0, completed = 0;
};

namespace {

std::atomic<size_t> g_device_bytes{0};
std::mutex g_alloc_mu;
std::vector<std::pair<void*, size_t>> g_allocs;

void stream_main(FakeStream* s) {
  std::minstd_rand rng((uint32_t)(uintptr_t)s);
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    s->cv.wait(lk, [&] { return s->stop || !s->ops.empty(); });
    if (s->ops.empty()) return;
    std::function<void()> op = std::move(s->ops.front());
    s->ops.pop_front();
    lk.unlock();
    if (rng() % 4 == 0) usleep(rng() % 200);  // vary the interleavings
    op();
    lk.lock();
    ++s->finished;
    s->cv.notify_all();
  }
}

}  // namespace

hipStream_t fake_stream_create() {
  auto* s = new FakeStream;
  s->th = std::thread(stream_main, s);
  return s;
}

void fake_stream_destroy(hipStream_t s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv.notify_all();
  s->th.join();
  delete s;
}

void fake_enqueue(hipStream_t s, std::function<void()> op) {
  std::lock_guard<std::mutex> lk(s->mu);
  s->ops.push_back(std::move(op));
  ++s->queued;
  s->cv.notify_all();
}

size_t fake_device_bytes() { return g_device_bytes.load(); }

hipError_t hipSetDevice(int device) { return device == 0 ? hipSuccess : hipErrorInvalidDevice; }
hipError_t hipGetLastError(void) { return hipSuccess; }

hipError_t hipMalloc(void** p, size_t bytes) {
  if (!p) return hipErrorInvalidValue;
  *p = aligned_alloc(256, (bytes + 255) / 256 * 256);
  if (!*p) return hipErrorOutOfMemory;
  std::lock_guard<std::mutex> lk(g_alloc_mu);
  g_allocs.emplace_back(*p, bytes);
  g_device_bytes += bytes;
  return hipSuccess;
}

hipError_t hipFree(void* p) {
  if (!p) return hipSuccess;
  std::lock_guard<std::mutex> lk(g_alloc_mu);
  for (size_t i = 0; i < g_allocs.size(); ++i)
    if (g_allocs[i].first == p) {
      g_device_bytes -= g_allocs[i].second;
      g_allocs.erase(g_allocs.begin() + (long)i);
      free(p);
      return hipSuccess;
    }
  return hipErrorInvalidValue;
}

hipError_t hipHostMalloc(void** p, size_t bytes, unsigned) {
  if (!p) return hipErrorInvalidValue;
  *p = malloc(bytes ? bytes : 1);
  return *p ? hipSuccess : hipErrorOutOfMemory;
}

hipError_t hipHostFree(void* p) {
  free(p);
  return hipSuccess;
}

hipError_t hipMemcpyAsync(void* dst, const void* src, size_t bytes, hipMemcpyKind,
                          hipStream_t stream) {
  if (!stream || (bytes && (!dst || !src))) return hipErrorInvalidValue;
  fake_enqueue(stream, [=] { memcpy(dst, src, bytes); });
  return hipSuccess;
}

hipError_t hipStreamSynchronize(hipStream_t s) {
  if (!s) return hipErrorInvalidValue;
  std::unique_lock<std::mutex> lk(s->mu);
  const uint64_t target = s->queued;
  s->cv.wait(lk, [&] { return s->finished >= target; });
  return hipSuccess;
}

hipError_t hipEventCreateWithFlags(hipEvent_t* ev, unsigned) {
  if (!ev) return hipErrorInvalidValue;
  *ev = new FakeEvent;
  return hipSuccess;
}

hipError_t hipEventDestroy(hipEvent_t ev) {
  delete ev;
  return hipSuccess;
}

hipError_t hipEventRecord(hipEvent_t ev, hipStream_t stream) {
  if (!ev || !stream) return hipErrorInvalidValue;
  uint64_t seq;
  {
    std::lock_guard<std::mutex> lk(ev->mu);
    seq = ++ev->recorded;
  }
  fake_enqueue(stream, [ev, seq] {
    // notified under the lock: a waiter that returns may destroy the event
    std::lock_guard<std::mutex> lk(ev->mu);
    ev->completed = std::max(ev->completed, seq);
    ev->cv.notify_all();
  });
  return hipSuccess;
}

hipError_t hipEventSynchronize(hipEvent_t ev) {
  if (!ev) return hipErrorInvalidValue;
  std::unique_lock<std::mutex> lk(ev->mu);
  const uint64_t target = ev->recorded;
  ev->cv.wait(lk, [&] { return ev->completed >= target; });
  return hipSuccess;
}

// ---- the fake index ----------------------------------------------------------

uint32_t fake_word(uint64_t count, uint32_t rid_base) {
  return (uint32_t)(count * 2654435761ull) ^ rid_base;
}

// Fake index of one column from the values in `sorted` (sorted in place).
static void fake_index(uint32_t* sorted, uint64_t n, uint32_t rid_base, uint32_t* keys,
                       uint32_t* offsets, uint32_t* lengths, uint32_t* words,
                       wahgpu_counts* counts) {
  std::sort(sorted, sorted + n);
  uint64_t k = 0;
  for (uint64_t i = 0; i < n;) {
    uint64_t j = i;
    while (j < n && sorted[j] == sorted[i]) ++j;
    keys[k] = sorted[i];
    offsets[k] = (uint32_t)k;
    lengths[k] = 1;
    words[k] = fake_word(j - i, rid_base);
    ++k;
    i = j;
  }
  wahgpu_counts c{};
  c.k = k;
  c.length = k;
  c.keycnt = k;
  c.bound = n ? sorted[n - 1] : 0;
  c.passes = 1;
  *counts = c;  // callers raise fault for a column that starts with the marker
}

uint64_t fake_distinct(const uint32_t* values, uint64_t n) {
  std::vector<uint32_t> v(values, values + n);
  std::sort(v.begin(), v.end());
  return (uint64_t)(std::unique(v.begin(), v.end()) - v.begin());
}

void fake_expected(const uint32_t* values, uint64_t n, uint32_t rid_base,
                   std::vector<uint32_t>* keys, std::vector<uint32_t>* words) {
  std::vector<uint32_t> v(values, values + n);
  const uint64_t k = fake_distinct(values, n);
  keys->assign(k, 0);
  words->assign(k, 0);
  std::vector<uint32_t> off(k), len(k);
  wahgpu_counts c;
  fake_index(v.data(), n, rid_base, keys->data(), off.data(), len.data(), words->data(), &c);
}

// ---- contexts -----------------------------------------------------------------

struct wahgpu_ctx {
  hipStream_t stream = nullptr;
  std::string err;
};

extern "C" {

int wahgpu_create(wahgpu_ctx** out, int device_ordinal, uint64_t) {
  if (!out) return WAHGPU_EINVAL;
  if (device_ordinal != 0) return WAHGPU_ENODEV;
  auto* c = new wahgpu_ctx;
  c->stream = fake_stream_create();
  *out = c;
  return WAHGPU_OK;
}

void wahgpu_destroy(wahgpu_ctx* c) {
  if (!c) return;
  fake_stream_destroy(c->stream);
  delete c;
}

const char* wahgpu_last_error(const wahgpu_ctx* c) { return c ? c->err.c_str() : ""; }

void* wahgpu_stream(wahgpu_ctx* c) { return c ? c->stream : nullptr; }

int wahgpu_build(wahgpu_ctx* c, const uint32_t* values, uint64_t n, uint32_t rid_base,
                 uint32_t* keys, uint32_t* offsets, uint32_t* lengths, uint64_t keys_cap,
                 uint32_t* words, uint64_t words_cap, wahgpu_counts* counts,
                 wahgpu_phase_times*) {
  if (!c || !counts || (n && !values)) return WAHGPU_EINVAL;
  const uint64_t k = fake_distinct(values, n);
  if (k > keys_cap || k > words_cap) {
    c->err = "result capacity too small";
    return WAHGPU_EINVAL;
  }
  std::vector<uint32_t> tmp(values, values + n);
  fake_enqueue(c->stream, [&] {
    fake_index(tmp.data(), n, rid_base, keys, offsets, lengths, words, counts);
    if (n && values[0] == kFakeFaultMarker) counts->fault = 1;
  });
  (void)hipStreamSynchronize(c->stream);
  if (counts->fault) {
    c->err = "kernel reported fault";
    return WAHGPU_EHIP;
  }
  return WAHGPU_OK;
}

void wah_batch_acc_add(wah_batch_acc* acc, uint64_t n) {
  acc->values += (n + 63) / 64 * 64;
  acc->cols += 1;
  acc->sort_tiles += (n + 4095) / 4096;
  acc->tail_tiles += (n + 2047) / 2048;
  acc->hist_chunks += (n + 65535) / 65536;
}

uint64_t wah_batch_acc_scratch(const wah_batch_acc* acc) {
  return 4 * acc->values + 256ull * acc->cols + 4096;
}

int wah_build_batch_ws(wahgpu_ctx* c, const uint32_t* d_values, uint64_t values_len,
                       const wahgpu_column* cols, uint32_t ncols, uint32_t* d_keys,
                       uint32_t* d_offsets, uint32_t* d_lengths, uint32_t* d_words,
                       wahgpu_counts* d_counts, void* d_scratch, uint64_t scratch_bytes) {
  if (!c) return WAHGPU_EINVAL;
  if (ncols == 0) return WAHGPU_OK;
  if (!cols || !d_counts || !d_scratch) {
    c->err = "null argument";
    return WAHGPU_EINVAL;
  }
  wah_batch_acc acc{};
  for (uint32_t i = 0; i < ncols; ++i) {
    if (cols[i].offset + cols[i].n > values_len) {
      c->err = "column past values_len";
      return WAHGPU_EINVAL;
    }
    wah_batch_acc_add(&acc, cols[i].n);
  }
  const uint64_t need = wah_batch_acc_scratch(&acc);
  if (need > scratch_bytes) {
    c->err = "workspace too small";
    return WAHGPU_ENOMEM;
  }
  std::vector<wahgpu_column> cv(cols, cols + ncols);
  char* ws = static_cast<char*>(d_scratch);
  fake_enqueue(c->stream, [=] {
    // the whole workspace is written, as a build's scratch would be
    memset(ws, 0xA5, need);
    uint32_t* sorted = reinterpret_cast<uint32_t*>(ws);
    for (uint32_t i = 0; i < ncols; ++i) {
      const wahgpu_column& col = cv[i];
      const uint64_t o = col.offset;
      memcpy(sorted + o, d_values + o, col.n * 4);
      fake_index(sorted + o, col.n, col.rid_base, d_keys + o, d_offsets + o, d_lengths + o,
                 d_words + 2 * o, &d_counts[i]);
      if (col.n && d_values[o] == kFakeFaultMarker) d_counts[i].fault = 1;
    }
  });
  return WAHGPU_OK;
}

}  // extern "C"
