// wahgpu_api.hip -- the C ABI of libwahgpu (include/wahgpu.h): contexts with a
// stream and a pooled device arena, the fused and staged builds, the stage
// entry points and the scan / compaction primitives.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "wah_internal.h"
#include "wah_kernels.h"
#include "wahgpu.h"

using wah::BuildArgs;
using wah::ScanArgs;

struct wahgpu_ctx {
  int device = 0;
  hipStream_t stream = nullptr;      // stream all work is ordered on
  hipStream_t own_stream = nullptr;  // created with the context
  hipEvent_t switch_ev = nullptr;    // orders a new stream after the old one
  char* arena = nullptr;
  uint64_t arena_cap = 0;
  bool fixed = false;
  uint64_t bump = 0;
  uint32_t max_resident = 1024;
  std::string err;
  wahgpu_counts* pinned_counts = nullptr;  // host-pinned landing slot
  // per-kernel timing (wahgpu_timing_begin/end)
  std::vector<hipEvent_t> tev;  // max_calls x (WAHGPU_TIMING_GROUPS + 1)
  uint32_t tmax = 0, tcalls = 0;
  bool timing = false;
  // batched builds: pinned staging of the descriptor, reusable once its copy ran
  char* desc_stage = nullptr;
  uint64_t desc_stage_cap = 0;
  hipEvent_t desc_copied = nullptr;
  // graph replay (wahgpu_batch_reserve)
  uint64_t res_values = 0;
  uint32_t res_cols = 0;
  hipGraphExec_t gexec = nullptr;
  std::vector<uintptr_t> gkey;  // buffers and sizes the graph was captured with
  uint64_t captures = 0, replays = 0;
  bool atomic_rank = false;  // radix ranking by LDS fetch-add (self-tested)
  // radix tile shape per column-length bucket (wahgpu_autotune / set_tuning)
  uint32_t tune_shape[WAHGPU_TUNE_BUCKETS] = {0, 0, 0, 0, 0};
  double tune_ms[WAHGPU_TUNE_BUCKETS][WAHGPU_TUNE_SHAPES] = {};
  int force_shape = -1;  // autotune runs: every column in this shape
  // radix digit width per bucket: 8, or 11 = wide digits allowed (when they
  // take fewer passes)
  uint32_t tune_digits[WAHGPU_TUNE_BUCKETS] = {8, 8, 8, 8, 8};
  double tune_digit_ms[WAHGPU_TUNE_BUCKETS][2] = {};
  int force_digits = 0;  // autotune runs: 8 or 11 for every column
  char* noalloc = nullptr;  // this device's copy of the static scratch array
};

namespace {

int fail(wahgpu_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

#define HIPCHK(ctx, call)                                                             \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess)                                                             \
      return fail(ctx, WAHGPU_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define LAUNCHCHK(ctx)                                                                 \
  do {                                                                                 \
    hipError_t e_ = hipGetLastError();                                                 \
    if (e_ != hipSuccess)                                                              \
      return fail(ctx, WAHGPU_EHIP, std::string("kernel launch: ") + hipGetErrorString(e_)); \
  } while (0)

// ---- one look-back launch sequence at a time per device -----------------
// The single-pass kernels (radix passes, tail, merge, compaction, scans,
// decode) wait for tiles that other workgroups of the same launch draw, and
// a tile-counter lane is drained only by workgroups placed on its XCD.  Two
// such launches on different streams can each fill the XCD the other one
// still needs while waiting on their own undrained lane: a circular wait that
// only the bounded wait ends, by reporting a fault.  So on one device these
// kernels never overlap.  Each launch sequence, from any context and on any
// stream, waits for a per-device event that the previous sequence recorded
// and records it after itself; the host lock keeps wait, launches and record
// of one sequence together (host waits come after end()).  Copies and other
// kernels still overlap them.
// Sequences enqueued while their stream is being captured into a graph are
// not ordered: a caller replays such graphs on one stream at a time.
constexpr int kOrderDevices = 64;
struct OrderSlot {
  std::recursive_mutex mu;
  hipEvent_t last = nullptr;  // recorded after the latest sequence
};
OrderSlot g_order[kOrderDevices];

class DeviceOrder {
 public:
  explicit DeviceOrder(wahgpu_ctx* c) : c_(c) {
    if (c->device < 0 || c->device >= kOrderDevices) {
      st_ = hipErrorInvalidDevice;
      return;
    }
    slot_ = &g_order[c->device];
    slot_->mu.lock();
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    st_ = hipStreamIsCapturing(c->stream, &cs);
    if (st_ != hipSuccess || cs != hipStreamCaptureStatusNone) return;
    if (!slot_->last) st_ = hipEventCreateWithFlags(&slot_->last, hipEventDisableTiming);
    if (st_ == hipSuccess) st_ = hipStreamWaitEvent(c->stream, slot_->last, 0);
    record_ = st_ == hipSuccess;
  }
  ~DeviceOrder() { end(); }
  // Records the event after the sequence and releases the lock; called
  // before the caller synchronizes its stream, so no host wait ever holds
  // the lock (a stream may wait for work that another thread enqueues).
  void end() {
    if (!slot_) return;
    if (record_) (void)hipEventRecord(slot_->last, c_->stream);
    slot_->mu.unlock();
    slot_ = nullptr;
  }
  DeviceOrder(const DeviceOrder&) = delete;
  DeviceOrder& operator=(const DeviceOrder&) = delete;
  hipError_t status() const { return st_; }

 private:
  wahgpu_ctx* c_;
  OrderSlot* slot_ = nullptr;
  hipError_t st_ = hipSuccess;
  bool record_ = false;
};

#define ORDERED(ctx)                  \
  DeviceOrder order_guard_(ctx);      \
  HIPCHK(ctx, order_guard_.status())

constexpr uint64_t kAlign = 256;
uint64_t align_up(uint64_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

// Reserve `bytes` of arena for the current call (grows unless fixed).
int arena_reserve(wahgpu_ctx* c, uint64_t bytes) {
  c->bump = 0;
  if (bytes <= c->arena_cap) return WAHGPU_OK;
  if (c->fixed)
    return fail(c, WAHGPU_ENOMEM, "arena of " + std::to_string(c->arena_cap) +
                                      " bytes is too small for " + std::to_string(bytes));
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (c->arena) {
    (void)hipFree(c->arena);
    c->arena = nullptr;
    c->arena_cap = 0;
  }
  uint64_t cap = (bytes + (2u << 20) - 1) / (2u << 20) * (2u << 20);
  if (hipMalloc(&c->arena, cap) != hipSuccess) {
    (void)hipGetLastError();
    c->arena = nullptr;
    return fail(c, WAHGPU_ENOMEM, "hipMalloc of " + std::to_string(cap) + " bytes failed");
  }
  c->arena_cap = cap;
  return WAHGPU_OK;
}

template <class T>
T* take(wahgpu_ctx* c, uint64_t count) {
  T* p = reinterpret_cast<T*>(c->arena + c->bump);
  c->bump += align_up(count * sizeof(T));
  return p;
}

// Sizes a fused build's scratch and grids are laid out for.
struct Caps {
  uint64_t n;  // values of the concatenated input (its highest end)
  uint32_t cols, sort_tiles, tail_tiles, hist_chunks;
  uint32_t wide_slots = 0, wide_tiles = 0;  // columns that may take wide digits, their tiles
  bool operator==(const Caps& o) const {
    return n == o.n && cols == o.cols && sort_tiles == o.sort_tiles &&
           tail_tiles == o.tail_tiles && hist_chunks == o.hist_chunks &&
           wide_slots == o.wide_slots && wide_tiles == o.wide_tiles;
  }
  bool covers(const Caps& o) const {
    return n >= o.n && cols >= o.cols && sort_tiles >= o.sort_tiles &&
           tail_tiles >= o.tail_tiles && hist_chunks >= o.hist_chunks &&
           wide_slots >= o.wide_slots && wide_tiles >= o.wide_tiles;
  }
};
// Column-length bucket of n: < 2^14, < 2^17, < 2^20, the rest.
constexpr uint32_t kBucketLog2[WAHGPU_TUNE_BUCKETS] = {14, 17, 20, 24, 64};
uint32_t bucket_of(uint64_t n) {
  for (uint32_t b = 0; b + 1 < WAHGPU_TUNE_BUCKETS; ++b)
    if (n < (1ull << kBucketLog2[b])) return b;
  return WAHGPU_TUNE_BUCKETS - 1;
}
int shape_for(const wahgpu_ctx* c, uint64_t n) {
  return c->force_shape >= 0 ? c->force_shape : (int)c->tune_shape[bucket_of(n)];
}

// A column of n values may take wide digits (k_plan decides from its values).
bool wide_ok(const wahgpu_ctx* c, uint64_t n) {
  const uint32_t d = c->force_digits ? (uint32_t)c->force_digits : c->tune_digits[bucket_of(n)];
  return d == 11u && n >= wah::kWideMinValues;
}
uint32_t wide_tiles_of(uint64_t n) { return (uint32_t)((n + wah::kWideTile - 1) / wah::kWideTile); }

Caps caps_one(uint64_t n, int shape, bool wide = false) {
  Caps k{n, 1u, wah::sort_tiles_of(n, wah::radix_tile(shape)), wah::tail_tiles_of(n),
         wah::hist_chunks_of(n)};
  if (wide) {
    k.wide_slots = 1;
    k.wide_tiles = wide_tiles_of(n);
  }
  return k;
}
// The fused build's caps for one column under the context's tuning.
Caps caps_for(const wahgpu_ctx* c, uint64_t n) {
  return caps_one(n, shape_for(c, n), wide_ok(c, n));
}

// Bytes of scratch a fused build needs (besides outputs), with `ndesc`
// descriptor regions of k.cols columns each.
uint64_t build_scratch_bytes(const Caps& k, uint32_t ndesc = 1) {
  const uint64_t nn = k.n ? k.n : 1;
  uint64_t b = 0;
  b += 6 * align_up(nn * 4);                       // keys, rids and lits of A and B
  b += ndesc * align_up(sizeof(wah::DescHead) + (uint64_t)k.cols * sizeof(wah::Column));
  b += align_up(sizeof(wah::Header));
  b += align_up((uint64_t)k.cols * sizeof(wah::ColState));
  b += align_up(2ull * k.sort_tiles * 256 * 4);
  b += align_up((uint64_t)k.tail_tiles * 8 + 8);
  b += align_up((uint64_t)k.tail_tiles * 4 + 4);
  b += align_up((uint64_t)k.sort_tiles * 4 + 4);
  b += align_up((uint64_t)k.wide_slots * sizeof(wah::WideState));
  b += align_up(2ull * k.wide_tiles * wah::kWideBins * 4);
  return b;
}

// Carves the shared scratch and the first descriptor region; further regions
// (one per sub-batch) follow with carve_desc.
void carve_build(wahgpu_ctx* c, BuildArgs* a, const Caps& k, int shape) {
  const uint64_t nn = k.n ? k.n : 1;
  a->keysA = take<uint32_t>(c, nn);
  a->ridsA = take<uint32_t>(c, nn);
  a->keysB = take<uint32_t>(c, nn);
  a->ridsB = take<uint32_t>(c, nn);
  a->litsA = take<uint32_t>(c, nn);
  a->litsB = take<uint32_t>(c, nn);
  a->desc = reinterpret_cast<wah::DescHead*>(
      take<char>(c, sizeof(wah::DescHead) + (uint64_t)k.cols * sizeof(wah::Column)));
  a->hdr = take<wah::Header>(c, 1);
  a->cs = take<wah::ColState>(c, k.cols);
  a->sort_status_stride = (uint64_t)k.sort_tiles * 256;
  a->sort_status = take<uint32_t>(c, 2 * a->sort_status_stride);
  a->tail_status = take<uint64_t>(c, (uint64_t)k.tail_tiles + 1);
  a->tile_entries = take<uint32_t>(c, (uint64_t)k.tail_tiles + 1);
  a->tile_runs = take<uint32_t>(c, (uint64_t)k.sort_tiles + 1);
  a->wide = take<wah::WideState>(c, k.wide_slots);
  a->wide_status_stride = (uint64_t)k.wide_tiles * wah::kWideBins;
  a->wide_status = take<uint32_t>(c, 2 * a->wide_status_stride);
  a->wide_slots_cap = k.wide_slots;
  a->wide_tiles_cap = k.wide_tiles;
  a->max_resident = c->max_resident;
  a->atomic_rank = c->atomic_rank;
  a->shape = shape;
  a->n_cap = k.n;
  a->cols_cap = k.cols;
  a->sort_tiles_cap = k.sort_tiles;
  a->tail_tiles_cap = k.tail_tiles;
  a->hist_chunks_cap = k.hist_chunks;
}
wah::DescHead* carve_desc(wahgpu_ctx* c, uint32_t cols) {
  return reinterpret_cast<wah::DescHead*>(
      take<char>(c, sizeof(wah::DescHead) + (uint64_t)cols * sizeof(wah::Column)));
}

// A single column at offset 0: its descriptor travels in k_begin's arguments.
// wide: it may take wide digits (slot 0, tiles from 0).  A long column takes
// the long tail tiles (fewer than the caps' kTailTile count).
void set_one(BuildArgs* a, uint64_t n, uint32_t rid_base, bool wide = false) {
  a->inline_desc = true;
  a->long_tail = n >= wah::kLongTailMin;
  a->one_dims = wah::Dims{1u, wah::sort_tiles_of(n, wah::radix_tile(a->shape)),
                          wah::tail_tiles_of(n, a->long_tail ? wah::kLongTailTile : wah::kTailTile),
                          wah::hist_chunks_of(n), wide ? wide_tiles_of(n) : 0u, 0u};
  a->one = wah::Column{0, n, rid_base, 0u, 0u, 0u, 0u, wide ? 1u : 0u, 0u, 0u};
}

// Scan scratch, zeroed before use: tile-counter lanes (kLaneWords words),
// then [fault u32, pad u32, result u64], then the u64 status words.
struct ScanScratch {
  uint32_t* base;
  uint64_t words;
  ScanArgs args;
  uint32_t* header;  // fault, pad, result lo, result hi
  uint64_t* result;
};
constexpr uint64_t kScanHeadWords = wah::kLaneWords + 8;
uint64_t scan_scratch_bytes(uint64_t n) {
  return align_up(kScanHeadWords * 4 + wah::scan_status_words(n) * 8);
}
ScanScratch carve_scan(wahgpu_ctx* c, uint64_t n) {
  ScanScratch s;
  uint64_t bytes = kScanHeadWords * 4 + wah::scan_status_words(n) * 8;
  s.base = take<uint32_t>(c, bytes / 4);
  s.words = bytes / 4;
  s.args.tile_counter = s.base;
  s.header = s.base + wah::kLaneWords;
  s.args.fault = s.header;
  s.result = reinterpret_cast<uint64_t*>(s.header + 2);
  s.args.status = reinterpret_cast<uint64_t*>(s.base + kScanHeadWords);
  s.args.max_resident = c->max_resident;
  return s;
}

int check_n(wahgpu_ctx* c, uint64_t n, uint32_t rid_base) {
  if (n > WAHGPU_MAX_VALUES)
    return fail(c, WAHGPU_ERANGE, "column of " + std::to_string(n) + " values exceeds " +
                                      std::to_string(WAHGPU_MAX_VALUES));
  if ((uint64_t)rid_base + n > ((uint64_t)1 << 32))
    return fail(c, WAHGPU_ERANGE, "rid_base + n exceeds 2^32 rows");
  return WAHGPU_OK;
}

int read_counts(wahgpu_ctx* c, const wahgpu_counts* d_counts, wahgpu_counts* out) {
  HIPCHK(c, hipMemcpyAsync(c->pinned_counts, d_counts, sizeof(wahgpu_counts),
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  *out = *c->pinned_counts;
  if (out->fault)
    return fail(c, WAHGPU_EHIP, "kernel reported fault " + std::to_string(out->fault) +
                                    " (look-back wait gave up or index out of range)");
  return WAHGPU_OK;
}

// Reads a scan scratch header [fault, pad, result] and its result.
int read_scan_result(wahgpu_ctx* c, const uint32_t* header, uint64_t* out) {
  HIPCHK(c, hipMemcpyAsync(c->pinned_counts, header, 16, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  uint32_t hdr[4];
  memcpy(hdr, c->pinned_counts, 16);
  if (hdr[0])
    return fail(c, WAHGPU_EHIP, "kernel reported fault " + std::to_string(hdr[0]) +
                                    " (look-back wait gave up or index out of range)");
  memcpy(out, &hdr[2], 8);
  return WAHGPU_OK;
}

// Events bracketing phases; elapsed microseconds between consecutive marks.
struct Marks {
  hipEvent_t ev[12];
  int n = 0;
  bool on = false;
  void init(bool enable) {
    on = enable;
    if (on)
      for (auto& e : ev) (void)hipEventCreate(&e);
  }
  void mark(hipStream_t s) {
    if (on && n < 12) (void)hipEventRecord(ev[n++], s);
  }
  double us(int a, int b) {
    if (!on || b >= n) return 0.0;
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ev[a], ev[b]);
    return ms * 1000.0;
  }
  ~Marks() {
    if (on)
      for (auto& e : ev) (void)hipEventDestroy(e);
  }
};

// Scratch of wahgpu_build_device_noalloc: a static device array, one copy
// per GPU (loaded with the code object and held until the process ends).
// Host code never uses the name itself (on the host it is only a stand-in
// symbol): each device's address comes from hipGetSymbolAddress, made once
// per device with that device current.
constexpr uint64_t kNoallocScratchBytes = 112ull << 20;
static __device__ __align__(256) char scratch[kNoallocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
char* g_noalloc_ptr[kMaxDevices] = {};

// Called by wahgpu_create with `device` current.
int noalloc_lookup(wahgpu_ctx* c, int device) {
  if (device >= kMaxDevices) return WAHGPU_OK;  // the entry point then refuses
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (!g_noalloc_ptr[device]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(c, WAHGPU_EHIP, std::string("hipGetSymbolAddress(scratch): ") +
                                      hipGetErrorString(e));
    g_noalloc_ptr[device] = static_cast<char*>(p);
  }
  c->noalloc = g_noalloc_ptr[device];
  return WAHGPU_OK;
}

}  // namespace

extern "C" {

int wahgpu_device_count(int* count) {
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) {
    (void)hipGetLastError();
    c = 0;
  }
  if (count) *count = c;
  return c > 0 ? WAHGPU_OK : WAHGPU_ENODEV;
}

int wahgpu_device_name(int ordinal, char* buf, size_t cap) {
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, ordinal) != hipSuccess) {
    (void)hipGetLastError();
    return WAHGPU_ENODEV;
  }
  if (buf && cap) {
    snprintf(buf, cap, "%s", prop.name);
  }
  return WAHGPU_OK;
}

int wahgpu_device_compute_units(int ordinal) {
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, ordinal) != hipSuccess) {
    (void)hipGetLastError();
    return WAHGPU_ENODEV;
  }
  return prop.multiProcessorCount;
}

int wahgpu_find_device(const char* name) {
  int count = 0;
  wahgpu_device_count(&count);
  for (int i = 0; i < count; ++i) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, i) != hipSuccess) continue;
    if (!name || !*name) return i;
    std::string arch(prop.gcnArchName);
    arch = arch.substr(0, arch.find(':'));
    if (std::string(prop.name) == name || arch == name) return i;
  }
  return -1;
}

int wahgpu_create(wahgpu_ctx** out, int device_ordinal, uint64_t arena_bytes) {
  if (!out) return WAHGPU_EINVAL;
  *out = nullptr;
  int count = 0;
  wahgpu_device_count(&count);
  if (device_ordinal < 0 || device_ordinal >= count) return WAHGPU_ENODEV;
  wahgpu_ctx* c = new wahgpu_ctx();
  c->device = device_ordinal;
  if (hipSetDevice(device_ordinal) != hipSuccess ||
      hipStreamCreateWithFlags(&c->own_stream, hipStreamDefault) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void**>(&c->pinned_counts), 64, hipHostMallocDefault) !=
          hipSuccess) {
    (void)hipGetLastError();
    delete c;
    return WAHGPU_EHIP;
  }
  c->stream = c->own_stream;
  if (noalloc_lookup(c, device_ordinal) != WAHGPU_OK) {
    (void)hipGetLastError();
    wahgpu_destroy(c);
    return WAHGPU_EHIP;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device_ordinal) == hipSuccess)
    c->max_resident = (uint32_t)prop.multiProcessorCount * 4u;
  // Radix ranking: LDS fetch-add when the device ranks same-address lanes in
  // lane order (checked here), else the ballot match.  WAHGPU_RANK=match or
  // =atomic overrides (the latter still requires the check to pass).
  {
    const char* env = getenv("WAHGPU_RANK");
    const std::string want = env ? env : "auto";
    uint32_t* d_ok = nullptr;
    if (want != "match" && hipMalloc(&d_ok, 4) == hipSuccess) {
      *c->pinned_counts = wahgpu_counts{};
      c->pinned_counts->bound = 1;
      uint32_t one = 1;
      if (hipMemcpyAsync(d_ok, &c->pinned_counts->bound, 4, hipMemcpyHostToDevice,
                         c->stream) == hipSuccess) {
        wah::launch_rank_selftest(d_ok, c->stream);
        if (hipMemcpyAsync(&c->pinned_counts->bound, d_ok, 4, hipMemcpyDeviceToHost,
                           c->stream) == hipSuccess &&
            hipStreamSynchronize(c->stream) == hipSuccess)
          c->atomic_rank = c->pinned_counts->bound == one;
      }
      (void)hipFree(d_ok);
    }
    (void)hipGetLastError();
  }
  if (arena_bytes) {
    c->fixed = false;
    int rc = arena_reserve(c, arena_bytes);
    if (rc != WAHGPU_OK) {
      wahgpu_destroy(c);
      return rc;
    }
    c->fixed = true;
  }
  *out = c;
  return WAHGPU_OK;
}

void wahgpu_destroy(wahgpu_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  for (auto e : c->tev) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->desc_copied) (void)hipEventDestroy(c->desc_copied);
  if (c->switch_ev) (void)hipEventDestroy(c->switch_ev);
  if (c->desc_stage) (void)hipHostFree(c->desc_stage);
  if (c->arena) (void)hipFree(c->arena);
  if (c->pinned_counts) (void)hipHostFree(c->pinned_counts);
  if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
  delete c;
}

const char* wahgpu_last_error(const wahgpu_ctx* c) { return c ? c->err.c_str() : ""; }
void* wahgpu_stream(wahgpu_ctx* c) { return c ? (void*)c->stream : nullptr; }

// All of a context's work shares its arena (scratch, look-back state, the
// descriptor), so work on the new stream must not start before the work
// already queued on the old one: an event recorded on the old stream is
// waited for by the new one.  Arena regrowth and graph teardown then only
// need to synchronize the current stream.  While either stream is being
// captured into a graph the event cannot be used, and the switch is refused.
int wahgpu_set_stream(wahgpu_ctx* c, void* stream) {
  if (!c) return WAHGPU_EINVAL;
  hipStream_t next = stream ? (hipStream_t)stream : c->own_stream;
  if (next == c->stream) return WAHGPU_OK;
  HIPCHK(c, hipSetDevice(c->device));
  hipStreamCaptureStatus cs_old = hipStreamCaptureStatusNone, cs_new = hipStreamCaptureStatusNone;
  HIPCHK(c, hipStreamIsCapturing(c->stream, &cs_old));
  HIPCHK(c, hipStreamIsCapturing(next, &cs_new));
  if (cs_old != hipStreamCaptureStatusNone || cs_new != hipStreamCaptureStatusNone)
    return fail(c, WAHGPU_EINVAL, "cannot switch streams while either stream is capturing");
  if (!c->switch_ev) HIPCHK(c, hipEventCreateWithFlags(&c->switch_ev, hipEventDisableTiming));
  HIPCHK(c, hipEventRecord(c->switch_ev, c->stream));
  HIPCHK(c, hipStreamWaitEvent(next, c->switch_ev, 0));
  c->stream = next;
  return WAHGPU_OK;
}
uint64_t wahgpu_arena_bytes(const wahgpu_ctx* c) { return c ? c->arena_cap : 0; }
int wahgpu_rank_mode(const wahgpu_ctx* c) { return c && c->atomic_rank ? 1 : 0; }

int wahgpu_synchronize(wahgpu_ctx* c) {
  if (!c) return WAHGPU_EINVAL;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return WAHGPU_OK;
}

// ---- fused build ---------------------------------------------------------

namespace {
struct EventHook : wah::LaunchHook {
  hipEvent_t* ev;
  int n = 0;
  void mark(hipStream_t s) override { (void)hipEventRecord(ev[n++], s); }
};
}  // namespace

static int build_device_impl(wahgpu_ctx* c, const uint32_t* d_values, uint64_t n,
                             uint32_t rid_base, uint32_t* d_keys, uint32_t* d_offsets,
                             uint32_t* d_lengths, uint32_t* d_words, wahgpu_counts* d_counts,
                             Marks* marks, char* scratch_base = nullptr) {
  BuildArgs a{};
  if (scratch_base) {
    // carve from the caller's scratch instead of the arena
    char* const arena = c->arena;
    const uint64_t bump = c->bump;
    c->arena = scratch_base;
    c->bump = 0;
    carve_build(c, &a, caps_for(c, n), shape_for(c, n));
    c->arena = arena;
    c->bump = bump;
  } else {
    carve_build(c, &a, caps_for(c, n), shape_for(c, n));
  }
  set_one(&a, n, rid_base, wide_ok(c, n));
  a.values = d_values;
  a.out_keys = d_keys;
  a.out_offsets = d_offsets;
  a.out_lengths = d_lengths;
  a.out_words = d_words;
  a.counts = d_counts;
  EventHook hook;
  EventHook* hp = nullptr;
  if (c->timing && c->tcalls < c->tmax) {
    hook.ev = &c->tev[(size_t)c->tcalls * (WAHGPU_TIMING_GROUPS + 1)];
    hp = &hook;
    ++c->tcalls;
    hook.mark(c->stream);
  }
  ORDERED(c);
  wah::launch_sort(a, c->stream, hp);
  LAUNCHCHK(c);
  if (marks) marks->mark(c->stream);
  wah::launch_tail(a, c->stream, hp);
  LAUNCHCHK(c);
  if (hp) hook.mark(c->stream);
  return WAHGPU_OK;
}

int wahgpu_timing_begin(wahgpu_ctx* c, uint32_t max_calls) {
  if (!c || !max_calls) return WAHGPU_EINVAL;
  HIPCHK(c, hipSetDevice(c->device));
  for (auto e : c->tev) (void)hipEventDestroy(e);
  c->tev.assign((size_t)max_calls * (WAHGPU_TIMING_GROUPS + 1), nullptr);
  for (auto& e : c->tev) HIPCHK(c, hipEventCreate(&e));
  c->tmax = max_calls;
  c->tcalls = 0;
  c->timing = true;
  return WAHGPU_OK;
}

int wahgpu_timing_end(wahgpu_ctx* c, double* ms, uint32_t* calls) {
  if (!c || !ms) return WAHGPU_EINVAL;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int g = 0; g < WAHGPU_TIMING_GROUPS; ++g) ms[g] = 0.0;
  for (uint32_t i = 0; i < c->tcalls; ++i) {
    hipEvent_t* ev = &c->tev[(size_t)i * (WAHGPU_TIMING_GROUPS + 1)];
    for (int g = 0; g < WAHGPU_TIMING_GROUPS; ++g) {
      float t = 0.f;
      HIPCHK(c, hipEventElapsedTime(&t, ev[g], ev[g + 1]));
      ms[g] += t;
    }
  }
  if (calls) *calls = c->tcalls;
  c->timing = false;
  for (auto e : c->tev) (void)hipEventDestroy(e);
  c->tev.clear();
  c->tcalls = c->tmax = 0;
  return WAHGPU_OK;
}

int wahgpu_build_device(wahgpu_ctx* c, const uint32_t* d_values, uint64_t n, uint32_t rid_base,
                        uint32_t* d_keys, uint32_t* d_offsets, uint32_t* d_lengths,
                        uint32_t* d_words, wahgpu_counts* d_counts) {
  if (!c || !d_counts) return fail(c, WAHGPU_EINVAL, "null context or counts");
  if (n && (!d_values || !d_keys || !d_offsets || !d_lengths || !d_words))
    return fail(c, WAHGPU_EINVAL, "null buffer");
  int rc = check_n(c, n, rid_base);
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  rc = arena_reserve(c, build_scratch_bytes(caps_for(c, n)));
  if (rc) return rc;
  return build_device_impl(c, d_values, n, rid_base, d_keys, d_offsets, d_lengths, d_words,
                           d_counts, nullptr);
}

int wahgpu_build_device_noalloc(wahgpu_ctx* c, const uint32_t* d_values, uint64_t n,
                                uint32_t rid_base, uint32_t* d_keys, uint32_t* d_offsets,
                                uint32_t* d_lengths, uint32_t* d_words,
                                wahgpu_counts* d_counts) {
  if (!c || !d_counts) return fail(c, WAHGPU_EINVAL, "null context or counts");
  if (n && (!d_values || !d_keys || !d_offsets || !d_lengths || !d_words))
    return fail(c, WAHGPU_EINVAL, "null buffer");
  if (n > WAHGPU_NOALLOC_MAX_VALUES)
    return fail(c, WAHGPU_ERANGE, "wahgpu_build_device_noalloc: column of " +
                                      std::to_string(n) + " values exceeds the limit of " +
                                      std::to_string(WAHGPU_NOALLOC_MAX_VALUES));
  int rc = check_n(c, n, rid_base);
  if (rc) return rc;
  if (!c->noalloc) return fail(c, WAHGPU_EHIP, "no static scratch on this device");
  // every tile shape fits: 24 bytes per value plus per-tile state
  if (build_scratch_bytes(caps_for(c, n)) > kNoallocScratchBytes)
    return fail(c, WAHGPU_ERANGE, "wahgpu_build_device_noalloc: scratch of " +
                                      std::to_string(kNoallocScratchBytes) + " bytes too small");
  HIPCHK(c, hipSetDevice(c->device));
  // The build zeroes the state it uses itself (k_begin, k_histogram): the
  // array keeps the previous call's contents and nothing is cleared here.
  return build_device_impl(c, d_values, n, rid_base, d_keys, d_offsets, d_lengths, d_words,
                           d_counts, nullptr, c->noalloc);
}

int wahgpu_build(wahgpu_ctx* c, const uint32_t* values, uint64_t n, uint32_t rid_base,
                 uint32_t* keys, uint32_t* offsets, uint32_t* lengths, uint64_t keys_cap,
                 uint32_t* words, uint64_t words_cap, wahgpu_counts* counts,
                 wahgpu_phase_times* times) {
  if (!c || !counts) return fail(c, WAHGPU_EINVAL, "null context or counts");
  if (n && !values) return fail(c, WAHGPU_EINVAL, "null values");
  int rc = check_n(c, n, rid_base);
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  uint64_t nn = n ? n : 1;
  uint64_t bytes = build_scratch_bytes(caps_for(c, n)) + align_up(nn * 4) * 4 + align_up(2 * nn * 4) +
                   align_up(sizeof(wahgpu_counts));
  rc = arena_reserve(c, bytes);
  if (rc) return rc;
  auto t0 = std::chrono::steady_clock::now();
  uint32_t* d_values = take<uint32_t>(c, nn);
  uint32_t* d_keys = take<uint32_t>(c, nn);
  uint32_t* d_offsets = take<uint32_t>(c, nn);
  uint32_t* d_lengths = take<uint32_t>(c, nn);
  uint32_t* d_words = take<uint32_t>(c, 2 * nn);
  wahgpu_counts* d_counts = take<wahgpu_counts>(c, 1);
  Marks m;
  m.init(times != nullptr);
  m.mark(c->stream);
  if (n) HIPCHK(c, hipMemcpyAsync(d_values, values, n * 4, hipMemcpyHostToDevice, c->stream));
  m.mark(c->stream);
  rc = build_device_impl(c, d_values, n, rid_base, d_keys, d_offsets, d_lengths, d_words,
                         d_counts, &m);
  if (rc) return rc;
  m.mark(c->stream);
  rc = read_counts(c, d_counts, counts);
  if (rc) return rc;
  if (counts->keycnt > keys_cap || counts->length > words_cap)
    return fail(c, WAHGPU_EINVAL, "output capacity too small");
  if (counts->length && !words) return fail(c, WAHGPU_EINVAL, "null words");
  if (counts->keycnt && (!keys || !offsets || !lengths))
    return fail(c, WAHGPU_EINVAL, "null keys/offsets/lengths");
  if (counts->length)
    HIPCHK(c, hipMemcpyAsync(words, d_words, counts->length * 4, hipMemcpyDeviceToHost,
                             c->stream));
  if (counts->keycnt) {
    HIPCHK(c, hipMemcpyAsync(keys, d_keys, counts->keycnt * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(offsets, d_offsets, counts->keycnt * 4, hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipMemcpyAsync(lengths, d_lengths, counts->keycnt * 4, hipMemcpyDeviceToHost,
                             c->stream));
  }
  m.mark(c->stream);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (times) {
    memset(times, 0, sizeof(*times));
    times->transfer_to = m.us(0, 1);
    times->sort = m.us(1, 2);
    times->column_length = m.us(2, 3);  // fused tail: chunks..column length
    times->transfer_back = m.us(3, 4);
    times->total = std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0)
                       .count();
  }
  return WAHGPU_OK;
}

// ---- batched build -----------------------------------------------------

namespace {

// One launch set of a batch: the columns whose length bucket maps to one
// radix tile shape, with their descriptor (dims + columns with tile and
// chunk prefixes in that shape's tiles) in host memory.
struct SubBatch {
  int shape;
  std::vector<char> desc;
  Caps caps;
};

// Validates the columns and splits them into sub-batches by tuned shape
// (every compiled shape when `all_shapes`, empty ones included, so that a
// captured graph fits any batch).  *shared = element-wise maximum of the
// sub-batches' caps (the scratch they share, one after another).
int plan_batch(wahgpu_ctx* c, const wahgpu_column* cols, uint32_t ncols, uint64_t values_len,
               bool all_shapes, std::vector<SubBatch>* subs, Caps* shared) {
  uint64_t end = 0;
  std::vector<uint32_t> members[wah::kRadixShapes];
  for (uint32_t i = 0; i < ncols; ++i) {
    const wahgpu_column& col = cols[i];
    int rc = check_n(c, col.n, col.rid_base);
    if (rc) return rc;
    if (col.offset < end || col.offset > values_len || col.n > values_len - col.offset)
      return fail(c, WAHGPU_EINVAL, "column " + std::to_string(i) +
                                        " overlaps the previous one or exceeds values_len");
    end = col.offset + col.n;
    members[shape_for(c, col.n)].push_back(i);
  }
  subs->clear();
  *shared = Caps{values_len, 0, 0, 0, 0};
  bool tuned[wah::kRadixShapes] = {};
  for (uint32_t b = 0; b < WAHGPU_TUNE_BUCKETS; ++b) tuned[c->tune_shape[b]] = true;
  for (int sh = 0; sh < wah::kRadixShapes; ++sh) {
    if (members[sh].empty() && !(all_shapes && tuned[sh])) continue;
    SubBatch sb;
    sb.shape = sh;
    const uint32_t m = (uint32_t)members[sh].size();
    sb.desc.assign(sizeof(wah::DescHead) + (uint64_t)m * sizeof(wah::Column), 0);
    auto* head = reinterpret_cast<wah::DescHead*>(sb.desc.data());
    wah::Column* dc = wah::desc_cols(head);
    uint64_t st = 0, tt = 0, hc = 0, wt = 0;
    uint32_t ws = 0;  // columns that may take wide digits, each with a WideState
    const uint32_t tile = wah::radix_tile(sh);
    for (uint32_t j = 0; j < m; ++j) {
      const wahgpu_column& col = cols[members[sh][j]];
      const bool wide = wide_ok(c, col.n);
      dc[j] = wah::Column{col.offset, col.n, col.rid_base, (uint32_t)st, (uint32_t)tt,
                          (uint32_t)hc, members[sh][j], wide ? ws + 1u : 0u, (uint32_t)wt, 0u};
      st += wah::sort_tiles_of(col.n, tile);
      tt += wah::tail_tiles_of(col.n);
      hc += wah::hist_chunks_of(col.n);
      if (wide) {
        ++ws;
        wt += wide_tiles_of(col.n);
      }
    }
    if (st > 0xFFFFFFF0ull || hc > 0xFFFFFFF0ull)
      return fail(c, WAHGPU_ERANGE, "batch too large for 32-bit tile numbering");
    head->dims = wah::Dims{m, (uint32_t)st, (uint32_t)tt, (uint32_t)hc, (uint32_t)wt, 0u};
    sb.caps = Caps{values_len, m, (uint32_t)st, (uint32_t)tt, (uint32_t)hc};
    sb.caps.wide_slots = ws;
    sb.caps.wide_tiles = (uint32_t)wt;
    shared->cols = std::max(shared->cols, m);
    shared->sort_tiles = std::max(shared->sort_tiles, (uint32_t)st);
    shared->tail_tiles = std::max(shared->tail_tiles, (uint32_t)tt);
    shared->hist_chunks = std::max(shared->hist_chunks, (uint32_t)hc);
    shared->wide_slots = std::max(shared->wide_slots, ws);
    shared->wide_tiles = std::max(shared->wide_tiles, (uint32_t)wt);
    subs->push_back(std::move(sb));
  }
  return WAHGPU_OK;
}

// Copies the descriptor to the device on the context's stream through pinned
// staging (waits for the previous copy out of the staging buffer first).
int upload_desc(wahgpu_ctx* c, const std::vector<char>& d, wah::DescHead* dst) {
  if (!c->desc_copied) HIPCHK(c, hipEventCreateWithFlags(&c->desc_copied, hipEventDisableTiming));
  HIPCHK(c, hipEventSynchronize(c->desc_copied));
  if (d.size() > c->desc_stage_cap) {
    if (c->desc_stage) (void)hipHostFree(c->desc_stage);
    c->desc_stage = nullptr;
    c->desc_stage_cap = 0;
    uint64_t cap = align_up(d.size()) * 2;
    HIPCHK(c, hipHostMalloc(reinterpret_cast<void**>(&c->desc_stage), cap, hipHostMallocDefault));
    c->desc_stage_cap = cap;
  }
  memcpy(c->desc_stage, d.data(), d.size());
  HIPCHK(c, hipMemcpyAsync(dst, c->desc_stage, d.size(), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipEventRecord(c->desc_copied, c->stream));
  return WAHGPU_OK;
}

// Caps of the reserved layout for graph replay.
Caps reserved_caps(const wahgpu_ctx* c) {
  const uint64_t v = c->res_values;
  uint32_t min_tile = wah::radix_tile(0);
  for (int sh = 1; sh < wah::kRadixShapes; ++sh) min_tile = std::min(min_tile, wah::radix_tile(sh));
  Caps k{v, c->res_cols, (uint32_t)(v / min_tile + c->res_cols),
         (uint32_t)(v / wah::kTailTile + c->res_cols),
         (uint32_t)(v / wah::kHistChunk + c->res_cols)};
  bool wide = false;  // some bucket allows wide digits
  for (uint32_t b = 1; b < WAHGPU_TUNE_BUCKETS; ++b) wide |= c->tune_digits[b] == 11u;
  if (wide) {
    k.wide_slots = (uint32_t)std::min<uint64_t>(c->res_cols, v / wah::kWideMinValues);
    k.wide_tiles = (uint32_t)(v / wah::kWideTile + k.wide_slots);
  }
  return k;
}

// Scratch bytes of a planned batch (shared scratch + one descriptor region
// per sub-batch).
uint64_t batch_scratch_bytes(const Caps& shared, size_t nsubs) {
  return build_scratch_bytes(shared, (uint32_t)(nsubs ? nsubs : 1));
}

// Carving from a caller's buffer instead of the arena, for the lifetime of
// the guard.
struct ScratchSwap {
  wahgpu_ctx* c;
  char* arena;
  uint64_t bump;
  bool on;
  ScratchSwap(wahgpu_ctx* ctx, char* ext) : c(ctx), arena(ctx->arena), bump(ctx->bump), on(ext) {
    if (on) {
      c->arena = ext;
      c->bump = 0;
    }
  }
  ~ScratchSwap() {
    if (on) {
      c->arena = arena;
      c->bump = bump;
    }
  }
};

// ext: scratch supplied by the caller (ext_bytes long; no arena, no graph).
int batch_device_impl(wahgpu_ctx* c, const uint32_t* d_values, uint64_t values_len,
                      const wahgpu_column* cols, uint32_t ncols, uint32_t* d_keys,
                      uint32_t* d_offsets, uint32_t* d_lengths, uint32_t* d_words,
                      wahgpu_counts* d_counts, uint64_t extra_bytes, char* ext = nullptr,
                      uint64_t ext_bytes = 0) {
  std::vector<SubBatch> subs;
  Caps shared;
  const bool reserve = !ext && c->res_cols != 0 && c->force_shape < 0;
  int rc = plan_batch(c, cols, ncols, values_len, reserve, &subs, &shared);
  if (rc) return rc;
  const bool graph = reserve && reserved_caps(c).covers(shared) && ncols <= c->res_cols;
  const Caps k = graph ? reserved_caps(c) : shared;
  const uint64_t need = batch_scratch_bytes(k, subs.size());
  if (ext) {
    if (need > ext_bytes)
      return fail(c, WAHGPU_ENOMEM, "workspace of " + std::to_string(ext_bytes) +
                                        " bytes is too small for " + std::to_string(need));
  } else {
    // extra_bytes: room the caller carves after the scratch (host batches)
    rc = arena_reserve(c, need + extra_bytes);
    if (rc) return rc;
  }
  ScratchSwap swap(c, ext);
  BuildArgs base{};
  carve_build(c, &base, k, 0);
  std::vector<BuildArgs> args;
  for (size_t i = 0; i < subs.size(); ++i) {
    BuildArgs a = base;
    a.shape = subs[i].shape;
    if (i > 0) a.desc = carve_desc(c, k.cols);
    a.inline_desc = false;
    a.values = d_values;
    a.out_keys = d_keys;
    a.out_offsets = d_offsets;
    a.out_lengths = d_lengths;
    a.out_words = d_words;
    a.counts = d_counts;
    rc = upload_desc(c, subs[i].desc, a.desc);
    if (rc) return rc;
    args.push_back(a);
  }
  if (graph) {
    std::vector<uintptr_t> key = {(uintptr_t)c->arena, (uintptr_t)c->stream,
                                  (uintptr_t)d_values, (uintptr_t)d_keys,
                                  (uintptr_t)d_offsets, (uintptr_t)d_lengths,
                                  (uintptr_t)d_words, (uintptr_t)d_counts,
                                  (uintptr_t)k.n, (uintptr_t)k.cols, (uintptr_t)subs.size()};
    if (!c->gexec || key != c->gkey) {
      if (c->gexec) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        (void)hipGraphExecDestroy(c->gexec);
        c->gexec = nullptr;
      }
      HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      for (const BuildArgs& a : args) {
        wah::launch_sort(a, c->stream);
        wah::launch_tail(a, c->stream);
      }
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamEndCapture(c->stream, &g);
      if (e != hipSuccess) return fail(c, WAHGPU_EHIP, std::string("capture: ") + hipGetErrorString(e));
      e = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) {
        c->gexec = nullptr;
        return fail(c, WAHGPU_EHIP, std::string("graph instantiate: ") + hipGetErrorString(e));
      }
      c->gkey = key;
      ++c->captures;
    }
    ORDERED(c);  // after the capture (which enqueues nothing) and its synchronization
    HIPCHK(c, hipGraphLaunch(c->gexec, c->stream));
    ++c->replays;
    return WAHGPU_OK;
  }
  ORDERED(c);
  for (const BuildArgs& a : args) {
    EventHook hook;
    EventHook* hp = nullptr;
    if (c->timing && c->tcalls < c->tmax && args.size() == 1) {
      hook.ev = &c->tev[(size_t)c->tcalls * (WAHGPU_TIMING_GROUPS + 1)];
      hp = &hook;
      ++c->tcalls;
      hook.mark(c->stream);
    }
    wah::launch_sort(a, c->stream, hp);
    LAUNCHCHK(c);
    wah::launch_tail(a, c->stream, hp);
    LAUNCHCHK(c);
    if (hp) hook.mark(c->stream);
  }
  return WAHGPU_OK;
}

}  // namespace

int wahgpu_build_batch_device(wahgpu_ctx* c, const uint32_t* d_values, uint64_t values_len,
                              const wahgpu_column* cols, uint32_t ncols, uint32_t* d_keys,
                              uint32_t* d_offsets, uint32_t* d_lengths, uint32_t* d_words,
                              wahgpu_counts* d_counts) {
  if (!c) return WAHGPU_EINVAL;
  if (ncols == 0) return WAHGPU_OK;
  if (!cols || !d_counts) return fail(c, WAHGPU_EINVAL, "null columns or counts");
  if (values_len && (!d_values || !d_keys || !d_offsets || !d_lengths || !d_words))
    return fail(c, WAHGPU_EINVAL, "null buffer");
  HIPCHK(c, hipSetDevice(c->device));
  return batch_device_impl(c, d_values, values_len, cols, ncols, d_keys, d_offsets, d_lengths,
                           d_words, d_counts, 0);
}

int wahgpu_build_batch(wahgpu_ctx* c, const uint32_t* const* values, const uint64_t* n,
                       const uint32_t* rid_base, uint32_t ncols, wahgpu_result* results) {
  if (!c) return WAHGPU_EINVAL;
  if (ncols == 0) return WAHGPU_OK;
  if (!values || !n || !results) return fail(c, WAHGPU_EINVAL, "null argument");
  HIPCHK(c, hipSetDevice(c->device));
  // columns at 64-value aligned offsets (16-byte vector loads per column)
  std::vector<wahgpu_column> cols(ncols);
  uint64_t total = 0;
  for (uint32_t i = 0; i < ncols; ++i) {
    if (n[i] && !values[i]) return fail(c, WAHGPU_EINVAL, "null values of a column");
    cols[i] = wahgpu_column{total, n[i], rid_base ? rid_base[i] : 0u, 0u};
    total += (n[i] + 63) / 64 * 64;
  }
  const uint64_t tn = total ? total : 1;
  const uint64_t io = align_up(tn * 4) * 4 + align_up(2 * tn * 4) +
                      align_up((uint64_t)ncols * sizeof(wahgpu_counts));
  // The scratch is carved first by batch_device_impl; the io buffers follow
  // it.  Reserve both, then carve io at the end of the scratch.
  std::vector<SubBatch> subs;
  Caps shared;
  const bool reserve = c->res_cols != 0 && c->force_shape < 0;
  int rc = plan_batch(c, cols.data(), ncols, total, reserve, &subs, &shared);
  if (rc) return rc;
  const bool graph = reserve && reserved_caps(c).covers(shared) && ncols <= c->res_cols;
  const uint64_t scratch = batch_scratch_bytes(graph ? reserved_caps(c) : shared, subs.size());
  rc = arena_reserve(c, scratch + io);
  if (rc) return rc;
  c->bump = scratch;
  uint32_t* d_values = take<uint32_t>(c, tn);
  uint32_t* d_keys = take<uint32_t>(c, tn);
  uint32_t* d_offsets = take<uint32_t>(c, tn);
  uint32_t* d_lengths = take<uint32_t>(c, tn);
  uint32_t* d_words = take<uint32_t>(c, 2 * tn);
  wahgpu_counts* d_counts = take<wahgpu_counts>(c, ncols);
  for (uint32_t i = 0; i < ncols; ++i)
    if (n[i])
      HIPCHK(c, hipMemcpyAsync(d_values + cols[i].offset, values[i], n[i] * 4,
                               hipMemcpyHostToDevice, c->stream));
  rc = batch_device_impl(c, d_values, total, cols.data(), ncols, d_keys, d_offsets, d_lengths,
                         d_words, d_counts, io);
  if (rc) return rc;
  std::vector<wahgpu_counts> hc(ncols);
  HIPCHK(c, hipMemcpyAsync(hc.data(), d_counts, ncols * sizeof(wahgpu_counts),
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (uint32_t i = 0; i < ncols; ++i) {
    const wahgpu_counts& k = hc[i];
    wahgpu_result& r = results[i];
    r.counts = k;
    if (k.fault)
      return fail(c, WAHGPU_EHIP, "kernel reported fault " + std::to_string(k.fault) +
                                      " (look-back wait gave up or index out of range)");
    if (k.keycnt > r.keys_cap || k.length > r.words_cap)
      return fail(c, WAHGPU_EINVAL, "output capacity too small for column " + std::to_string(i));
    if ((k.length && !r.words) || (k.keycnt && (!r.keys || !r.offsets || !r.lengths)))
      return fail(c, WAHGPU_EINVAL, "null output of column " + std::to_string(i));
    const uint64_t o = cols[i].offset;
    if (k.length)
      HIPCHK(c, hipMemcpyAsync(r.words, d_words + 2 * o, k.length * 4, hipMemcpyDeviceToHost,
                               c->stream));
    if (k.keycnt) {
      HIPCHK(c, hipMemcpyAsync(r.keys, d_keys + o, k.keycnt * 4, hipMemcpyDeviceToHost, c->stream));
      HIPCHK(c, hipMemcpyAsync(r.offsets, d_offsets + o, k.keycnt * 4, hipMemcpyDeviceToHost,
                               c->stream));
      HIPCHK(c, hipMemcpyAsync(r.lengths, d_lengths + o, k.keycnt * 4, hipMemcpyDeviceToHost,
                               c->stream));
    }
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return WAHGPU_OK;
}

int wahgpu_batch_reserve(wahgpu_ctx* c, uint64_t max_values, uint32_t max_cols) {
  if (!c) return WAHGPU_EINVAL;
  if ((max_values == 0) != (max_cols == 0))
    return fail(c, WAHGPU_EINVAL, "max_values and max_cols are both zero or both nonzero");
  HIPCHK(c, hipSetDevice(c->device));
  if (c->gexec) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    (void)hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
    c->gkey.clear();
  }
  c->res_values = max_values;
  c->res_cols = max_cols;
  if (max_cols) {
    int rc = arena_reserve(c, build_scratch_bytes(reserved_caps(c)));
    if (rc) return rc;
  }
  return WAHGPU_OK;
}

int wahgpu_get_tuning(const wahgpu_ctx* c, wahgpu_tuning* out) {
  if (!c || !out) return WAHGPU_EINVAL;
  for (uint32_t b = 0; b < WAHGPU_TUNE_BUCKETS; ++b) {
    out->bucket_log2[b] = kBucketLog2[b];
    out->shape[b] = c->tune_shape[b];
    out->digit_bits[b] = c->tune_digits[b];
    for (uint32_t s = 0; s < WAHGPU_TUNE_SHAPES; ++s) out->ms[b][s] = c->tune_ms[b][s];
    out->digit_ms[b][0] = c->tune_digit_ms[b][0];
    out->digit_ms[b][1] = c->tune_digit_ms[b][1];
  }
  for (int sh = 0; sh < WAHGPU_TUNE_SHAPES; ++sh) {
    out->shape_threads[sh] = (uint32_t)wah::radix_threads(sh);
    out->shape_items[sh] = (uint32_t)wah::radix_items(sh);
  }
  return WAHGPU_OK;
}

static void drop_graph(wahgpu_ctx* c) {
  if (c->gexec) {
    (void)hipStreamSynchronize(c->stream);
    (void)hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
    c->gkey.clear();
  }
}

static_assert(WAHGPU_TUNE_SHAPES == wah::kRadixShapes, "public shape count");

int wahgpu_set_tuning(wahgpu_ctx* c, const wahgpu_tuning* in) {
  if (!c || !in) return WAHGPU_EINVAL;
  for (uint32_t b = 0; b < WAHGPU_TUNE_BUCKETS; ++b) {
    if (in->shape[b] >= WAHGPU_TUNE_SHAPES) return fail(c, WAHGPU_EINVAL, "unknown tile shape");
    if (in->digit_bits[b] != 0 && in->digit_bits[b] != 8 && in->digit_bits[b] != 11)
      return fail(c, WAHGPU_EINVAL, "digit width must be 8 or 11");
  }
  HIPCHK(c, hipSetDevice(c->device));
  drop_graph(c);  // the set of launch sets a batch runs may change
  for (uint32_t b = 0; b < WAHGPU_TUNE_BUCKETS; ++b) {
    c->tune_shape[b] = in->shape[b];
    c->tune_digits[b] = in->digit_bits[b] == 11 ? 11u : 8u;
  }
  return WAHGPU_OK;
}

// Times every tile shape on a synthetic batch of each length bucket and keeps
// the fastest: columns of one length per bucket (2^13, 2^15.5, 2^18.5, 2^22,
// 2^26 values), uniform over 4096 values, 2^24 values per bucket (one 2^26
// column for the last).
int wahgpu_autotune(wahgpu_ctx* c, wahgpu_tuning* out) {
  if (!c) return WAHGPU_EINVAL;
  HIPCHK(c, hipSetDevice(c->device));
  drop_graph(c);
  const double len_log2[WAHGPU_TUNE_BUCKETS] = {13.0, 15.5, 18.5, 22.0, 26.0};
  const uint64_t total = 1ull << 26;  // device buffers: the largest bucket's values
  uint32_t *d_v = nullptr, *d_k = nullptr, *d_o = nullptr, *d_l = nullptr, *d_w = nullptr;
  wahgpu_counts* d_c = nullptr;
  const uint64_t maxcols = 4096;
  int rc = WAHGPU_OK;
  hipEvent_t ev[2] = {nullptr, nullptr};
  std::vector<uint32_t> host(total);
  auto cleanup = [&]() {
    if (d_v) (void)hipFree(d_v);
    if (d_k) (void)hipFree(d_k);
    if (d_o) (void)hipFree(d_o);
    if (d_l) (void)hipFree(d_l);
    if (d_w) (void)hipFree(d_w);
    if (d_c) (void)hipFree(d_c);
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
    c->force_shape = -1;
    c->force_digits = 0;
  };
  if (hipMalloc(&d_v, total * 4) != hipSuccess || hipMalloc(&d_k, total * 4) != hipSuccess ||
      hipMalloc(&d_o, total * 4) != hipSuccess || hipMalloc(&d_l, total * 4) != hipSuccess ||
      hipMalloc(&d_w, 2 * total * 4) != hipSuccess ||
      hipMalloc(&d_c, maxcols * sizeof(wahgpu_counts)) != hipSuccess ||
      hipEventCreate(&ev[0]) != hipSuccess || hipEventCreate(&ev[1]) != hipSuccess) {
    (void)hipGetLastError();
    cleanup();
    return fail(c, WAHGPU_ENOMEM, "autotune: device allocation failed");
  }
  // best of three timed builds (after one warm-up) of the batch `cols`; a
  // build that reported a fault is not a valid timing and fails the tune
  std::string fault_msg;
  auto time_batch = [&](const std::vector<wahgpu_column>& cols, uint64_t values,
                        std::vector<wahgpu_counts>& hc, float* best_ms) {
    const uint32_t ncols = (uint32_t)cols.size();
    *best_ms = 1e30f;
    for (int rep = 0; rep < 4 && rc == WAHGPU_OK; ++rep) {
      (void)hipEventRecord(ev[0], c->stream);
      rc = batch_device_impl(c, d_v, values, cols.data(), ncols, d_k, d_o, d_l, d_w, d_c, 0);
      (void)hipEventRecord(ev[1], c->stream);
      if (rc == WAHGPU_OK && hipEventSynchronize(ev[1]) != hipSuccess) rc = WAHGPU_EHIP;
      float ms = 0.f;
      if (rc == WAHGPU_OK) (void)hipEventElapsedTime(&ms, ev[0], ev[1]);
      if (rc == WAHGPU_OK &&
          hipMemcpy(hc.data(), d_c, ncols * sizeof(wahgpu_counts), hipMemcpyDeviceToHost) !=
              hipSuccess)
        rc = WAHGPU_EHIP;
      for (uint32_t i = 0; rc == WAHGPU_OK && i < ncols; ++i)
        if (hc[i].fault) {
          fault_msg = "autotune: shape " + std::to_string(c->force_shape) + ", digits " +
                      std::to_string(c->force_digits) + " reported fault " +
                      std::to_string(hc[i].fault);
          rc = WAHGPU_EHIP;
        }
      if (rc == WAHGPU_OK && rep > 0 && ms < *best_ms) *best_ms = ms;  // first run warms up
    }
  };
  auto generate = [&](uint64_t values, uint32_t to, uint64_t seed) {
    wahgpu_gen_params gp{};
    gp.dist = WAHGPU_DIST_UNIFORM;
    gp.from = 0;
    gp.to = to;
    gp.seed = seed;
    if (wahgpu_generate(&gp, host.data(), values) != WAHGPU_OK) rc = WAHGPU_EINVAL;
    if (rc == WAHGPU_OK &&
        hipMemcpy(d_v, host.data(), values * 4, hipMemcpyHostToDevice) != hipSuccess)
      rc = WAHGPU_EHIP;
  };
  for (uint32_t b = 0; b < WAHGPU_TUNE_BUCKETS && rc == WAHGPU_OK; ++b) {
    const uint64_t len = (uint64_t)std::pow(2.0, len_log2[b]);
    const uint64_t step = (len + 63) / 64 * 64;
    const uint64_t values = std::max<uint64_t>(len, 1ull << 24);  // per bucket
    const uint32_t ncols = (uint32_t)std::min<uint64_t>(maxcols, values / step);
    std::vector<wahgpu_column> cols(ncols);
    std::vector<wahgpu_counts> hc(ncols);
    for (uint32_t i = 0; i < ncols; ++i) cols[i] = wahgpu_column{i * step, len, 0u, 0u};
    // tile shape, with 8-bit digits, on keys of 12 bits
    generate(values, 4095, 1000 + b);
    double best = 1e30;
    c->force_digits = 8;
    for (int sh = 0; sh < WAHGPU_TUNE_SHAPES && rc == WAHGPU_OK; ++sh) {
      c->force_shape = sh;
      float ms = 0.f;
      time_batch(cols, values, hc, &ms);
      c->tune_ms[b][sh] = ms;
      if (rc == WAHGPU_OK && ms < best) {
        best = ms;
        c->tune_shape[b] = (uint32_t)sh;
      }
    }
    // digit width in that shape, on keys of 10 bits (one wide pass against
    // two 8-bit ones); columns too short for wide digits keep 8 bits
    c->tune_digits[b] = 8;
    c->tune_digit_ms[b][0] = c->tune_digit_ms[b][1] = 0.0;
    if (rc == WAHGPU_OK && len >= wah::kWideMinValues) {
      generate(values, 1023, 2000 + b);
      c->force_shape = (int)c->tune_shape[b];
      float ms8 = 0.f, ms11 = 0.f;
      c->force_digits = 8;
      time_batch(cols, values, hc, &ms8);
      c->force_digits = 11;
      if (rc == WAHGPU_OK) time_batch(cols, values, hc, &ms11);
      for (uint32_t i = 0; rc == WAHGPU_OK && i < ncols; ++i)
        if (hc[i].mode != wah::kModeWideReport) {
          fault_msg = "autotune: a 10-bit column did not take wide digits";
          rc = WAHGPU_EHIP;
        }
      c->tune_digit_ms[b][0] = ms8;
      c->tune_digit_ms[b][1] = ms11;
      if (rc == WAHGPU_OK && ms11 < ms8) c->tune_digits[b] = 11;
    }
  }
  cleanup();
  if (rc != WAHGPU_OK && !fault_msg.empty()) return fail(c, rc, fault_msg);
  if (rc != WAHGPU_OK) return rc == WAHGPU_EHIP ? fail(c, rc, "autotune: build failed") : rc;
  return out ? wahgpu_get_tuning(c, out) : WAHGPU_OK;
}

int wahgpu_graph_stats(const wahgpu_ctx* c, uint64_t* captures, uint64_t* replays) {
  if (!c) return WAHGPU_EINVAL;
  if (captures) *captures = c->captures;
  if (replays) *replays = c->replays;
  return WAHGPU_OK;
}

// ---- staged build (one kernel group per reference stage) ------------------

int wahgpu_build_staged(wahgpu_ctx* c, const uint32_t* values, uint64_t n, uint32_t* keys,
                        uint32_t* offsets, uint32_t* lengths, uint64_t keys_cap,
                        uint32_t* words, uint64_t words_cap, wahgpu_counts* counts,
                        wahgpu_phase_times* times) {
  if (!c || !counts) return fail(c, WAHGPU_EINVAL, "null context or counts");
  if (n && !values) return fail(c, WAHGPU_EINVAL, "null values");
  int rc = check_n(c, n, 0);
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  uint64_t nn = n ? n : 1;
  uint64_t bytes = build_scratch_bytes(caps_one(n, shape_for(c, n))) + align_up(nn * 4) * 12 + align_up(2 * nn * 4) +
                   align_up(sizeof(wahgpu_counts)) + 4 * scan_scratch_bytes(2 * nn);
  rc = arena_reserve(c, bytes);
  if (rc) return rc;
  auto t0 = std::chrono::steady_clock::now();
  BuildArgs a{};
  carve_build(c, &a, caps_one(n, shape_for(c, n)), shape_for(c, n));
  set_one(&a, n, 0);
  uint32_t* d_values = take<uint32_t>(c, nn);
  uint32_t* s_keys = take<uint32_t>(c, nn);
  uint32_t* s_rids = take<uint32_t>(c, nn);
  uint32_t* chids = take<uint32_t>(c, nn);
  uint32_t* lits = take<uint32_t>(c, nn);
  uint32_t* mv = take<uint32_t>(c, nn);
  uint32_t* mc = take<uint32_t>(c, nn);
  uint32_t* ml = take<uint32_t>(c, nn);
  uint32_t* fills = take<uint32_t>(c, nn);
  uint32_t* o_keys = take<uint32_t>(c, nn);
  uint32_t* o_offsets = take<uint32_t>(c, nn);
  uint32_t* o_lengths = take<uint32_t>(c, nn);
  uint32_t* index = take<uint32_t>(c, 2 * nn);
  wahgpu_counts* d_counts = take<wahgpu_counts>(c, 1);
  ScanScratch sm = carve_scan(c, nn), sf = carve_scan(c, 2 * nn), sc = carve_scan(c, nn);
  Marks m;
  m.init(times != nullptr);
  m.mark(c->stream);  // 0
  if (n) HIPCHK(c, hipMemcpyAsync(d_values, values, n * 4, hipMemcpyHostToDevice, c->stream));
  m.mark(c->stream);  // 1 transfer to
  a.values = d_values;
  a.counts = d_counts;
  a.sorted_keys = s_keys;
  a.sorted_rids = s_rids;
  a.force_pair = true;  // the stages below take one (value, rid) pair per row
  // The passes ping-pong between (keysA, ridsA) and (keysB, ridsB); finalize
  // copies whichever holds the result into (s_keys, s_rids).
  // each segment of launches up to a host read is one ordered sequence
  {
    ORDERED(c);
    wah::launch_sort(a, c->stream);
    wah::launch_sort_finalize(a, c->stream);
    LAUNCHCHK(c);
    m.mark(c->stream);  // 2 sort
    wah::launch_chunks(s_rids, n, chids, lits, c->stream);
    LAUNCHCHK(c);
    m.mark(c->stream);  // 3 chunks + literals
    wah::launch_fill32(sm.base, sm.words, 0, c->stream);
    wah::launch_merge(sm.args, s_keys, chids, lits, n, mv, mc, ml, sm.result, c->stream);
    LAUNCHCHK(c);
  }
  uint64_t k = 0;
  rc = read_scan_result(c, sm.header, &k);  // k sizes the next launches (src/phases.cpp:890-903)
  if (rc) return rc;
  m.mark(c->stream);  // 4 merge
  {
    ORDERED(c);
    wah::launch_fills(mv, mc, k, fills, c->stream);
    LAUNCHCHK(c);
    m.mark(c->stream);  // 5 fills
    wah::launch_fill32(sf.base, sf.words, 0, c->stream);
    wah::launch_compact(sf.args, fills, ml, k, 1, index, sf.result, c->stream);
    LAUNCHCHK(c);
  }
  uint64_t L = 0;
  rc = read_scan_result(c, sf.header, &L);
  if (rc) return rc;
  m.mark(c->stream);  // 6 fuse
  {
    ORDERED(c);
    wah::launch_fill32(sc.base, sc.words, 0, c->stream);
    wah::launch_colen(sc.args, mv, fills, k, o_keys, o_offsets, o_lengths, d_counts, c->stream);
    LAUNCHCHK(c);
  }
  wahgpu_counts cc;
  uint64_t unused;
  rc = read_scan_result(c, sc.header, &unused);  // fault word of the column kernel
  if (rc) return rc;
  rc = read_counts(c, d_counts, &cc);
  if (rc) return rc;
  m.mark(c->stream);  // 7 column length
  counts->k = k;
  counts->length = L;
  counts->keycnt = cc.keycnt;
  counts->bound = cc.bound;
  counts->passes = cc.passes;
  counts->fault = 0;
  counts->mode = 0;
  if (counts->keycnt > keys_cap || counts->length > words_cap)
    return fail(c, WAHGPU_EINVAL, "output capacity too small");
  if (L) HIPCHK(c, hipMemcpyAsync(words, index, L * 4, hipMemcpyDeviceToHost, c->stream));
  if (cc.keycnt) {
    HIPCHK(c, hipMemcpyAsync(keys, o_keys, cc.keycnt * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(offsets, o_offsets, cc.keycnt * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(lengths, o_lengths, cc.keycnt * 4, hipMemcpyDeviceToHost, c->stream));
  }
  m.mark(c->stream);  // 8 transfer back
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (times) {
    times->transfer_to = m.us(0, 1);
    times->sort = m.us(1, 2);
    times->chunks = m.us(2, 3);
    times->merge = m.us(3, 4);
    times->fills = m.us(4, 5);
    times->fuse = m.us(5, 6);
    times->column_length = m.us(6, 7);
    times->transfer_back = m.us(7, 8);
    times->total = std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0)
                       .count();
  }
  return WAHGPU_OK;
}

// ---- stages ----------------------------------------------------------------

int wahgpu_sort_rids_by_value(wahgpu_ctx* c, const uint32_t* d_values, uint64_t n,
                              uint32_t rid_base, uint32_t* d_keys_out, uint32_t* d_rids_out) {
  if (!c) return WAHGPU_EINVAL;
  if (n && (!d_values || !d_keys_out || !d_rids_out))
    return fail(c, WAHGPU_EINVAL, "null buffer");
  int rc = check_n(c, n, rid_base);
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  rc = arena_reserve(c, build_scratch_bytes(caps_one(n, shape_for(c, n))) + align_up(sizeof(wahgpu_counts)));
  if (rc) return rc;
  BuildArgs a{};
  carve_build(c, &a, caps_one(n, shape_for(c, n)), shape_for(c, n));
  set_one(&a, n, rid_base);
  a.values = d_values;
  a.counts = take<wahgpu_counts>(c, 1);
  a.sorted_keys = d_keys_out;
  a.sorted_rids = d_rids_out;
  a.force_pair = true;
  ORDERED(c);
  wah::launch_sort(a, c->stream);
  wah::launch_sort_finalize(a, c->stream);
  LAUNCHCHK(c);
  return WAHGPU_OK;
}

int wahgpu_produce_chunk_id_literals(wahgpu_ctx* c, const uint32_t* d_rids, uint64_t n,
                                     uint32_t* d_chids, uint32_t* d_lits) {
  if (!c) return WAHGPU_EINVAL;
  if (n && (!d_rids || !d_chids || !d_lits)) return fail(c, WAHGPU_EINVAL, "null buffer");
  HIPCHK(c, hipSetDevice(c->device));
  wah::launch_chunks(d_rids, n, d_chids, d_lits, c->stream);
  LAUNCHCHK(c);
  return WAHGPU_OK;
}

int wahgpu_merge_lit_by_val_chids(wahgpu_ctx* c, const uint32_t* d_values,
                                  const uint32_t* d_chids, const uint32_t* d_lits, uint64_t n,
                                  uint32_t* d_values_out, uint32_t* d_chids_out,
                                  uint32_t* d_lits_out, uint64_t* k) {
  if (!c || !k) return WAHGPU_EINVAL;
  if (n && (!d_values || !d_chids || !d_lits || !d_values_out || !d_chids_out || !d_lits_out))
    return fail(c, WAHGPU_EINVAL, "null buffer");
  int rc = check_n(c, n, 0);
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  rc = arena_reserve(c, scan_scratch_bytes(n));
  if (rc) return rc;
  ScanScratch s = carve_scan(c, n);
  ORDERED(c);
  wah::launch_fill32(s.base, s.words, 0, c->stream);
  wah::launch_merge(s.args, d_values, d_chids, d_lits, n, d_values_out, d_chids_out, d_lits_out,
                    s.result, c->stream);
  LAUNCHCHK(c);
  order_guard_.end();  // no host wait holds the device order
  return read_scan_result(c, s.header, k);
}

int wahgpu_produce_fills(wahgpu_ctx* c, const uint32_t* d_values, const uint32_t* d_chids,
                         uint64_t k, uint32_t* d_fills) {
  if (!c) return WAHGPU_EINVAL;
  if (k && (!d_values || !d_chids || !d_fills)) return fail(c, WAHGPU_EINVAL, "null buffer");
  HIPCHK(c, hipSetDevice(c->device));
  wah::launch_fills(d_values, d_chids, k, d_fills, c->stream);
  LAUNCHCHK(c);
  return WAHGPU_OK;
}

int wahgpu_fuse_fill_literals(wahgpu_ctx* c, const uint32_t* d_fills, const uint32_t* d_lits,
                              uint64_t k, uint32_t* d_index, uint64_t* index_length) {
  if (!c || !index_length) return WAHGPU_EINVAL;
  if (k && (!d_fills || !d_lits || !d_index)) return fail(c, WAHGPU_EINVAL, "null buffer");
  int rc = check_n(c, k, 0);
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  rc = arena_reserve(c, scan_scratch_bytes(k));
  if (rc) return rc;
  ScanScratch s = carve_scan(c, k);
  ORDERED(c);
  wah::launch_fill32(s.base, s.words, 0, c->stream);
  wah::launch_compact(s.args, d_fills, d_lits, k, 1, d_index, s.result, c->stream);
  LAUNCHCHK(c);
  order_guard_.end();  // no host wait holds the device order
  return read_scan_result(c, s.header, index_length);
}

int wahgpu_compute_column_length(wahgpu_ctx* c, const uint32_t* d_values,
                                 const uint32_t* d_fills, uint64_t k, uint32_t* d_keys,
                                 uint32_t* d_offsets, uint32_t* d_lengths, uint64_t* keycnt) {
  if (!c || !keycnt) return WAHGPU_EINVAL;
  if (k && (!d_values || !d_fills || !d_keys || !d_offsets || !d_lengths))
    return fail(c, WAHGPU_EINVAL, "null buffer");
  int rc = check_n(c, k, 0);
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  rc = arena_reserve(c, scan_scratch_bytes(k) + align_up(sizeof(wahgpu_counts)));
  if (rc) return rc;
  ScanScratch s = carve_scan(c, k);
  wahgpu_counts* d_counts = take<wahgpu_counts>(c, 1);
  wah::launch_fill32(s.base, s.words, 0, c->stream);
  wah::launch_fill32(reinterpret_cast<uint32_t*>(d_counts), sizeof(wahgpu_counts) / 4, 0,
                     c->stream);
  ORDERED(c);
  wah::launch_colen(s.args, d_values, d_fills, k, d_keys, d_offsets, d_lengths, d_counts,
                    c->stream);
  LAUNCHCHK(c);
  wahgpu_counts cc;
  uint64_t unused;
  order_guard_.end();  // no host wait holds the device order
  rc = read_scan_result(c, s.header, &unused);
  if (rc) return rc;
  rc = read_counts(c, d_counts, &cc);
  if (rc) return rc;
  *keycnt = cc.keycnt;
  return WAHGPU_OK;
}

int wahgpu_stream_compaction(wahgpu_ctx* c, const uint32_t* d_data, const uint32_t* d_valid,
                             uint64_t n, uint32_t* d_out, uint64_t* count) {
  if (!c || !count) return WAHGPU_EINVAL;
  if (n && (!d_data || !d_valid || !d_out)) return fail(c, WAHGPU_EINVAL, "null buffer");
  int rc = check_n(c, n, 0);
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  rc = arena_reserve(c, scan_scratch_bytes(n));
  if (rc) return rc;
  ScanScratch s = carve_scan(c, n);
  ORDERED(c);
  wah::launch_fill32(s.base, s.words, 0, c->stream);
  wah::launch_compact(s.args, d_data, d_valid, n, 0, d_out, s.result, c->stream);
  LAUNCHCHK(c);
  order_guard_.end();  // no host wait holds the device order
  return read_scan_result(c, s.header, count);
}

int wahgpu_exclusive_scan(wahgpu_ctx* c, const uint32_t* d_in, uint64_t n, uint32_t* d_out) {
  if (!c) return WAHGPU_EINVAL;
  if (n && (!d_in || !d_out)) return fail(c, WAHGPU_EINVAL, "null buffer");
  int rc = check_n(c, n, 0);
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  rc = arena_reserve(c, scan_scratch_bytes(n));
  if (rc) return rc;
  ScanScratch s = carve_scan(c, n);
  ORDERED(c);
  wah::launch_fill32(s.base, s.words, 0, c->stream);
  wah::launch_scan(s.args, d_in, nullptr, n, d_out, false, c->stream);
  LAUNCHCHK(c);
  return WAHGPU_OK;
}

int wahgpu_segmented_inclusive_scan(wahgpu_ctx* c, const uint32_t* d_in, const uint32_t* d_heads,
                                    uint64_t n, uint32_t* d_out) {
  if (!c) return WAHGPU_EINVAL;
  if (n && (!d_in || !d_heads || !d_out)) return fail(c, WAHGPU_EINVAL, "null buffer");
  int rc = check_n(c, n, 0);
  if (rc) return rc;
  HIPCHK(c, hipSetDevice(c->device));
  rc = arena_reserve(c, scan_scratch_bytes(n));
  if (rc) return rc;
  ScanScratch s = carve_scan(c, n);
  ORDERED(c);
  wah::launch_fill32(s.base, s.words, 0, c->stream);
  wah::launch_scan(s.args, d_in, d_heads, n, d_out, true, c->stream);
  LAUNCHCHK(c);
  return WAHGPU_OK;
}

// ---- internal: batched build on a caller's workspace (the server) -------

void wah_batch_acc_add(wah_batch_acc* acc, uint64_t n) {
  uint32_t min_tile = wah::radix_tile(0);
  for (int sh = 1; sh < wah::kRadixShapes; ++sh) min_tile = std::min(min_tile, wah::radix_tile(sh));
  acc->values += (n + 63) / 64 * 64;
  acc->cols += 1;
  acc->sort_tiles += wah::sort_tiles_of(n, min_tile);
  acc->tail_tiles += wah::tail_tiles_of(n);
  acc->hist_chunks += wah::hist_chunks_of(n);
}

uint64_t wah_batch_acc_scratch(const wah_batch_acc* acc) {
  // one descriptor region per compiled tile shape: the most sub-batches
  const Caps k{acc->values, acc->cols, (uint32_t)acc->sort_tiles, (uint32_t)acc->tail_tiles,
               (uint32_t)acc->hist_chunks};
  return build_scratch_bytes(k, wah::kRadixShapes);
}

int wah_build_batch_ws(wahgpu_ctx* c, const uint32_t* d_values, uint64_t values_len,
                       const wahgpu_column* cols, uint32_t ncols, uint32_t* d_keys,
                       uint32_t* d_offsets, uint32_t* d_lengths, uint32_t* d_words,
                       wahgpu_counts* d_counts, void* d_scratch, uint64_t scratch_bytes) {
  if (!c) return WAHGPU_EINVAL;
  if (ncols == 0) return WAHGPU_OK;
  if (!cols || !d_counts || !d_scratch) return fail(c, WAHGPU_EINVAL, "null argument");
  HIPCHK(c, hipSetDevice(c->device));
  return batch_device_impl(c, d_values, values_len, cols, ncols, d_keys, d_offsets, d_lengths,
                           d_words, d_counts, 0, static_cast<char*>(d_scratch), scratch_bytes);
}

// ---- query side: decode, range query, encode, combine --------------------

namespace {
constexpr uint64_t kMaxRows = 1ull << 32;  // rows are uint32 (src/cpu.cpp:182)
uint64_t bitmap_words(uint64_t nrows) { return (nrows + 31) / 32; }
}  // namespace

int wahgpu_decode_device(wahgpu_ctx* c, const uint32_t* d_words, uint64_t len, uint64_t nrows,
                         uint32_t* d_bitmap, uint64_t* last_row) {
  if (!c) return WAHGPU_EINVAL;
  if ((len && !d_words) || (nrows && !d_bitmap)) return fail(c, WAHGPU_EINVAL, "null buffer");
  if (nrows > kMaxRows || len > ((uint64_t)1 << 32))
    return fail(c, WAHGPU_ERANGE, "more than 2^32 rows or words");
  HIPCHK(c, hipSetDevice(c->device));
  int rc = arena_reserve(c, scan_scratch_bytes(len) + kAlign);
  if (rc) return rc;
  ScanScratch sc = carve_scan(c, len);
  uint64_t* d_last = take<uint64_t>(c, 2);
  ORDERED(c);
  wah::launch_fill32(sc.base, sc.words, 0, c->stream);
  wah::launch_fill32(reinterpret_cast<uint32_t*>(d_last), 4, 0, c->stream);
  if (nrows) wah::launch_fill32(d_bitmap, bitmap_words(nrows), 0, c->stream);
  wah::launch_decode(sc.args, d_words, nullptr, len, d_bitmap, nrows, d_last, c->stream);
  LAUNCHCHK(c);
  uint64_t unused;
  order_guard_.end();  // no host wait holds the device order
  rc = read_scan_result(c, sc.header, &unused);  // fault word (synchronizes)
  if (rc) return rc;
  if (last_row) {
    HIPCHK(c, hipMemcpyAsync(c->pinned_counts, d_last, 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    memcpy(last_row, c->pinned_counts, 8);
  }
  return WAHGPU_OK;
}

int wahgpu_query_range_device(wahgpu_ctx* c, const uint32_t* d_keys, const uint32_t* d_offsets,
                              const uint32_t* d_lengths, uint64_t keycnt,
                              const uint32_t* d_words, uint32_t lo, uint32_t hi, uint64_t nrows,
                              uint32_t* d_bitmap, uint64_t* keys_matched) {
  if (!c) return WAHGPU_EINVAL;
  if ((keycnt && (!d_keys || !d_offsets || !d_lengths || !d_words)) || (nrows && !d_bitmap))
    return fail(c, WAHGPU_EINVAL, "null buffer");
  if (nrows > kMaxRows) return fail(c, WAHGPU_ERANGE, "more than 2^32 rows");
  HIPCHK(c, hipSetDevice(c->device));
  if (nrows) wah::launch_fill32(d_bitmap, bitmap_words(nrows), 0, c->stream);
  if (keys_matched) *keys_matched = 0;
  if (!keycnt || lo > hi) return WAHGPU_OK;
  int rc = arena_reserve(c, align_up(4 * 8));
  if (rc) return rc;
  uint64_t* d_range = take<uint64_t>(c, 4);
  wah::launch_key_range(d_keys, d_offsets, d_lengths, keycnt, lo, hi, d_range, c->stream);
  LAUNCHCHK(c);
  uint64_t range[4];
  HIPCHK(c, hipMemcpyAsync(c->pinned_counts, d_range, 32, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  memcpy(range, c->pinned_counts, 32);
  if (keys_matched) *keys_matched = range[1] - range[0];
  const uint64_t w0 = range[2], len = range[3] - range[2];
  if (range[1] <= range[0] || len == 0) return WAHGPU_OK;
  // the selected keys' words are contiguous: decode them as one stream whose
  // segments restart at each key's first word
  rc = arena_reserve(c, scan_scratch_bytes(len) + align_up(len * 4));
  if (rc) return rc;
  ScanScratch sc = carve_scan(c, len);
  uint32_t* heads = take<uint32_t>(c, len);
  ORDERED(c);
  wah::launch_fill32(sc.base, sc.words, 0, c->stream);
  wah::launch_fill32(heads, len, 0, c->stream);
  wah::launch_mark_heads(d_offsets, range[0], range[1], w0, heads, c->stream);
  wah::launch_decode(sc.args, d_words + w0, heads, len, d_bitmap, nrows, nullptr, c->stream);
  LAUNCHCHK(c);
  uint64_t unused;
  order_guard_.end();  // no host wait holds the device order
  return read_scan_result(c, sc.header, &unused);
}

int wahgpu_encode_device(wahgpu_ctx* c, const uint32_t* d_bitmap, uint64_t nrows,
                         uint32_t* d_words, uint64_t cap, uint64_t* len) {
  if (!c || !len) return WAHGPU_EINVAL;
  if (nrows && (!d_bitmap || (cap && !d_words))) return fail(c, WAHGPU_EINVAL, "null buffer");
  if (nrows > kMaxRows) return fail(c, WAHGPU_ERANGE, "more than 2^32 rows");
  HIPCHK(c, hipSetDevice(c->device));
  *len = 0;
  if (!nrows) return WAHGPU_OK;
  const uint64_t tiles = wah::encode_tiles(nrows);
  int rc = arena_reserve(c, align_up(tiles * 24) + kAlign);
  if (rc) return rc;
  uint64_t* tile_sum = take<uint64_t>(c, 3 * tiles);
  uint64_t* d_len = take<uint64_t>(c, 1);
  wah::launch_encode(d_bitmap, nullptr, 4, nrows, tile_sum, d_words, cap, d_len, c->stream);
  LAUNCHCHK(c);
  HIPCHK(c, hipMemcpyAsync(c->pinned_counts, d_len, 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  memcpy(len, c->pinned_counts, 8);
  if (*len > cap) return fail(c, WAHGPU_EINVAL, "output capacity too small: " +
                                                    std::to_string(*len) + " words");
  return WAHGPU_OK;
}

int wahgpu_combine_device(wahgpu_ctx* c, int op, const uint32_t* d_a, uint64_t la,
                          const uint32_t* d_b, uint64_t lb, uint64_t nrows, uint32_t* d_out,
                          uint64_t cap, uint64_t* len) {
  if (!c || !len) return WAHGPU_EINVAL;
  if (op < WAHGPU_OP_AND || op > WAHGPU_OP_ANDNOT) return fail(c, WAHGPU_EINVAL, "unknown op");
  if ((la && !d_a) || (lb && !d_b) || (nrows && cap && !d_out))
    return fail(c, WAHGPU_EINVAL, "null buffer");
  if (nrows > kMaxRows || la > ((uint64_t)1 << 32) || lb > ((uint64_t)1 << 32))
    return fail(c, WAHGPU_ERANGE, "more than 2^32 rows or words");
  HIPCHK(c, hipSetDevice(c->device));
  *len = 0;
  if (!nrows) return WAHGPU_OK;
  const uint64_t bw = bitmap_words(nrows), tiles = wah::encode_tiles(nrows);
  int rc = arena_reserve(c, 2 * align_up(bw * 4) + scan_scratch_bytes(la) +
                                scan_scratch_bytes(lb) + align_up(tiles * 24) + kAlign);
  if (rc) return rc;
  uint32_t* bm_a = take<uint32_t>(c, bw);
  uint32_t* bm_b = take<uint32_t>(c, bw);
  ScanScratch sa = carve_scan(c, la), sb = carve_scan(c, lb);
  uint64_t* tile_sum = take<uint64_t>(c, 3 * tiles);
  uint64_t* d_len = take<uint64_t>(c, 1);
  ORDERED(c);
  wah::launch_fill32(sa.base, sa.words, 0, c->stream);
  wah::launch_fill32(sb.base, sb.words, 0, c->stream);
  wah::launch_fill32(bm_a, bw, 0, c->stream);
  wah::launch_fill32(bm_b, bw, 0, c->stream);
  wah::launch_decode(sa.args, d_a, nullptr, la, bm_a, nrows, nullptr, c->stream);
  wah::launch_decode(sb.args, d_b, nullptr, lb, bm_b, nrows, nullptr, c->stream);
  wah::launch_encode(bm_a, bm_b, op, nrows, tile_sum, d_out, cap, d_len, c->stream);
  LAUNCHCHK(c);
  uint64_t unused;
  order_guard_.end();  // no host wait holds the device order
  rc = read_scan_result(c, sa.header, &unused);
  if (rc) return rc;
  rc = read_scan_result(c, sb.header, &unused);
  if (rc) return rc;
  HIPCHK(c, hipMemcpyAsync(c->pinned_counts, d_len, 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  memcpy(len, c->pinned_counts, 8);
  if (*len > cap) return fail(c, WAHGPU_EINVAL, "output capacity too small: " +
                                                    std::to_string(*len) + " words");
  return WAHGPU_OK;
}

}  // extern "C"
