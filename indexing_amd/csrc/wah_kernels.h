// wah_kernels.h -- device-state layout and launch entry points shared by the
// kernels (wah_kernels.hip) and the C ABI (wahgpu_api.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "wahgpu.h"

typedef wahgpu_counts wahgpu_counts_dev;

namespace wah {

// Tuning knobs (compile-time; defaults are the measured best on MI355X).
#ifndef WAH_SORT_THREADS
#define WAH_SORT_THREADS 512
#endif
#ifndef WAH_SORT_ITEMS
#define WAH_SORT_ITEMS 16
#endif
#ifndef WAH_TAIL_THREADS
#define WAH_TAIL_THREADS 512
#endif
#ifndef WAH_TAIL_ITEMS
#define WAH_TAIL_ITEMS 8
#endif
#ifndef WAH_SORT_WAVES_PER_EU
#define WAH_SORT_WAVES_PER_EU 4
#endif

constexpr int kSortThreads = WAH_SORT_THREADS;  // threads of a radix-pass block
constexpr int kSortItems = WAH_SORT_ITEMS;      // keys per thread in a radix pass
constexpr int kSortTile = kSortThreads * kSortItems;
constexpr int kTailThreads = WAH_TAIL_THREADS;
constexpr int kTailItems = WAH_TAIL_ITEMS;
constexpr int kTailTile = kTailThreads * kTailItems;  // sorted pairs per tail tile
constexpr int kScanItems = 8;
constexpr int kScanTile = 256 * kScanItems;
constexpr uint32_t kHistMaxBlocks = 2048;
constexpr int kCounters = 8;
constexpr int kCounterTail = 4;

// Which buffer a pass reads / writes.
constexpr uint32_t kBufValues = 0;  // the input column, rids implicit (rid_base + i)
constexpr uint32_t kBufA = 1;
constexpr uint32_t kBufB = 2;

struct PassInfo {
  uint32_t active;      // digit not constant over the column
  uint32_t shift;       // 8 * digit position
  uint32_t src, dst;    // kBuf*
  uint32_t status_sel;  // which of the two status arrays this pass uses
  uint32_t ordinal;     // index among active passes
};

constexpr uint32_t kHistCopies = 32;        // striped copies of the global histogram
constexpr uint32_t kHistCopyStride = 1088;  // words between copies (4 KiB + 256 B)
constexpr uint32_t kLaneWords = 8 * 64;     // tile-counter lanes of one kernel

struct SortState {
  // --- zeroed before every build (kSortHeaderWords words) ---
  uint32_t hist[kHistCopies][kHistCopyStride];  // [copy][digit * 256 + bin]
  uint32_t tiles[kCounters][kLaneWords];        // tile-counter lanes per kernel
  uint32_t maxval;
  // --- written by k_plan ---
  uint32_t digit_offset[4][256];  // exclusive prefix of the digit histogram
  PassInfo pass[4];
  uint32_t num_active;
  uint32_t final_buf;             // buffer holding the sorted pairs
};
constexpr uint64_t kSortHeaderWords =
    (uint64_t)kHistCopies * kHistCopyStride + (uint64_t)kCounters * kLaneWords + 1;

struct BuildArgs {
  const uint32_t* values;
  uint64_t n;
  uint32_t rid_base;
  uint32_t *keysA, *ridsA, *keysB, *ridsB;
  SortState* st;
  uint32_t* sort_status;   // sort_status_words(n)
  uint64_t* tail_status;   // tail_status_words(n)
  uint32_t *out_keys, *out_offsets, *out_lengths, *out_words;
  uint32_t *sorted_keys, *sorted_rids;  // stage API sort output
  wahgpu_counts_dev* counts;
  uint32_t max_resident;   // resident 256-thread blocks on the device
};

struct ScanArgs {
  uint64_t* status;        // scan_status_words(n), zeroed
  uint32_t* tile_counter;  // kLaneWords tile-counter lanes, zeroed
  uint32_t* fault;         // zeroed; nonzero after a failed wait / bad index
  uint32_t max_resident;
};

uint64_t sort_status_words(uint64_t n);
uint64_t tail_status_words(uint64_t n);
uint64_t scan_status_words(uint64_t n);

// Optional hook called between kernel groups of the fused build (timing).
struct LaunchHook {
  virtual void mark(hipStream_t s) = 0;
};
void launch_sort(const BuildArgs& a, hipStream_t s, LaunchHook* hook = nullptr);
void launch_tail(const BuildArgs& a, hipStream_t s, LaunchHook* hook = nullptr);
void launch_sort_finalize(const BuildArgs& a, hipStream_t s);
void launch_fill32(uint32_t* p, uint64_t words, uint32_t v, hipStream_t s);
void launch_chunks(const uint32_t* rids, uint64_t n, uint32_t* chids, uint32_t* lits,
                   hipStream_t s);
void launch_fills(const uint32_t* vals, const uint32_t* chids, uint64_t k, uint32_t* fills,
                  hipStream_t s);
void launch_merge(const ScanArgs& a, const uint32_t* vals, const uint32_t* chids,
                  const uint32_t* lits, uint64_t n, uint32_t* ov, uint32_t* oc, uint32_t* ol,
                  uint64_t* result, hipStream_t s);
void launch_compact(const ScanArgs& a, const uint32_t* data, const uint32_t* valid, uint64_t n,
                    int mode, uint32_t* out, uint64_t* result, hipStream_t s);
void launch_colen(const ScanArgs& a, const uint32_t* vals, const uint32_t* fills, uint64_t k,
                  uint32_t* keys, uint32_t* offsets, uint32_t* lengths,
                  wahgpu_counts_dev* counts, hipStream_t s);
void launch_scan(const ScanArgs& a, const uint32_t* in, const uint32_t* heads, uint64_t n,
                 uint32_t* out, bool segmented, hipStream_t s);

}  // namespace wah
