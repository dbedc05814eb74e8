// wah_kernels.h -- device-state layout and launch entry points shared by the
// kernels (wah_kernels.hip) and the C ABI (wahgpu_api.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "wahgpu.h"

typedef wahgpu_counts wahgpu_counts_dev;

namespace wah {

constexpr int kSortItems = 16;                  // keys per thread in a radix pass
constexpr int kSortTile = 256 * kSortItems;     // 4096 keys per tile
constexpr int kTailItems = 8;
constexpr int kTailTile = 256 * kTailItems;     // 2048 sorted pairs per tail tile
constexpr int kScanItems = 8;
constexpr int kScanTile = 256 * kScanItems;
constexpr uint32_t kHistMaxBlocks = 1024;
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

struct SortState {
  uint32_t digit_offset[4][256];  // exclusive prefix of the digit histogram
  PassInfo pass[4];
  uint32_t num_active;
  uint32_t final_buf;             // buffer holding the sorted pairs
  uint32_t maxval;
  uint32_t tile_counter[kCounters];
};

struct BuildArgs {
  const uint32_t* values;
  uint64_t n;
  uint32_t rid_base;
  uint32_t *keysA, *ridsA, *keysB, *ridsB;
  SortState* st;
  uint32_t* part_hist;     // kHistMaxBlocks x 1024
  uint32_t* part_max;      // kHistMaxBlocks
  uint32_t* sort_status;   // sort_status_words(n)
  uint64_t* tail_status;   // tail_status_words(n)
  uint32_t *out_keys, *out_offsets, *out_lengths, *out_words;
  uint32_t *sorted_keys, *sorted_rids;  // stage API sort output
  wahgpu_counts_dev* counts;
  uint32_t max_resident;   // persistent grid size
};

struct ScanArgs {
  uint64_t* status;        // scan_status_words(n), zeroed
  uint32_t* tile_counter;  // zeroed
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
