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
#define WAH_SORT_THREADS 1024
#endif
#ifndef WAH_SORT_ITEMS
#define WAH_SORT_ITEMS 16
#endif
#ifndef WAH_TAIL_THREADS
#define WAH_TAIL_THREADS 1024
#endif
#ifndef WAH_TAIL_ITEMS
#define WAH_TAIL_ITEMS 8
#endif
#ifndef WAH_SORT_WAVES_PER_EU
#define WAH_SORT_WAVES_PER_EU 4
#endif
// Consecutive tiles drawn by one tile-counter lane (one XCD) in the radix
// passes and the tail (see next_tile).
#ifndef WAH_SORT_TILE_GROUP
#define WAH_SORT_TILE_GROUP 8
#endif
#ifndef WAH_TAIL_TILE_GROUP
#define WAH_TAIL_TILE_GROUP 4
#endif

constexpr int kSortThreads = WAH_SORT_THREADS;  // threads of a radix-pass block
constexpr int kSortItems = WAH_SORT_ITEMS;      // keys per thread in a radix pass
constexpr int kSortTile = kSortThreads * kSortItems;
// Radix tile shapes compiled in: threads per block x keys per thread (shape
// 0 is the default).  A build runs one shape per column, chosen per
// column-length bucket by the tuning table.  Fewer keys per thread make more,
// smaller tiles (more blocks in flight, more look-back steps); more keys per
// thread the reverse.  Every tile is a whole number of the histogram's
// 4096-row chunks (and of the run cut's 4096-row steps), which count and cut
// the runs of one radix tile at a time: so no 256 x 8 or 256 x 24, and no
// 1024 x 24 (its 96 KiB of staged keys and 16 waves leave too few registers).
constexpr int kRadixShapes = 6;
__host__ __device__ constexpr int radix_threads(int shape) {
  return shape == 1 || shape == 4 || shape == 5 ? 512
         : shape == 2                           ? 256
         : shape == 3                           ? 1024
                                                : kSortThreads;
}
__host__ __device__ constexpr int radix_items(int shape) {
  return shape == 3 || shape == 4 ? 8 : shape == 5 ? 24 : shape == 0 ? kSortItems : 16;
}
__host__ __device__ constexpr uint32_t radix_tile(int shape) {
  return (uint32_t)(radix_threads(shape) * radix_items(shape));
}
// Waves per SIMD a radix instantiation is compiled for: 8 keys per thread
// fit in 80 registers (6 waves), 16 in 128 (4), 24 in 256 (2).
__host__ __device__ constexpr int sort_waves_per_eu(int items) {
  return items <= 8 ? 6 : items >= 24 ? 2 : WAH_SORT_WAVES_PER_EU;
}
constexpr int kTailThreads = WAH_TAIL_THREADS;
constexpr int kTailItems = WAH_TAIL_ITEMS;
constexpr int kTailTile = kTailThreads * kTailItems;  // sorted pairs per tail tile
// Tail occupancy: 8 waves per SIMD (64 VGPRs), so two 1024-thread blocks
// (2 x 64 KiB of staged words) share a CU and one loads while the other
// scans, looks back or stores.
#ifndef WAH_TAIL_LANES  // tile counters of the tail (0: as many as the grid allows, up to 8)
#define WAH_TAIL_LANES 0
#endif
// Tail look-back window: 64 x this many tiles per round trip.  The wait is
// for the aggregates of the tiles just before (their loads finish at
// different times), not for the walk, so one slot is best (1.32 ms on 2^28
// values against 1.38 / 1.41 / 1.51 ms for 2 / 4 / 8).
#ifndef WAH_TAIL_LOOKBACK_SLOTS
#define WAH_TAIL_LOOKBACK_SLOTS 1
#endif
#ifndef WAH_TAIL_WAVES_PER_EU
#define WAH_TAIL_WAVES_PER_EU 8
#endif
constexpr int kTailWavesPerEU = WAH_TAIL_WAVES_PER_EU;
constexpr int kTailBlocksPerCU = kTailWavesPerEU * 4 * 64 / kTailThreads;
// Run-mode tail: the same tiles as 16 items per thread (512 threads for the
// 8192-item tile) at 4 waves per SIMD (128 VGPRs), two blocks per CU.
constexpr int kRunTailItems = 16;
constexpr int kRunTailThreads = kTailTile / kRunTailItems;
constexpr int kRunTailBlocksPerCU = 4 * 4 * 64 / kRunTailThreads;
static_assert(kRunTailThreads * kRunTailItems == kTailTile, "run tail tiles = pair tail tiles");
// Long single columns (at least 2^WAH_LONG_TAIL_MIN_LOG2 values) take tail
// tiles of 16384 items, 1024 threads x 16 at one block per CU (4 waves per
// SIMD), in both modes: 2^28 uniform values, tail 0.95 ms against 0.99 ms in
// 8192-item tiles.  Shorter columns and batches keep the 8192-item tiles,
// which give more CUs a tile (2^20 values: 17 us against 20 us).
#ifndef WAH_LONG_TAIL_MIN_LOG2
#define WAH_LONG_TAIL_MIN_LOG2 24
#endif
constexpr int kLongTailThreads = 1024;
constexpr int kLongTailItems = 16;
constexpr int kLongTailTile = kLongTailThreads * kLongTailItems;
constexpr int kLongTailBlocksPerCU = 4 * 4 * 64 / kLongTailThreads;
constexpr uint64_t kLongTailMin = 1ull << WAH_LONG_TAIL_MIN_LOG2;
// waves per SIMD a tail instantiation is compiled for
__host__ __device__ constexpr int tail_waves_per_eu(int items) {
  return items >= 16 ? 4 : kTailWavesPerEU;
}
constexpr int kScanItems = 8;
constexpr int kScanTile = 256 * kScanItems;
constexpr uint32_t kHistMaxBlocks = 2048;
// Tile counters: radix launches 0-3 of pair-mode columns, then their tail;
// radix launches 0-3 of run-mode columns, then their tail.
constexpr int kCounters = 14;
constexpr int kCounterTail = 4;
constexpr int kCounterRun = 5;
constexpr int kCounterRunTail = 9;
constexpr int kCounterCut = 10;
constexpr int kCounterWide = 11;  // wide-digit launches 0-2

// Wide digits.  A pair-mode column whose keys take fewer passes of up to 11
// bits than of 8 (9-11, 17-22 or 25-32 significant bits, trivial 8-bit
// digits skipped) sorts with 2048-bin digits when its length bucket allows
// it: a second histogram over its values (the first one only gives 8-bit
// digits) and up to three passes with 2^bits bins each, the key's
// significant bits split evenly between them.
constexpr int kWideBits = 11;
constexpr uint32_t kWideBins = 1u << kWideBits;
constexpr int kWidePasses = 3;
constexpr int kWideThreads = 512;
constexpr int kWideItems = 16;
constexpr uint32_t kWideTile = kWideThreads * kWideItems;
constexpr uint64_t kWideMinValues = 1u << 14;  // shorter columns keep 8-bit digits

// Which buffer a pass reads / writes.
constexpr uint32_t kBufValues = 0;  // the input column, rids implicit (rid_base + i)
constexpr uint32_t kBufA = 1;
constexpr uint32_t kBufB = 2;

constexpr uint32_t kLaneWords = 8 * 64;     // tile-counter lanes of one kernel
constexpr uint32_t kHistChunk = 4096;       // values per histogram work item
constexpr bool radix_tiles_whole_chunks(int sh = 0) {
  return sh == kRadixShapes || (radix_tile(sh) % kHistChunk == 0 && radix_tiles_whole_chunks(sh + 1));
}
static_assert(radix_tiles_whole_chunks(), "radix tiles hold whole histogram chunks");

// One column of a build.  Its values are values[off, off + n) of the
// concatenated input; its sorted pairs, keys, offsets and lengths use the same
// range of their buffers, its words [2 off, 2 off + 2 n).  Tiles and histogram
// chunks never straddle two columns; the *0 fields are the column's first
// global tile / chunk (prefix sums over the preceding columns).
struct Column {
  uint64_t off;
  uint64_t n;
  uint32_t rid_base;
  uint32_t sort_tile0;
  uint32_t tail_tile0;
  uint32_t hist_chunk0;
  uint32_t index;  // slot of the column's counts in the caller's array
  uint32_t wide_slot;   // 1 + index of its wide-digit state; 0: 8-bit digits only
  uint32_t wide_tile0;  // its first wide tile (prefix over the preceding columns)
  uint32_t pad;
};
static_assert(sizeof(Column) == 48, "descriptor layout");

struct Dims {
  uint32_t ncols, sort_tiles, tail_tiles, hist_chunks;
  uint32_t wide_tiles;  // kWideTile tiles of the columns that may take wide digits
  uint32_t pad;
};
// Device descriptor: dims, padding, then ncols Columns.
struct DescHead {
  Dims dims;
  uint32_t pad[2];
};
static_assert(sizeof(DescHead) == 32, "descriptor layout");
__host__ __device__ inline const Column* desc_cols(const DescHead* d) {
  return reinterpret_cast<const Column*>(d + 1);
}
__host__ __device__ inline Column* desc_cols(DescHead* d) {
  return reinterpret_cast<Column*>(d + 1);
}

// Build-wide state, zeroed by k_begin.
struct Header {
  uint32_t tiles[kCounters][kLaneWords];  // tile-counter lanes per kernel
  uint32_t fault;                         // nonzero after a failed wait / bad index
  uint32_t max_active[2];                 // most radix passes of a column, per mode
  uint32_t mode_cols[2];                  // nonempty columns per mode
  uint32_t max_wide;                      // most wide-digit passes of a column
  uint32_t pad[58];
};

// Per-column state.  hist and maxval are zeroed by k_begin; the rest is
// written by k_plan.
// Sort modes of a column.  Pair mode sorts one (value, rid) pair per row.
// Run mode first cuts the column into runs of equal values inside one
// 31-row chunk and sorts one item per run: (value, rid of its first row,
// literal bits of its rows); columns whose rows repeat their predecessor
// (runs, low cardinality in order) shrink to a fraction of n items.
constexpr uint32_t kModePair = 0;
constexpr uint32_t kModeRun = 1;
// counts.mode of a pair-mode column sorted with wide digits (ColState.mode
// stays kModePair)
constexpr uint32_t kModeWideReport = 2;

struct ColState {
  uint32_t hist[4][256];          // histogram of each 8-bit digit
  uint32_t rep[4][256];           // the same over rows that repeat their predecessor
  uint32_t maxval;
  uint32_t num_active;            // radix passes this column needs
  uint32_t final_buf;             // kBuf* holding its sorted items
  uint32_t shift[4];              // digit shift of its a-th active pass
  uint32_t mode;                  // kModePair / kModeRun
  uint32_t items;                 // items sorted: n, or the runs in run mode
  uint32_t narrow;                // 1: keys < 2^16 move as 16-bit words between passes
  uint32_t wide;                  // 1: sorts with wide digits (num_active of them)
  uint32_t wide_slot;             // its WideState
  uint32_t wbits[kWidePasses];    // wide digit widths
  uint32_t wshift[kWidePasses];   // and shifts
  uint32_t pad[46];
  uint32_t digit_offset[4][256];  // exclusive prefix of each item digit histogram
};
constexpr uint32_t kColZeroWords = 8 * 256 + 1;

// Per column that may take wide digits: the wide digit histograms, turned
// into exclusive prefixes (global digit offsets) by k_wplan.
struct WideState {
  uint32_t hist[kWidePasses][kWideBins];
};

struct BuildArgs {
  const uint32_t* values;            // concatenated input
  uint32_t *keysA, *ridsA, *keysB, *ridsB;
  uint32_t *litsA, *litsB;           // run mode: literal bits of each item
  DescHead* desc;                    // device descriptor
  Header* hdr;
  ColState* cs;                      // cols_cap entries
  uint32_t* sort_status;             // 2 arrays of sort_status_stride words
  uint64_t sort_status_stride;
  uint64_t* tail_status;             // tail_tiles_cap words
  uint32_t* tile_entries;            // tail_tiles_cap words: entries (k) per tail tile
  uint32_t* tile_runs;               // sort_tiles_cap words: runs per radix tile, then
                                     // their exclusive prefix within the column
  WideState* wide;                   // wide_slots_cap entries
  uint32_t* wide_status;             // 2 arrays of wide_status_stride words
  uint64_t wide_status_stride;
  uint32_t *out_keys, *out_offsets, *out_lengths, *out_words;
  uint32_t *sorted_keys, *sorted_rids;  // stage API sort output (one column)
  wahgpu_counts_dev* counts;         // per column
  uint32_t max_resident;             // resident 256-thread blocks on the device
  int shape;                         // radix tile shape (radix_threads x radix_items)
  // Upper bounds the grids are sized for (a captured graph replays any batch
  // within them; the kernels read the actual sizes from the descriptor).
  uint32_t cols_cap, sort_tiles_cap, tail_tiles_cap, hist_chunks_cap;
  uint32_t wide_slots_cap, wide_tiles_cap;  // 0: no column may take wide digits
  uint64_t n_cap;
  // Every column sorts (value, rid) pairs (the stage API needs them).
  bool force_pair;
  // Radix ranking by LDS fetch-add (checked on the device at context start).
  bool atomic_rank;
  // Single column: k_begin writes the descriptor from these (no host copy).
  bool inline_desc;
  // The column is long: tail tiles of kLongTailTile items (one_dims counts
  // them); tail_tiles_cap still bounds the kTailTile count.
  bool long_tail;
  Dims one_dims;
  Column one;
};

// Tile / chunk counts of a column of n values.
__host__ __device__ inline uint32_t sort_tiles_of(uint64_t n, uint32_t tile = kSortTile) {
  return (uint32_t)((n + tile - 1) / tile);
}
__host__ __device__ inline uint32_t tail_tiles_of(uint64_t n, uint32_t tile = kTailTile) {
  return (uint32_t)((n + tile - 1) / tile);
}
__host__ __device__ inline uint32_t hist_chunks_of(uint64_t n) {
  return (uint32_t)((n + kHistChunk - 1) / kHistChunk);
}

struct ScanArgs {
  uint64_t* status;        // scan_status_words(n), zeroed
  uint32_t* tile_counter;  // kLaneWords tile-counter lanes, zeroed
  uint32_t* fault;         // zeroed; nonzero after a failed wait / bad index
  uint32_t max_resident;
};

uint64_t scan_status_words(uint64_t n);

// Optional hook called between kernel groups of the fused build (timing).
struct LaunchHook {
  virtual void mark(hipStream_t s) = 0;
};
// The fused build: k_begin, k_histogram, k_plan, 4 radix launches (the a-th
// launch runs each column's a-th active pass), k_tail, k_finish.
void launch_sort(const BuildArgs& a, hipStream_t s, LaunchHook* hook = nullptr);
void launch_tail(const BuildArgs& a, hipStream_t s, LaunchHook* hook = nullptr);
void launch_sort_finalize(const BuildArgs& a, hipStream_t s);
// *ok must hold 1; stays 1 if LDS fetch-adds rank lanes in lane order here.
void launch_rank_selftest(uint32_t* ok, hipStream_t s);
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

// Query side (wah_query.hip): WAH streams -> row bitmaps -> WAH streams.
void launch_decode(const ScanArgs& s, const uint32_t* words, const uint32_t* heads,
                   uint64_t len, uint32_t* bitmap, uint64_t nrows, uint64_t* last_row,
                   hipStream_t st);
void launch_mark_heads(const uint32_t* offsets, uint64_t a, uint64_t b, uint64_t w0,
                       uint32_t* heads, hipStream_t st);
void launch_key_range(const uint32_t* keys, const uint32_t* offsets, const uint32_t* lengths,
                      uint64_t keycnt, uint32_t lo, uint32_t hi, uint64_t* range, hipStream_t st);
uint64_t encode_tiles(uint64_t nrows);
void launch_encode(const uint32_t* a, const uint32_t* b, int op, uint64_t nrows,
                   uint64_t* tile_sum, uint32_t* out, uint64_t cap, uint64_t* len,
                   hipStream_t st);

}  // namespace wah
