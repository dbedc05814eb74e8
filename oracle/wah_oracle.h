/*
 * wah_oracle.h -- CPU oracle for the WAH bitmap-index build.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the serial
 * reference program `src/cpu.cpp` (josephnoir/indexing).  It exists to check
 * the HIP implementation (libwahgpu) and to time the CPU baseline.  Only the
 * test suite, the smoke check and the CPU-baseline leg of the benchmark may
 * load it; no product code path calls it.
 *
 * Every function names the reference lines it restates.  All arithmetic is
 * uint32 (as in the reference's `uint32_t` vectors).
 *
 * Parity pins (what fixes each function independently of this source):
 *   - sort: bubble mode is the literal O(n^2) loop; packed mode sorts the
 *     64-bit words (value << 32 | rid), whose order equals a stable sort by
 *     value; tests check bubble == packed on random inputs.
 *   - whole build: hand-derived golden vectors (built-in 32-value sample and a
 *     fill example), a brute-force decode check (bit r of key v set iff
 *     values[r] == v), and the structural invariants of the encoding, which
 *     together determine the word sequence uniquely.
 *   - scans / compaction / reduce-by-key: brute-force definitions on tiny
 *     inputs and closed forms (sum of lengths == L, popcounts == occurrences).
 */
#ifndef WAH_ORACLE_H
#define WAH_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { WAH_ORACLE_SORT_BUBBLE = 0, WAH_ORACLE_SORT_PACKED = 1 };

/* src/cpu.cpp:180-197 (bubble) == src/phases.cpp:143-160 (stable_sort).
 * rids[i] = rid_base + i, then both arrays sorted by input, ties by rid. */
void wah_oracle_sort_rids_by_value(uint32_t* input, uint32_t* rids, size_t n,
                                   uint32_t rid_base, int mode);

/* src/cpu.cpp:201-211 */
void wah_oracle_produce_chunk_id_literals(const uint32_t* rids, uint32_t* chids,
                                          uint32_t* lits, size_t n);

/* src/cpu.cpp:78-116 (reduce_by_key, OR) via :215-219.  In place; returns k. */
size_t wah_oracle_merge_lit_by_val_chids(uint32_t* input, uint32_t* chids,
                                         uint32_t* lits, size_t n);

/* src/cpu.cpp:223-241.  In place: chids[0..k) become fills. */
void wah_oracle_produce_fills(const uint32_t* input, uint32_t* chids, size_t k);

/* src/cpu.cpp:245-257 with stream_compaction :170-174.  index needs 2k
 * words; returns index_length L. */
size_t wah_oracle_fuse_fill_literals(const uint32_t* fills, const uint32_t* lits,
                                     uint32_t* index, size_t k);

/* src/cpu.cpp:261-272 (reduce_by_key SUM :119-148, exclusive_scan :160-168).
 * In place: input[0..K) become the keys.  lengths[0..K) receive the per-key
 * word counts (the reference's `tmp` after the reduction) and offsets[0..K)
 * their exclusive scan.  Returns K. */
size_t wah_oracle_compute_column_length(uint32_t* input, const uint32_t* fills,
                                        uint32_t* offsets, uint32_t* lengths,
                                        size_t k);

/* src/cpu.cpp:170-174: remove every element equal to val; returns new size. */
size_t wah_oracle_stream_compaction(uint32_t* data, size_t n, uint32_t val);

/* src/sc.cpp:67-78 (compact by flags; the GPU kernel
 * include/stream_compaction.cl:124-160 keeps data[i] where valid[i] != 0). */
size_t wah_oracle_compact_valid(const uint32_t* data, const uint32_t* valid,
                                size_t n, uint32_t* out);

/* src/cpu.cpp:160-168 (n may be 0 here; the reference requires n >= 1). */
void wah_oracle_exclusive_scan(const uint32_t* in, uint32_t* out, size_t n);

/* src/segmented_scan.cpp:109-121 (heads[i] != 0 starts a segment; element 0
 * always starts one). */
void wah_oracle_segmented_inclusive_scan(const uint32_t* in, const uint32_t* heads,
                                         uint32_t* out, size_t n);

typedef struct wah_oracle_counts {
  uint64_t k;       /* merged (value, chunk) entries: "Merged to k values"   */
  uint64_t length;  /* index_length L                                        */
  uint64_t keycnt;  /* K distinct keys                                       */
} wah_oracle_counts;

/* The whole pipeline of src/cpu.cpp:313-328.  keys/offsets/lengths need n
 * entries, words 2n.  rid_base = 0 is the reference; other bases index the
 * column as rows [rid_base, rid_base + n).  Returns 0, or -1 on allocation
 * failure.  n = 0 gives an empty index (the reference is undefined there). */
int wah_oracle_build(const uint32_t* values, size_t n, uint32_t rid_base,
                     int sort_mode, uint32_t* keys, uint32_t* offsets,
                     uint32_t* lengths, uint32_t* words, wah_oracle_counts* out);

/* The same, with the seconds of each stage (CLOCK_MONOTONIC) written to
 * stage_s[0..5]: sort, chunk ids + literals, merge, fills, fuse, column
 * length (the buckets the reference's `phases` prints, src/phases.cpp:714-1127).
 * stage_s may be NULL. */
#define WAH_ORACLE_STAGES 6
int wah_oracle_build_timed(const uint32_t* values, size_t n, uint32_t rid_base,
                           int sort_mode, uint32_t* keys, uint32_t* offsets,
                           uint32_t* lengths, uint32_t* words, wah_oracle_counts* out,
                           double* stage_s);

/* The same build with the sort on `threads` threads (the packed words cut
 * into slices, sorted, then merged by rank ranges: the same order; CPU
 * baseline only, the reference is serial).  threads <= 1 is
 * wah_oracle_build_timed with the packed sort. */
int wah_oracle_build_threads(const uint32_t* values, size_t n, uint32_t rid_base,
                             int threads, uint32_t* keys, uint32_t* offsets,
                             uint32_t* lengths, uint32_t* words, wah_oracle_counts* out,
                             double* stage_s);

/* src/cpu.cpp:47-74: '0'/'1' text of the bitmap, trailing '0's trimmed.
 * Writes at most cap bytes (no terminator) and returns the full length. */
size_t wah_oracle_decoded_bitmap(const uint32_t* words, size_t len, char* out,
                                 size_t cap);

#ifdef __cplusplus
}
#endif
#endif
