/*
 * wahgpu.h -- C ABI of libwahgpu, the MI355X (gfx950) WAH bitmap-index build.
 *
 * This header is the drop-in boundary that replaces the CAF OpenCL module
 * (caf::opencl::manager::spawn + mem_ref<T>) of josephnoir/indexing for the
 * index-build path.  Only plain C types cross it: pointers, sizes, status
 * codes.  Every function returns WAHGPU_OK (0) or a negative WAHGPU_E* code;
 * wahgpu_last_error() gives the message of the last failure on a context.
 *
 * Semantics follow the reference's serial program src/cpu.cpp bit for bit:
 *   rid = row position; chid = rid / 31; literal = 0x80000000 | 1 << rid % 31;
 *   literals sharing (value, chid) are OR-merged; a key's first fill is its
 *   first chid, later fills are chid - chid_prev - 1; zero fills are dropped;
 *   keys ascend; offsets are the exclusive scan of per-key word counts.
 *
 * Pointers named d_* are device pointers (hipMalloc / torch CUDA tensors);
 * the others are host pointers.  Device-pointer functions are ordered on the
 * context's stream (wahgpu_stream) and return before the GPU finishes unless
 * they return counts to the host, which they say.
 *
 * Limits: one column holds at most WAHGPU_MAX_VALUES values (digit counts and
 * word positions are kept in 30-bit look-back fields); rid_base + n must not
 * exceed 2^32 (rows are uint32 as in the reference).
 */
#ifndef WAHGPU_H
#define WAHGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WAHGPU_OK 0
#define WAHGPU_EINVAL (-1)   /* bad argument (null pointer, capacity too small) */
#define WAHGPU_ENOMEM (-2)   /* device arena or host allocation failed          */
#define WAHGPU_EHIP (-3)     /* HIP runtime error                               */
#define WAHGPU_ERANGE (-4)   /* n above WAHGPU_MAX_VALUES or rids overflow      */
#define WAHGPU_ENODEV (-5)   /* no such device                                  */
#define WAHGPU_EIO (-6)      /* file could not be read or written               */
#define WAHGPU_EBUSY (-7)    /* serving queue full / request not finished       */

#define WAHGPU_MAX_VALUES ((((uint64_t)1) << 30) - 1u)

typedef struct wahgpu_ctx wahgpu_ctx;

/* Sizes the build produced.  The reference prints k ("Merged to k values",
 * src/cpu.cpp:321) and L, K ("Index has L elements with K keys", :327). */
typedef struct wahgpu_counts {
  uint64_t k;       /* merged (value, chunk) entries                 */
  uint64_t length;  /* index_length L: WAH words                     */
  uint64_t keycnt;  /* K: distinct values                            */
  uint32_t bound;   /* maximum value (0 for an empty column)          */
  uint32_t passes;  /* radix passes the sort ran (trivial digits skipped) */
  uint32_t fault;   /* nonzero: a kernel gave up waiting on a look-back
                       status or caught an out-of-range index; the
                       outputs are not valid (never expected)            */
  uint32_t mode;    /* sort mode of the fused build: 0 = one (value, rid)
                       pair per row, 1 = one item per run of equal values
                       inside a 31-row chunk (chosen when that sorts under
                       2/3 of the rows), 2 = pairs sorted with wide digits
                       (see wahgpu_tuning); staged paths report 0        */
} wahgpu_counts;

/* Microseconds per phase, the buckets `phases` prints (src/phases.cpp:714-1127). */
typedef struct wahgpu_phase_times {
  double transfer_to, sort, chunks, merge, fills, fuse, column_length,
      transfer_back, total;
} wahgpu_phase_times;

/* ---- devices and contexts (caf::opencl::manager, src/phases.cpp:373-388) */
int wahgpu_device_count(int* count);
/* Writes the device's marketing name ("AMD Instinct MI355X") into buf. */
int wahgpu_device_name(int ordinal, char* buf, size_t cap);
/* Compute units of the device (the `-j 0` default of src/gpu.cpp:341). */
int wahgpu_device_compute_units(int ordinal);
/* Ordinal of the first device whose name or gfx arch equals `name`, or -1
 * (find_device_if with a name match, src/phases.cpp:376-380). */
int wahgpu_find_device(const char* name);
/* One HIP stream plus a pooled device arena.  arena_bytes = 0 grows the arena
 * on demand; otherwise the arena is allocated once and requests beyond it
 * fail with WAHGPU_ENOMEM.  A context is used by one thread at a time.
 * Contexts and streams may run builds at the same time, but on one device the
 * library's single-pass kernels (radix passes, tail, scans, compaction,
 * decode) run one call's launch sequence at a time, in the order the calls
 * were made: each call's kernels wait (hipStreamWaitEvent) for the previous
 * call's, whatever its stream (a call that reads results back between its
 * kernels, such as the staged build, orders each part between two reads
 * separately, and no host wait holds the order).  Their look-back waits on tiles drawn by
 * workgroups placed on particular XCDs, and two such launches sharing the
 * device could each hold the CUs the other needs.  Copies and other work still
 * overlap them.  Calls made while their stream captures a graph are not
 * ordered: replay such graphs on one stream at a time. */
int wahgpu_create(wahgpu_ctx** out, int device_ordinal, uint64_t arena_bytes);
void wahgpu_destroy(wahgpu_ctx* ctx);
const char* wahgpu_last_error(const wahgpu_ctx* ctx);
void* wahgpu_stream(wahgpu_ctx* ctx);               /* the hipStream_t */
/* Orders the context's work on a caller's hipStream_t (e.g. the framework's
 * current stream); NULL returns to the context's own stream, a blocking
 * stream that is ordered with the legacy default stream (torch's default).
 * A switch makes the new stream wait (hipStreamWaitEvent) for the work
 * already queued on the old one, since both use the context's scratch; it is
 * refused (WAHGPU_EINVAL) while either stream is capturing a graph. */
int wahgpu_set_stream(wahgpu_ctx* ctx, void* stream);
int wahgpu_synchronize(wahgpu_ctx* ctx);
uint64_t wahgpu_arena_bytes(const wahgpu_ctx* ctx);
/* Radix ranking the context uses: 1 = LDS fetch-add (taken when a check at
 * wahgpu_create shows the device returns same-address fetch-adds of one wave
 * instruction in lane order, which makes the rank stable), 0 = ballot match.
 * The environment variable WAHGPU_RANK=match forces 0. */
int wahgpu_rank_mode(const wahgpu_ctx* ctx);

/* ---- whole build (the indexer's worker, src/gpu.cpp:356-367; the phases
 * pipeline, src/phases.cpp:720-1088) ---------------------------------------
 * Fused fast path on device pointers.  Caller capacities follow the
 * reference's worst-case sizing (src/gpu.cpp:347-352): d_keys, d_offsets,
 * d_lengths hold n entries, d_words 2n.  Counts are written to the device
 * struct d_counts; nothing is returned to the host (graph-capturable). */
int wahgpu_build_device(wahgpu_ctx* ctx, const uint32_t* d_values, uint64_t n,
                        uint32_t rid_base, uint32_t* d_keys, uint32_t* d_offsets,
                        uint32_t* d_lengths, uint32_t* d_words,
                        wahgpu_counts* d_counts);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): wahgpu_build_device with its scratch (radix ping-pong
 * buffers, look-back state, descriptor) taken from a static device array of
 * 112 MiB per GPU, held until the process ends, instead of the context's
 * arena.  It calls no hipMalloc, hipFree or other allocating routine; the
 * context (wahgpu_create, which looks the array up) and the caller's inputs
 * and outputs are made beforehand.  Columns of at most
 * WAHGPU_NOALLOC_MAX_VALUES values; larger ones fail with WAHGPU_ERANGE.
 * Every call on one GPU shares that GPU's array (the build clears what it
 * needs itself).  Calls from other streams or contexts are ordered like every
 * build (see wahgpu_create); calls captured into graphs are not, and graphs
 * replayed on several streams at once race on the array. */
#define WAHGPU_NOALLOC_MAX_VALUES (((uint64_t)1) << 22)
int wahgpu_build_device_noalloc(wahgpu_ctx* ctx, const uint32_t* d_values, uint64_t n,
                                uint32_t rid_base, uint32_t* d_keys, uint32_t* d_offsets,
                                uint32_t* d_lengths, uint32_t* d_words,
                                wahgpu_counts* d_counts);

/* Per-kernel timing of wahgpu_build_device calls, for measurement: between
 * wahgpu_timing_begin and wahgpu_timing_end every build_device call records
 * HIP events around its kernels on the context's stream (at most max_calls
 * calls).  timing_end waits for them and writes, per kernel group, the summed
 * milliseconds over the recorded calls:
 *   ms[0] histogram + digit scan, ms[1..4] radix passes of digits 0..3
 *   (inactive passes are near zero), ms[5] tail, ms[6] lengths.
 * *calls receives the number of recorded calls. */
#define WAHGPU_TIMING_GROUPS 7
int wahgpu_timing_begin(wahgpu_ctx* ctx, uint32_t max_calls);
int wahgpu_timing_end(wahgpu_ctx* ctx, double* ms, uint32_t* calls);

/* Host-pointer build: copies values in, runs the fused build, copies the
 * results out and returns the counts.  keys_cap bounds keys/offsets/lengths,
 * words_cap bounds words (n and 2n always suffice).  times may be NULL. */
int wahgpu_build(wahgpu_ctx* ctx, const uint32_t* values, uint64_t n,
                 uint32_t rid_base, uint32_t* keys, uint32_t* offsets,
                 uint32_t* lengths, uint64_t keys_cap, uint32_t* words,
                 uint64_t words_cap, wahgpu_counts* counts,
                 wahgpu_phase_times* times);

/* Host-pointer build through the staged kernels, one stage per phase, the
 * way `phases` runs; fills every bucket of times (may be NULL). */
int wahgpu_build_staged(wahgpu_ctx* ctx, const uint32_t* values, uint64_t n,
                        uint32_t* keys, uint32_t* offsets, uint32_t* lengths,
                        uint64_t keys_cap, uint32_t* words, uint64_t words_cap,
                        wahgpu_counts* counts, wahgpu_phase_times* times);

/* ---- batched build: many columns in one launch set ----------------------
 * The CAF indexer sends one column (slice) per worker request
 * (src/gpu.cpp:194-227); this indexes any number of columns of uneven length
 * and cardinality together: one histogram, one plan, four radix, one tail
 * and one finish launch in all.  Radix launch a runs every column's a-th
 * non-constant digit, with look-back chains per column; tiles never straddle
 * columns.  Each column's result equals wahgpu_build_device of that column
 * alone with the same rid_base. */
typedef struct wahgpu_column {
  uint64_t offset;    /* first value of the column in d_values */
  uint64_t n;         /* values, at most WAHGPU_MAX_VALUES */
  uint32_t rid_base;  /* row id of its first value */
  uint32_t reserved;  /* 0 */
} wahgpu_column;

/* Device pointers.  Columns ascend by offset and do not overlap.  Column c
 * writes keys, offsets and lengths to [offset, offset + n) of d_keys,
 * d_offsets, d_lengths, its words to [2 offset, 2 offset + 2 n) of d_words and
 * its counts to d_counts[c]; offsets are relative to the column's first word.
 * values_len is the length of d_values (d_keys, d_offsets, d_lengths hold
 * values_len entries, d_words 2 values_len).  Asynchronous on the context's
 * stream; `cols` is consumed before the call returns. */
int wahgpu_build_batch_device(wahgpu_ctx* ctx, const uint32_t* d_values,
                              uint64_t values_len, const wahgpu_column* cols,
                              uint32_t ncols, uint32_t* d_keys, uint32_t* d_offsets,
                              uint32_t* d_lengths, uint32_t* d_words,
                              wahgpu_counts* d_counts);

/* Host pointers: column c is values[c][0 .. n[c]) with row ids from
 * rid_base[c] (NULL: all 0); its outputs land in results[c] (capacities as
 * for wahgpu_build).  Synchronous. */
typedef struct wahgpu_result {
  uint32_t* keys;
  uint32_t* offsets;
  uint32_t* lengths;
  uint64_t keys_cap;
  uint32_t* words;
  uint64_t words_cap;
  wahgpu_counts counts; /* out */
} wahgpu_result;
int wahgpu_build_batch(wahgpu_ctx* ctx, const uint32_t* const* values, const uint64_t* n,
                       const uint32_t* rid_base, uint32_t ncols, wahgpu_result* results);

/* Graph replay for batched device builds.  After wahgpu_batch_reserve(ctx,
 * max_values, max_cols) the context lays its scratch out for any batch of at
 * most max_values values (values_len) and max_cols columns; the first such
 * build captures the launch set into a hipGraph, later ones with the same
 * buffers upload their descriptors and replay it.  (0, 0) turns it off. */
int wahgpu_batch_reserve(wahgpu_ctx* ctx, uint64_t max_values, uint32_t max_cols);
/* Number of graph captures and replays so far (for tests and tuning). */
int wahgpu_graph_stats(const wahgpu_ctx* ctx, uint64_t* captures, uint64_t* replays);

/* Tuning per column-length bucket: the radix tile shape and the digit width.
 * A build sorts each column with the tile shape of its length bucket
 * (columns of different shapes in one batch run as one launch set per shape):
 *   buckets:  n < 2^14, n < 2^17, n < 2^20, n < 2^24, larger;
 *   shapes:   threads per block x keys per thread of the radix passes,
 *             0 = 1024 x 16, 1 = 512 x 16, 2 = 256 x 16, 3 = 1024 x 8,
 *             4 = 512 x 8, 5 = 512 x 24 (shape_threads / shape_items; every
 *             tile a multiple of 4096 keys; the tail tile is 1024 x 8);
 *   digits:   8 = 8-bit digits (256 bins; digits every value shares are
 *             skipped), 11 = wide digits allowed: a pair-mode column of at
 *             least 2^14 values whose keys take fewer passes of up to 11 bits
 *             (keys of 9-11, 17-22 or 25-32 significant bits) sorts with
 *             those (a second histogram pass over its values, then one pass
 *             per wide digit, 512 x 16 tiles).  counts.mode reports 2 for it
 *             and counts.passes its wide passes.
 * wahgpu_autotune times, on this device, synthetic batches of each bucket
 * (columns of 2^13, 2^15.5, 2^18.5, 2^22 and 2^26 values, 2^24 values per
 * bucket and one 2^26 column for the last): every shape with 8-bit digits on
 * uniform values with 2^12 distinct (ms), then the fastest shape with 8-bit
 * and with wide digits on uniform values with 2^10 distinct (digit_ms), and
 * keeps the fastest of each per bucket in the context.  The default table is
 * shape 0 and 8-bit digits everywhere. */
#define WAHGPU_TUNE_BUCKETS 5
#define WAHGPU_TUNE_SHAPES 6
typedef struct wahgpu_tuning {
  uint32_t bucket_log2[WAHGPU_TUNE_BUCKETS]; /* bucket b: n < 2^bucket_log2[b] (64: any) */
  uint32_t shape[WAHGPU_TUNE_BUCKETS];       /* chosen shape per bucket */
  uint32_t digit_bits[WAHGPU_TUNE_BUCKETS];  /* digit width per bucket: 8, or 11 (wide allowed) */
  double ms[WAHGPU_TUNE_BUCKETS][WAHGPU_TUNE_SHAPES]; /* autotune: build time per shape */
  uint32_t shape_threads[WAHGPU_TUNE_SHAPES];  /* threads per block of each shape */
  uint32_t shape_items[WAHGPU_TUNE_SHAPES];    /* keys per thread of each shape */
  double digit_ms[WAHGPU_TUNE_BUCKETS][2];     /* autotune: 8-bit / wide digits (0: not timed) */
} wahgpu_tuning;
int wahgpu_autotune(wahgpu_ctx* ctx, wahgpu_tuning* out);
int wahgpu_get_tuning(const wahgpu_ctx* ctx, wahgpu_tuning* out);
/* Installs shape[] and digit_bits[] (0 reads as 8; other fields ignored);
 * shapes must be < WAHGPU_TUNE_SHAPES and digit widths 8 or 11. */
int wahgpu_set_tuning(wahgpu_ctx* ctx, const wahgpu_tuning* in);

/* ---- serving: continuous batching of build requests ----------------------
 * A host request loop in place of the indexer -> worker round trips
 * (src/gpu.cpp:194-261): requests (one column each, host memory) queue as
 * they arrive.  The server allocates one device arena of arena_bytes at
 * creation and carves it into per-batch blocks.  `slots` worker threads, each
 * with its own stream, take the queued requests, in arrival order, whose
 * combined device footprint (values, keys, offsets, lengths, words and the
 * batched build's scratch: about 49 bytes per value plus ~37 KiB per column)
 * fits the largest free block, within max_batch_cols columns and
 * max_batch_values values, and build them as one batched build in that block.
 * When the first queued request does not fit the free memory, it waits until a
 * running batch ends and frees its block (admission by HBM footprint).
 * Requests arriving while others are on the GPU join the next batch of the
 * first slot to free up; the copies of different slots overlap each other
 * and the kernels of the others (the kernels themselves run one batch at a
 * time, as every build on a device).  A request completes as soon as its own result
 * copies have ended, in batch order; its block is freed after the batch's last
 * copy.  With plain = 1 every slot instead serves one request at a time
 * through wahgpu_build (the per-column drop-in, with its own growing arena),
 * for comparison under the same arrivals. */
typedef struct wahgpu_server wahgpu_server;
typedef struct wahgpu_server_config {
  int device;
  uint32_t slots;            /* worker slots (batches in flight), >= 1 */
  uint64_t arena_bytes;      /* the device arena shared by all slots */
  uint32_t max_batch_cols;   /* columns per batch */
  uint64_t max_batch_values; /* values per batch; 0 = only the arena bounds it */
  uint32_t plain;            /* 1: one wahgpu_build per request */
  uint32_t reserved;
} wahgpu_server_config;

typedef struct wahgpu_server_stats {
  uint64_t requests;      /* completed */
  uint64_t batches;
  uint64_t values;        /* values of completed requests */
  uint64_t max_batch_cols;
  uint64_t slot_values;   /* max_batch_values of the configuration */
  double p50_us, p99_us, mean_us, max_us;  /* submit -> result in host buffers */
  double busy_s;          /* first submit -> last completion */
  uint64_t memory_waits;  /* times a slot waited for arena memory to free up */
  uint64_t arena_bytes;   /* the device arena */
  uint64_t arena_peak;    /* most bytes of it in use at once */
} wahgpu_server_stats;

int wahgpu_server_create(wahgpu_server** out, const wahgpu_server_config* cfg);
/* Stops the slots after the queued requests are served. */
void wahgpu_server_destroy(wahgpu_server* s);
/* Queues one column.  values must stay valid and the result buffers
 * (capacities as for wahgpu_build) untouched until the request completes;
 * result->counts is filled then.  Fails with WAHGPU_ERANGE for a column whose
 * footprint exceeds the whole arena (or max_batch_values).  *id names the
 * request. */
int wahgpu_server_submit(wahgpu_server* s, const uint32_t* values, uint64_t n,
                         uint32_t rid_base, wahgpu_result* result, uint64_t* id);
/* WAHGPU_OK when request id has completed, WAHGPU_EBUSY while pending,
 * the request's error code if it failed. */
int wahgpu_server_poll(wahgpu_server* s, uint64_t id);
/* Blocks until request id completes; returns as poll.  Completed requests
 * are forgotten once polled or waited for with a final status. */
int wahgpu_server_wait(wahgpu_server* s, uint64_t id);
/* Message of the last request that failed (empty if none); the pointer is
 * valid in the calling thread until its next call. */
const char* wahgpu_server_last_error(wahgpu_server* s);
/* Blocks until every submitted request has completed. */
int wahgpu_server_drain(wahgpu_server* s);
int wahgpu_server_get_stats(wahgpu_server* s, wahgpu_server_stats* out);
/* Clears the latency record and counters (between measurement phases). */
int wahgpu_server_reset_stats(wahgpu_server* s);

/* Page-locked host memory (hipHostMalloc) for request values and result
 * buffers: copies from and to it run at full link speed and asynchronously,
 * where pageable memory is staged through a bounce buffer. */
int wahgpu_host_alloc(uint64_t bytes, void** out);
void wahgpu_host_free(void* p);

/* ---- stages (device pointers; each synchronizes when it returns a count) */
/* include/sort_rids_by_value.cl:24-30 + include/radix.cl:65-233: stable sort
 * of (value, rid = rid_base + i) by value. */
int wahgpu_sort_rids_by_value(wahgpu_ctx* ctx, const uint32_t* d_values, uint64_t n,
                              uint32_t rid_base, uint32_t* d_keys_out,
                              uint32_t* d_rids_out);
/* include/produce_chunk_id_literals.cl:17-24 */
int wahgpu_produce_chunk_id_literals(wahgpu_ctx* ctx, const uint32_t* d_rids,
                                     uint64_t n, uint32_t* d_chids, uint32_t* d_lits);
/* include/merge_lit_by_val_chids.cl:32-66 + compaction (src/phases.cpp:864-903):
 * OR of lits over runs of equal (value, chid); returns k. */
int wahgpu_merge_lit_by_val_chids(wahgpu_ctx* ctx, const uint32_t* d_values,
                                  const uint32_t* d_chids, const uint32_t* d_lits,
                                  uint64_t n, uint32_t* d_values_out,
                                  uint32_t* d_chids_out, uint32_t* d_lits_out,
                                  uint64_t* k);
/* include/produce_fills.cl:24-43 */
int wahgpu_produce_fills(wahgpu_ctx* ctx, const uint32_t* d_values,
                         const uint32_t* d_chids, uint64_t k, uint32_t* d_fills);
/* include/fuse_fill_literals.cl:25-36 + compaction (src/phases.cpp:958-979):
 * interleave (fill, literal), drop zero words; d_index holds 2k; returns L. */
int wahgpu_fuse_fill_literals(wahgpu_ctx* ctx, const uint32_t* d_fills,
                              const uint32_t* d_lits, uint64_t k, uint32_t* d_index,
                              uint64_t* index_length);
/* include/compute_column_length.cl:34-70 + segmented scan + compaction + scan
 * (src/phases.cpp:989-1077): keys, per-key word counts, offsets; returns K. */
int wahgpu_compute_column_length(wahgpu_ctx* ctx, const uint32_t* d_values,
                                 const uint32_t* d_fills, uint64_t k,
                                 uint32_t* d_keys, uint32_t* d_offsets,
                                 uint32_t* d_lengths, uint64_t* keycnt);

/* ---- primitives -------------------------------------------------------- */
/* include/stream_compaction.cl:68-160: keep d_data[i] where d_valid[i] != 0,
 * in order; returns the count (the kernel's result[0]). */
int wahgpu_stream_compaction(wahgpu_ctx* ctx, const uint32_t* d_data,
                             const uint32_t* d_valid, uint64_t n, uint32_t* d_out,
                             uint64_t* count);
/* include/scan.cl:16-130: exclusive sum scan (uint32, wraps mod 2^32). */
int wahgpu_exclusive_scan(wahgpu_ctx* ctx, const uint32_t* d_in, uint64_t n,
                          uint32_t* d_out);
/* include/segmented_scan.cl:23-271: inclusive sum scan restarting where
 * d_heads[i] != 0 (element 0 always starts a segment). */
int wahgpu_segmented_inclusive_scan(wahgpu_ctx* ctx, const uint32_t* d_in,
                                    const uint32_t* d_heads, uint64_t n,
                                    uint32_t* d_out);

/* ---- stitching partial indexes (the merger actor, src/gpu.cpp:103-165) ---
 * Partial indexes built over consecutive row ranges (wahgpu_build with
 * rid_base = first row of the range) are appended in row order; each key's
 * words are concatenated with its first fill re-based on the key's previous
 * chunk, and a chunk shared across the boundary is OR-merged.  The result is
 * the index of the whole column.  The reference merger (src/gpu.cpp:140-149)
 * concatenates without re-basing; this one is exact.  Host only. */
typedef struct wahgpu_merger wahgpu_merger;
int wahgpu_merger_create(wahgpu_merger** out);
void wahgpu_merger_destroy(wahgpu_merger* m);
/* The `add` message: keycnt keys, offsets into words, `length` words total
 * (per-key lengths are implied, src/gpu.cpp:143-144). */
int wahgpu_merger_add(wahgpu_merger* m, uint64_t keycnt, uint64_t length,
                      const uint32_t* keys, const uint32_t* offsets,
                      const uint32_t* words);
uint64_t wahgpu_merger_keycnt(const wahgpu_merger* m);
uint64_t wahgpu_merger_length(const wahgpu_merger* m);
/* Writes the merged index: keys ascending, offsets, lengths, words. */
int wahgpu_merger_export(const wahgpu_merger* m, uint32_t* keys, uint32_t* offsets,
                         uint32_t* lengths, uint32_t* words);
/* Borrowed view of one key's words (NULL, 0 for an absent key). */
int wahgpu_merger_key_words(const wahgpu_merger* m, uint32_t key,
                            const uint32_t** words, uint64_t* len);

/* ---- query side (device pointers) ---------------------------------------
 * The reference decodes a key's words on the host, one '0'/'1' character per
 * row (decoded_bitmap, src/cpu.cpp:47-74; the merger's print,
 * src/gpu.cpp:151-158).  These run on the device over the index as the
 * builds leave it.  A row bitmap holds row r in bit r % 32 of word r / 32,
 * (nrows + 31) / 32 words; rows past nrows are dropped.  Row ids are the
 * build's (rid_base included).  Each call synchronizes its stream before it
 * returns. */

/* One key's WAH words -> its row bitmap (d_bitmap is overwritten).
 * *last_row (may be NULL) = one past the last row the words set: the length
 * of the reference's decoded string, whose trailing '0's are trimmed. */
int wahgpu_decode_device(wahgpu_ctx* ctx, const uint32_t* d_words, uint64_t len,
                         uint64_t nrows, uint32_t* d_bitmap, uint64_t* last_row);

/* Range query on an equality index (keys ascending; offsets, lengths and
 * words as a build writes them): d_bitmap = OR of the row bitmaps of every
 * key in [lo, hi].  *keys_matched (may be NULL) = keys in the range. */
int wahgpu_query_range_device(wahgpu_ctx* ctx, const uint32_t* d_keys,
                              const uint32_t* d_offsets, const uint32_t* d_lengths,
                              uint64_t keycnt, const uint32_t* d_words, uint32_t lo,
                              uint32_t hi, uint64_t nrows, uint32_t* d_bitmap,
                              uint64_t* keys_matched);

/* Row bitmap -> WAH words in the build's canonical form (src/cpu.cpp:223-257:
 * per nonempty chunk a 0-fill of the chunks skipped, when nonzero, then the
 * literal; no 1-fills, no trailing fill).  *len = words; fails with
 * WAHGPU_EINVAL when *len > cap (no word past cap is written). */
int wahgpu_encode_device(wahgpu_ctx* ctx, const uint32_t* d_bitmap, uint64_t nrows,
                         uint32_t* d_words, uint64_t cap, uint64_t* len);

/* Two WAH streams over rows [0, nrows) -> the canonical WAH words of
 * a op b.  cap = 2 (la + lb) always suffices. */
#define WAHGPU_OP_AND 0
#define WAHGPU_OP_OR 1
#define WAHGPU_OP_XOR 2
#define WAHGPU_OP_ANDNOT 3 /* a AND NOT b */
int wahgpu_combine_device(wahgpu_ctx* ctx, int op, const uint32_t* d_a, uint64_t la,
                          const uint32_t* d_b, uint64_t lb, uint64_t nrows, uint32_t* d_out,
                          uint64_t cap, uint64_t* len);

/* ---- host-side formats (no GPU needed) --------------------------------- */
/* Value files as the reference reads them (src/cpu.cpp:297-303): whitespace-
 * separated decimal, parsed like `ifstream >> uint32_t`, stopping silently at
 * the first failed extraction.  *values is malloc'ed; free with wahgpu_free.
 * The file is parsed in parallel (up to 16 threads): it is cut into spans at
 * white space, each span is parsed like the loop, and the values are those
 * before the first failed extraction, exactly as a serial read. */
int wahgpu_read_values_text(const char* path, uint32_t** values, uint64_t* n);
/* The same with a thread count (1 = one serial pass, 0 = as many as the
 * host has, at most 16). */
int wahgpu_read_values_text_threads(const char* path, int threads, uint32_t** values,
                                    uint64_t* n);
/* Writes values as the generator does (src/generate_test_data.cpp:48-54):
 * decimal, separated by sep, then a newline; formatted on `threads` threads
 * (0 = as many as the host has, at most 16). */
int wahgpu_write_values_text(const char* path, const uint32_t* values, uint64_t n,
                             const char* sep, int threads);
/* Raw little-endian uint32 files (opt-in extension). */
int wahgpu_read_values_binary(const char* path, uint32_t** values, uint64_t* n);
void wahgpu_free(void* p);
/* 64-bit order-sensitive checksum of an index (its keycnt keys, offsets and
 * lengths and its `length` words): equal indexes give equal checksums, so a
 * result can be compared with another build of the same column without
 * keeping both (e.g. a served request against the same column built alone). */
uint64_t wahgpu_index_checksum(const uint32_t* keys, const uint32_t* offsets,
                               const uint32_t* lengths, uint64_t keycnt, const uint32_t* words,
                               uint64_t length);
/* src/cpu.cpp:47-74: '0'/'1' text of a key's bitmap, trailing '0's trimmed.
 * Writes at most cap bytes (no terminator); returns the full length. */
uint64_t wahgpu_decode_bitmap(const uint32_t* words, uint64_t len, char* out,
                              uint64_t cap);

/* ---- synthetic columns (the `generate` program, src/generate_test_data.cpp)
 * Seeded and reproducible on any host: value i is a pure function of
 * (seed, i) for the uniform kind. */
#define WAHGPU_DIST_UNIFORM 0   /* i.i.d. uniform in [from, to]                  */
#define WAHGPU_DIST_ZIPF 1      /* Zipf(s) rank - 1 over `cardinality` values     */
#define WAHGPU_DIST_ZIPF_RUNS 2 /* Zipf(s) values in runs of Geometric(1/run_mean) */
typedef struct wahgpu_gen_params {
  uint32_t dist;
  uint32_t from, to;        /* uniform range (inclusive)                      */
  uint32_t cardinality;     /* zipf support size                              */
  double zipf_s;            /* zipf exponent                                  */
  double run_mean;          /* mean run length (zipf-runs)                     */
  uint64_t seed;
} wahgpu_gen_params;
int wahgpu_generate(const wahgpu_gen_params* p, uint32_t* out, uint64_t n);
/* Values at positions [start, start + count) of the same column. */
int wahgpu_generate_range(const wahgpu_gen_params* p, uint32_t* out, uint64_t start,
                          uint64_t count);

#ifdef __cplusplus
}
#endif
#endif
