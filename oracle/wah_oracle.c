/*
 * wah_oracle.c -- CPU oracle for the WAH bitmap-index build.
 *
 * TEST INFRASTRUCTURE ONLY: a plain-C restatement of `src/cpu.cpp` of the
 * reference, used to check libwahgpu and to time the CPU baseline.  It is
 * never linked into, loaded by or called from the product library or CLIs.
 * Each function cites the reference lines it follows (paths relative to the
 * reference repository root).
 *
 * Deviations, all outside the reference's defined behaviour:
 *   - n == 0 (and k == 0) return empty results; the reference underflows
 *     `size() - 1` (src/cpu.cpp:183) and reads `heads.front()` of an empty
 *     vector (:228).
 *   - rid_base: the reference always numbers rows from 0 (src/cpu.cpp:182);
 *     a non-zero base indexes the column as rows [base, base + n).
 */
#include "wah_oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ---- sort_rids_by_value ------------------------------------------------- */

/* Literal restatement of src/cpu.cpp:180-197: iota, then an O(n^2) bubble
 * sort that swaps on strict '>' only, hence stable. */
static void sort_bubble(uint32_t* input, uint32_t* rids, size_t n) {
  if (n < 2) return;
  for (size_t i = n - 1; i > 0; --i) {
    for (size_t j = 0; j < i; ++j) {
      if (input[j] > input[j + 1]) {
        uint32_t t = input[j]; input[j] = input[j + 1]; input[j + 1] = t;
        t = rids[j]; rids[j] = rids[j + 1]; rids[j + 1] = t;
      }
    }
  }
}

/* Top-down merge sort of 64-bit words.  The words are (value << 32 | rid)
 * with distinct, increasing rids, so their numeric order is exactly the
 * stable order by value of src/phases.cpp:143-160 (std::stable_sort). */
static void msort_u64(uint64_t* a, uint64_t* tmp, size_t n) {
  if (n < 32) {
    for (size_t i = 1; i < n; ++i) {
      uint64_t x = a[i];
      size_t j = i;
      while (j > 0 && a[j - 1] > x) { a[j] = a[j - 1]; --j; }
      a[j] = x;
    }
    return;
  }
  size_t h = n / 2;
  msort_u64(a, tmp, h);
  msort_u64(a + h, tmp, n - h);
  if (a[h - 1] <= a[h]) return;
  memcpy(tmp, a, h * sizeof(uint64_t));
  size_t i = 0, j = h, o = 0;
  while (i < h && j < n) a[o++] = (a[j] < tmp[i]) ? a[j++] : tmp[i++];
  while (i < h) a[o++] = tmp[i++];
}

static int sort_packed(uint32_t* input, uint32_t* rids, size_t n) {
  if (n < 2) return 0;
  uint64_t* w = (uint64_t*)malloc(n * sizeof(uint64_t));
  uint64_t* t = (uint64_t*)malloc((n / 2 + 1) * sizeof(uint64_t));
  if (!w || !t) { free(w); free(t); return -1; }
  for (size_t i = 0; i < n; ++i) w[i] = ((uint64_t)input[i] << 32) | rids[i];
  msort_u64(w, t, n);
  for (size_t i = 0; i < n; ++i) {
    input[i] = (uint32_t)(w[i] >> 32);
    rids[i] = (uint32_t)w[i];
  }
  free(w);
  free(t);
  return 0;
}

/* ---- the same order on several threads (CPU baseline only) -------------
 * Not a restatement of the reference (which is serial): the packed words are
 * cut into T slices sorted with msort_u64 on T threads; the output is then
 * cut into T ranges at ranks r_t = t n / T, and thread t merges the slices'
 * parts of range t.  The words are distinct, so the word of rank r is the
 * smallest v with count(words <= v) > r (a binary search over the 64-bit
 * values), and each slice's part below it is the count of its words < v. */
typedef struct {
  uint64_t* w;           /* slices, each sorted */
  uint64_t* out;         /* merged output (in the sort jobs: tmp space) */
  size_t n, lo, hi;      /* sort job: slice [lo, hi) */
  int threads, t;        /* merge job: range t of threads */
  const size_t* bounds;  /* slice bounds, threads + 1 */
} par_job;

static size_t count_below(const uint64_t* a, size_t n, uint64_t v, int inclusive) {
  size_t lo = 0, hi = n;  /* first index whose word is > v (inclusive) or >= v */
  while (lo < hi) {
    size_t mid = lo + (hi - lo) / 2;
    if (inclusive ? a[mid] <= v : a[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

/* cut[i] = words of slice i below the word of rank r (r = n: all). */
static void rank_cuts(const uint64_t* w, const size_t* b, int k, size_t n, size_t r,
                      size_t* cut) {
  if (r >= n) {
    for (int i = 0; i < k; ++i) cut[i] = b[i + 1] - b[i];
    return;
  }
  uint64_t lo = 0, hi = UINT64_MAX;  /* smallest v with count(<= v) > r */
  while (lo < hi) {
    uint64_t mid = lo + (hi - lo) / 2;
    size_t c = 0;
    for (int i = 0; i < k; ++i) c += count_below(w + b[i], b[i + 1] - b[i], mid, 1);
    if (c > r) hi = mid;
    else lo = mid + 1;
  }
  for (int i = 0; i < k; ++i) cut[i] = count_below(w + b[i], b[i + 1] - b[i], lo, 0);
}

static void* par_sort_slice(void* p) {
  par_job* j = (par_job*)p;
  msort_u64(j->w + j->lo, j->out + j->lo, j->hi - j->lo);
  return NULL;
}

static void* par_merge_range(void* p) {
  par_job* j = (par_job*)p;
  const int k = j->threads;
  size_t c0[64], c1[64];
  rank_cuts(j->w, j->bounds, k, j->n, j->n * (size_t)j->t / (size_t)k, c0);
  rank_cuts(j->w, j->bounds, k, j->n, j->n * (size_t)(j->t + 1) / (size_t)k, c1);
  size_t o = j->n * (size_t)j->t / (size_t)k;
  size_t pos[64], end[64];
  for (int i = 0; i < k; ++i) {
    pos[i] = j->bounds[i] + c0[i];
    end[i] = j->bounds[i] + c1[i];
  }
  for (;;) {
    int best = -1;
    for (int i = 0; i < k; ++i)
      if (pos[i] < end[i] && (best < 0 || j->w[pos[i]] < j->w[pos[best]])) best = i;
    if (best < 0) break;
    j->out[o++] = j->w[pos[best]++];
  }
  return NULL;
}

static int sort_packed_threads(uint32_t* input, uint32_t* rids, size_t n, int threads) {
  if (threads > 64) threads = 64;
  if (threads < 2 || n < (size_t)threads * 1024) return sort_packed(input, rids, n);
  uint64_t* w = (uint64_t*)malloc(n * sizeof(uint64_t));
  uint64_t* o = (uint64_t*)malloc(n * sizeof(uint64_t));
  if (!w || !o) { free(w); free(o); return -1; }
  for (size_t i = 0; i < n; ++i) w[i] = ((uint64_t)input[i] << 32) | rids[i];
  size_t b[65];
  for (int t = 0; t <= threads; ++t) b[t] = n * (size_t)t / (size_t)threads;
  pthread_t th[64];
  par_job jobs[64];
  for (int t = 0; t < threads; ++t) {
    jobs[t] = (par_job){w, o, n, b[t], b[t + 1], threads, t, b};
    pthread_create(&th[t], NULL, par_sort_slice, &jobs[t]);
  }
  for (int t = 0; t < threads; ++t) pthread_join(th[t], NULL);
  for (int t = 0; t < threads; ++t) pthread_create(&th[t], NULL, par_merge_range, &jobs[t]);
  for (int t = 0; t < threads; ++t) pthread_join(th[t], NULL);
  for (size_t i = 0; i < n; ++i) {
    input[i] = (uint32_t)(o[i] >> 32);
    rids[i] = (uint32_t)o[i];
  }
  free(w);
  free(o);
  return 0;
}

void wah_oracle_sort_rids_by_value(uint32_t* input, uint32_t* rids, size_t n,
                                   uint32_t rid_base, int mode) {
  for (size_t i = 0; i < n; ++i) rids[i] = rid_base + (uint32_t)i;  /* iota, :182 */
  if (mode == WAH_ORACLE_SORT_BUBBLE) sort_bubble(input, rids, n);
  else sort_packed(input, rids, n);
}

/* ---- produce_chunk_id_literals: src/cpu.cpp:201-211 -------------------- */
void wah_oracle_produce_chunk_id_literals(const uint32_t* rids, uint32_t* chids,
                                          uint32_t* lits, size_t n) {
  for (size_t i = 0; i < n; ++i) {
    lits[i] = 1u << (rids[i] % 31u);
    lits[i] |= 1u << 31;
    chids[i] = rids[i] / 31u;
  }
}

typedef struct {
  const uint32_t* rids;
  uint32_t *chids, *lits;
  size_t lo, hi;
} chunk_job;

static void* par_chunks(void* p) {
  chunk_job* j = (chunk_job*)p;
  wah_oracle_produce_chunk_id_literals(j->rids + j->lo, j->chids + j->lo, j->lits + j->lo,
                                       j->hi - j->lo);
  return NULL;
}

/* The same on `threads` threads, each on a contiguous slice (CPU baseline). */
static void produce_chunk_id_literals_threads(const uint32_t* rids, uint32_t* chids,
                                              uint32_t* lits, size_t n, int threads) {
  if (threads > 64) threads = 64;
  if (threads < 2 || n < (size_t)threads * 1024) {
    wah_oracle_produce_chunk_id_literals(rids, chids, lits, n);
    return;
  }
  pthread_t th[64];
  chunk_job jobs[64];
  for (int t = 0; t < threads; ++t) {
    jobs[t] = (chunk_job){rids, chids, lits, n * (size_t)t / (size_t)threads,
                          n * (size_t)(t + 1) / (size_t)threads};
    pthread_create(&th[t], NULL, par_chunks, &jobs[t]);
  }
  for (int t = 0; t < threads; ++t) pthread_join(th[t], NULL);
}

/* ---- merge_lit_by_val_chids: reduce_by_key OR, src/cpu.cpp:78-116 ------ */
size_t wah_oracle_merge_lit_by_val_chids(uint32_t* input, uint32_t* chids,
                                         uint32_t* lits, size_t n) {
  size_t from = 0, out = 0;
  while (from < n) {
    uint32_t v = input[from], c = chids[from];
    size_t to = from;
    while (to < n && input[to] == v && chids[to] == c) ++to;
    uint32_t merged = 0;
    while (from < to) merged |= lits[from++];
    input[out] = v;
    chids[out] = c;
    lits[out] = merged;
    ++out;
  }
  return out;
}

/* ---- produce_fills: src/cpu.cpp:223-241 --------------------------------- */
void wah_oracle_produce_fills(const uint32_t* input, uint32_t* chids, size_t k) {
  if (k == 0) return;
  /* heads = adjacent_difference(input); heads.front() = 1 (:227-228) */
  uint32_t prev_chid = chids[0];
  /* i == 0 is a head: fill = chid (:233-234) */
  for (size_t i = 1; i < k; ++i) {
    uint32_t head = input[i] - input[i - 1];
    uint32_t c = chids[i];
    chids[i] = (head == 0) ? c - prev_chid - 1u : c;  /* :231-232 */
    prev_chid = c;
  }
}

/* ---- stream_compaction: src/cpu.cpp:170-174 ----------------------------- */
size_t wah_oracle_stream_compaction(uint32_t* data, size_t n, uint32_t val) {
  size_t o = 0;
  for (size_t i = 0; i < n; ++i)
    if (data[i] != val) data[o++] = data[i];
  return o;
}

/* ---- fuse_fill_literals: src/cpu.cpp:245-257 ---------------------------- */
size_t wah_oracle_fuse_fill_literals(const uint32_t* fills, const uint32_t* lits,
                                     uint32_t* index, size_t k) {
  for (size_t i = 0; i < k; ++i) {
    index[2 * i] = fills[i];
    index[2 * i + 1] = lits[i];
  }
  return wah_oracle_stream_compaction(index, 2 * k, 0u);
}

/* ---- exclusive_scan: src/cpu.cpp:160-168 -------------------------------- */
void wah_oracle_exclusive_scan(const uint32_t* in, uint32_t* out, size_t n) {
  if (n == 0) return;
  uint32_t run = 0;
  for (size_t i = 0; i < n; ++i) {
    uint32_t x = in[i];
    out[i] = run;
    run += x;
  }
}

/* ---- reduce_by_key SUM: src/cpu.cpp:119-148 (in place, returns K) ------- */
static size_t reduce_by_key_sum(uint32_t* keys, uint32_t* vals, size_t n) {
  size_t from = 0, out = 0;
  while (from < n) {
    uint32_t key = keys[from];
    size_t to = from;
    while (to < n && keys[to] == key) ++to;
    uint32_t sum = 0;
    while (from < to) sum += vals[from++];
    keys[out] = key;
    vals[out] = sum;
    ++out;
  }
  return out;
}

/* ---- compute_colum_length: src/cpu.cpp:261-272 -------------------------- */
size_t wah_oracle_compute_column_length(uint32_t* input, const uint32_t* fills,
                                        uint32_t* offsets, uint32_t* lengths,
                                        size_t k) {
  for (size_t i = 0; i < k; ++i) lengths[i] = 1u + (fills[i] == 0 ? 0u : 1u);
  size_t keycnt = reduce_by_key_sum(input, lengths, k);
  wah_oracle_exclusive_scan(lengths, offsets, keycnt);
  return keycnt;
}

/* ---- compaction by flags: src/sc.cpp:67-78 ------------------------------ */
size_t wah_oracle_compact_valid(const uint32_t* data, const uint32_t* valid,
                                size_t n, uint32_t* out) {
  size_t o = 0;
  for (size_t i = 0; i < n; ++i)
    if (valid[i] != 0) out[o++] = data[i];
  return o;
}

/* ---- segmented inclusive scan: src/segmented_scan.cpp:109-121 ----------- */
void wah_oracle_segmented_inclusive_scan(const uint32_t* in, const uint32_t* heads,
                                         uint32_t* out, size_t n) {
  for (size_t i = 0; i < n; ++i) {
    if (i == 0 || heads[i] != 0) out[i] = in[i];
    else out[i] = out[i - 1] + in[i];
  }
}

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* ---- whole build: src/cpu.cpp:313-328 ----------------------------------- */
static int build_impl(const uint32_t* values, size_t n, uint32_t rid_base, int sort_mode,
                      int threads, uint32_t* keys, uint32_t* offsets, uint32_t* lengths,
                      uint32_t* words, wah_oracle_counts* out, double* stage_s) {
  double t[7] = {0};
  out->k = out->length = out->keycnt = 0;
  if (stage_s)
    for (int i = 0; i < WAH_ORACLE_STAGES; ++i) stage_s[i] = 0.0;
  if (n == 0) return 0;
  uint32_t* input = (uint32_t*)malloc(n * sizeof(uint32_t));
  uint32_t* rids = (uint32_t*)malloc(n * sizeof(uint32_t));
  uint32_t* chids = (uint32_t*)malloc(n * sizeof(uint32_t));
  uint32_t* lits = (uint32_t*)malloc(n * sizeof(uint32_t));
  if (!input || !rids || !chids || !lits) {
    free(input); free(rids); free(chids); free(lits);
    return -1;
  }
  memcpy(input, values, n * sizeof(uint32_t));                   /* :313 */
  t[0] = now_s();
  if (threads > 1) {
    for (size_t i = 0; i < n; ++i) rids[i] = rid_base + (uint32_t)i;  /* iota, :182 */
    if (sort_packed_threads(input, rids, n, threads)) {
      free(input); free(rids); free(chids); free(lits);
      return -1;
    }
  } else {
    wah_oracle_sort_rids_by_value(input, rids, n, rid_base, sort_mode); /* :318 */
  }
  t[1] = now_s();
  produce_chunk_id_literals_threads(rids, chids, lits, n, threads); /* :319 */
  t[2] = now_s();
  size_t k = wah_oracle_merge_lit_by_val_chids(input, chids, lits, n); /* :320 */
  t[3] = now_s();
  wah_oracle_produce_fills(input, chids, k);                     /* :322 */
  t[4] = now_s();
  size_t len = wah_oracle_fuse_fill_literals(chids, lits, words, k); /* :323-324 */
  t[5] = now_s();
  size_t keycnt = wah_oracle_compute_column_length(input, chids, offsets,
                                                   lengths, k);  /* :325-326 */
  memcpy(keys, input, keycnt * sizeof(uint32_t));                /* keys = input[i], :332 */
  t[6] = now_s();
  if (stage_s)
    for (int i = 0; i < WAH_ORACLE_STAGES; ++i) stage_s[i] = t[i + 1] - t[i];
  out->k = k;
  out->length = len;
  out->keycnt = keycnt;
  free(input); free(rids); free(chids); free(lits);
  return 0;
}

int wah_oracle_build_timed(const uint32_t* values, size_t n, uint32_t rid_base,
                           int sort_mode, uint32_t* keys, uint32_t* offsets,
                           uint32_t* lengths, uint32_t* words, wah_oracle_counts* out,
                           double* stage_s) {
  return build_impl(values, n, rid_base, sort_mode, 1, keys, offsets, lengths, words, out,
                    stage_s);
}

int wah_oracle_build_threads(const uint32_t* values, size_t n, uint32_t rid_base,
                             int threads, uint32_t* keys, uint32_t* offsets,
                             uint32_t* lengths, uint32_t* words, wah_oracle_counts* out,
                             double* stage_s) {
  return build_impl(values, n, rid_base, WAH_ORACLE_SORT_PACKED, threads, keys, offsets,
                    lengths, words, out, stage_s);
}

int wah_oracle_build(const uint32_t* values, size_t n, uint32_t rid_base,
                     int sort_mode, uint32_t* keys, uint32_t* offsets,
                     uint32_t* lengths, uint32_t* words, wah_oracle_counts* out) {
  return wah_oracle_build_timed(values, n, rid_base, sort_mode, keys, offsets, lengths,
                                words, out, NULL);
}

/* ---- decoded_bitmap: src/cpu.cpp:47-74 ---------------------------------- */
size_t wah_oracle_decoded_bitmap(const uint32_t* words, size_t len, char* out,
                                 size_t cap) {
  size_t pos = 0, last_one = 0; /* length up to and including the last '1' */
  for (size_t w = 0; w < len; ++w) {
    uint32_t block = words[w];
    if (block & 0x80000000u) {
      for (int i = 0; i < 31; ++i) {
        char ch = (block >> i) & 1u ? '1' : '0';
        if (pos < cap) out[pos] = ch;
        ++pos;
        if (ch == '1') last_one = pos;
      }
    } else {
      char bit = (block & 0x40000000u) ? '1' : '0';
      uint32_t times = block & 0x3FFFFFFFu;
      for (uint64_t i = 0; i < (uint64_t)times * 31u; ++i) {
        if (pos < cap) out[pos] = bit;
        ++pos;
        if (bit == '1') last_one = pos;
      }
    }
  }
  return last_one; /* res.erase(res.find_last_not_of("0") + 1), :72 */
}
