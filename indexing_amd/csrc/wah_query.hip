// wah_query.hip -- gfx950 kernels of the query side of a WAH index: decoding
// WAH word streams into row bitmaps, and encoding row bitmaps (or a bitwise
// combination of two) back into WAH words in the build's canonical form.
//
// Word format (src/cpu.cpp:47-74): bit 31 set = literal, whose payload bits
// 0..30 are the rows 31 c + b of its chunk c; bit 31 clear = fill of
// (bits 0..29) chunks, all of value bit 30.  Each key's stream starts at
// chunk 0, so a word's first chunk is the sum of the chunk counts of the
// words before it in its key's stream.
//
// Row bitmap: row r is bit r % 32 of 32-bit word r / 32.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "wah_device.h"
#include "wah_kernels.h"

namespace wah {

namespace {

constexpr int kDecItems = 8;
constexpr int kDecTile = kBlock * kDecItems;
constexpr int kEncItems = 8;                  // chunks per thread
constexpr int kEncTile = kBlock * kEncItems;  // chunks per tile

__device__ __forceinline__ uint32_t word_chunks(uint32_t w) {
  return (w & 0x80000000u) ? 1u : (w & 0x3FFFFFFFu);
}

// ORs the 31 bits `bits` (bit b = row row0 + b) into the bitmap, dropping
// rows at or past nrows.
__device__ __forceinline__ void or_bits31(uint32_t* bitmap, uint64_t nrows, uint64_t row0,
                                          uint32_t bits) {
  if (row0 >= nrows || bits == 0) return;
  if (nrows - row0 < 31) bits &= (1u << (uint32_t)(nrows - row0)) - 1u;
  if (!bits) return;
  const uint64_t wi = row0 >> 5;
  const uint32_t sh = (uint32_t)(row0 & 31u);
  const uint32_t lo = bits << sh;
  if (lo) atomicOr(&bitmap[wi], lo);
  if (sh > 1) {
    const uint32_t hi = bits >> (32u - sh);
    if (hi) atomicOr(&bitmap[wi + 1], hi);
  }
}

// Sets rows [row0, row0 + count) (a 1-fill; the build never writes one, the
// decoder of src/cpu.cpp:60-66 accepts it).
__device__ void set_rows(uint32_t* bitmap, uint64_t nrows, uint64_t row0, uint64_t count) {
  uint64_t end = row0 + count < nrows ? row0 + count : nrows;
  for (uint64_t r = row0; r < end;) {
    const uint32_t sh = (uint32_t)(r & 31u);
    const uint64_t take = (32u - sh) < end - r ? (32u - sh) : end - r;
    const uint32_t m = take == 32 ? 0xFFFFFFFFu : ((1u << (uint32_t)take) - 1u) << sh;
    atomicOr(&bitmap[r >> 5], m);
    r += take;
  }
}

// Chunk count of word i, with the segment-start flag in bit 32.
__device__ __forceinline__ uint64_t dec_elem(const uint32_t* words, const uint32_t* heads,
                                             uint64_t i) {
  return (uint64_t)word_chunks(words[i]) |
         ((uint64_t)(i == 0 || (heads && heads[i] != 0u)) << 32);
}

}  // namespace

// One pass over a stream of WAH words (a key's, or the concatenated streams
// of several keys with heads[i] != 0 where a key starts): per word its first
// chunk (segmented scan of chunk counts, decoupled look-back across tiles),
// then its literal bits or 1-fill rows ORed into the row bitmap.
__global__ __launch_bounds__(kBlock) void k_q_decode(const uint32_t* __restrict__ words,
                                                     const uint32_t* __restrict__ heads,
                                                     uint64_t len, uint32_t* __restrict__ bitmap,
                                                     uint64_t nrows, uint64_t* status,
                                                     uint32_t num_tiles, uint32_t* tile_counter,
                                                     uint32_t lanes, uint32_t* fault,
                                                     uint64_t* last_row) {
  __shared__ uint64_t s_scan[kWaves];
  __shared__ uint64_t s_prefix;
  __shared__ uint32_t s_tile;
  const int t = threadIdx.x;
  uint32_t iter = 0;
  for (;;) {
    const uint32_t tile = next_tile(tile_counter, lanes, &s_tile, iter);
    if (tile >= num_tiles) break;
    const uint64_t g0 = (uint64_t)tile * kDecTile + (uint64_t)t * kDecItems;
    uint64_t agg = OpSegAdd::identity();
#pragma unroll
    for (int i = 0; i < kDecItems; ++i)
      if (g0 + i < len) agg = OpSegAdd::combine(agg, dec_elem(words, heads, g0 + i));
    uint64_t tile_agg;
    const uint64_t ex = block_exclusive<OpSegAdd>(agg, s_scan, &tile_agg);
    if (t == 0) s_prefix = lookback64<OpSegAdd>(status, tile, tile_agg, fault);
    __syncthreads();
    uint64_t run = OpSegAdd::combine(s_prefix, ex);
    uint64_t my_last = 0;  // row after the last set row of this thread's words
#pragma unroll
    for (int i = 0; i < kDecItems; ++i) {
      const uint64_t g = g0 + i;
      if (g < len) {
        const uint32_t w = words[g];
        const uint64_t e = dec_elem(words, heads, g);
        // first chunk of this word = inclusive sum before adding it
        const uint32_t c0 = (e >> 32) ? 0u : (uint32_t)run;
        run = OpSegAdd::combine(run, e);
        const uint64_t row0 = 31ull * c0;
        if (w & 0x80000000u) {
          const uint32_t bits = w & 0x7FFFFFFFu;
          or_bits31(bitmap, nrows, row0, bits);
          if (bits) {
            const uint64_t r = row0 + 32u - (uint32_t)__clz(bits);
            my_last = r > my_last ? r : my_last;
          }
        } else if ((w & 0x40000000u) && (w & 0x3FFFFFFFu)) {
          const uint64_t cnt = 31ull * (w & 0x3FFFFFFFu);
          set_rows(bitmap, nrows, row0, cnt);
          my_last = row0 + cnt > my_last ? row0 + cnt : my_last;
        }
      }
    }
    if (last_row && my_last) atomicMax((unsigned long long*)last_row, (unsigned long long)my_last);
    __syncthreads();
  }
}

// Marks the first word of each key k in [a, b) of an index: heads[offsets[k]
// - w0] = 1 (heads zeroed before).
__global__ __launch_bounds__(kBlock) void k_q_mark_heads(const uint32_t* __restrict__ offsets,
                                                         uint64_t a, uint64_t b, uint64_t w0,
                                                         uint32_t* __restrict__ heads) {
  for (uint64_t k = a + (uint64_t)blockIdx.x * kBlock + threadIdx.x; k < b;
       k += (uint64_t)gridDim.x * kBlock)
    heads[offsets[k] - w0] = 1u;
}

// Locates the keys of an index in [lo, hi] (keys ascend): range[0] = first
// key index >= lo, range[1] = first > hi, range[2] = first word of key
// range[0], range[3] = one past the last word of key range[1] - 1.
__global__ void k_q_key_range(const uint32_t* __restrict__ keys,
                              const uint32_t* __restrict__ offsets,
                              const uint32_t* __restrict__ lengths, uint64_t keycnt,
                              uint32_t lo, uint32_t hi, uint64_t* __restrict__ range) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  uint64_t a = 0, e = keycnt;
  while (a < e) {  // lower bound of lo
    const uint64_t m = (a + e) / 2;
    if (keys[m] < lo) a = m + 1;
    else e = m;
  }
  uint64_t b = a;
  e = keycnt;
  while (b < e) {  // upper bound of hi
    const uint64_t m = (b + e) / 2;
    if (keys[m] <= hi) b = m + 1;
    else e = m;
  }
  range[0] = a;
  range[1] = b;
  range[2] = a < b ? offsets[a] : 0;
  range[3] = a < b ? (uint64_t)offsets[b - 1] + lengths[b - 1] : 0;
}

// ---- encoding a row bitmap (or op(a, b) of two) into WAH words -------------

namespace {

// Payload of chunk c: rows 31 c .. 31 c + 30 of the bitmap (rows >= nrows
// are zero).
__device__ __forceinline__ uint32_t chunk_bits(const uint32_t* bm, uint64_t nrows, uint64_t c) {
  const uint64_t r = 31ull * c;
  if (r >= nrows) return 0u;
  const uint64_t wi = r >> 5;
  const uint32_t sh = (uint32_t)(r & 31u);
  const uint64_t nwords = (nrows + 31) / 32;
  uint64_t x = bm[wi];
  if (wi + 1 < nwords) x |= (uint64_t)bm[wi + 1] << 32;
  uint32_t bits = (uint32_t)(x >> sh) & 0x7FFFFFFFu;
  if (nrows - r < 31) bits &= (1u << (uint32_t)(nrows - r)) - 1u;
  return bits;
}

__device__ __forceinline__ uint32_t combine_bits(int op, uint32_t a, uint32_t b) {
  switch (op) {
    case 0: return a & b;   // AND
    case 1: return a | b;   // OR
    case 2: return a ^ b;   // XOR
    case 3: return a & ~b;  // AND NOT
    default: return a;      // the first bitmap alone
  }
}

struct OpMax {
  static __device__ __forceinline__ uint64_t identity() { return 0; }
  static __device__ __forceinline__ uint64_t combine(uint64_t a, uint64_t b) {
    return a > b ? a : b;
  }
};

// Per thread: its kEncItems chunks' payloads, the words they emit counting a
// fill only where the previous emitted chunk is this thread's own, and its
// first and last emitted chunk (+1; 0 = none).
struct ThreadChunks {
  uint32_t bits[kEncItems];
  uint32_t words;
  uint64_t first, last;
};

__device__ __forceinline__ ThreadChunks thread_chunks(const uint32_t* a, const uint32_t* b,
                                                      int op, uint64_t nrows, uint64_t chunks,
                                                      uint64_t c0) {
  ThreadChunks r;
  r.words = 0;
  r.first = r.last = 0;
#pragma unroll
  for (int i = 0; i < kEncItems; ++i) {
    const uint64_t c = c0 + i;
    uint32_t x = 0;
    if (c < chunks) {
      const uint32_t pa = chunk_bits(a, nrows, c);
      x = b ? combine_bits(op, pa, chunk_bits(b, nrows, c)) : pa;
    }
    r.bits[i] = x;
    if (x) {
      // previous emitted chunk of this thread is r.last - 1: the fill
      // c - r.last is a word when nonzero; the fill of the thread's first
      // entry is counted by the caller
      r.words += 1u + (r.last && c != r.last ? 1u : 0u);
      if (!r.first) r.first = c + 1;
      r.last = c + 1;
    }
  }
  return r;
}

}  // namespace

// Pass 1: per tile of kEncTile chunks, the words its entries emit except the
// fill of its first entry, and its first and last emitted chunk (+1).
__global__ __launch_bounds__(kBlock) void k_q_enc_count(const uint32_t* __restrict__ a,
                                                        const uint32_t* __restrict__ b, int op,
                                                        uint64_t nrows, uint64_t chunks,
                                                        uint64_t* __restrict__ tile_sum) {
  __shared__ uint64_t s_scan[kWaves];
  const int t = threadIdx.x;
  const uint64_t c0 = (uint64_t)blockIdx.x * kEncTile + (uint64_t)t * kEncItems;
  const ThreadChunks tc = thread_chunks(a, b, op, nrows, chunks, c0);
  // fill of this thread's first entry, when an earlier thread of the tile
  // emitted: chunk gap to that thread's last entry
  uint64_t tot;
  const uint64_t prev = block_exclusive<OpMax>(tc.last, s_scan, &tot);
  uint32_t w = tc.words;
  if (tc.first && prev) w += (tc.first != prev + 1) ? 1u : 0u;
  uint64_t wsum;
  (void)block_exclusive<OpAdd>(w, s_scan, &wsum);
  uint64_t first_any;
  (void)block_exclusive<OpMax>(tc.first ? (uint64_t)(0xFFFFFFFFFFull - tc.first) : 0ull,
                               s_scan, &first_any);
  if (t == 0) {
    tile_sum[3 * (uint64_t)blockIdx.x + 0] = wsum;
    tile_sum[3 * (uint64_t)blockIdx.x + 1] = first_any ? 0xFFFFFFFFFFull - first_any : 0ull;
    tile_sum[3 * (uint64_t)blockIdx.x + 2] = tot;
  }
}

// Pass 2 (one block): per tile its first word and the last emitted chunk
// (+1) before it; the total word count into *len (L).
__global__ __launch_bounds__(1024) void k_q_enc_scan(uint64_t* __restrict__ tile_sum,
                                                     uint32_t tiles, uint64_t* __restrict__ len) {
  __shared__ uint64_t s_part[1024];
  __shared__ uint64_t s_last[1024];
  const uint32_t t = threadIdx.x;
  const uint32_t per = (tiles + 1023) / 1024;
  const uint32_t s0 = t * per < tiles ? t * per : tiles;
  const uint32_t s1 = s0 + per < tiles ? s0 + per : tiles;
  // strip summary: last emitted chunk of the strip
  uint64_t last = 0;
  for (uint32_t s = s0; s < s1; ++s) {
    const uint64_t l = tile_sum[3 * (uint64_t)s + 2];
    last = l > last ? l : last;
  }
  s_last[t] = last;
  __syncthreads();
  // exclusive max over strips (chunks ascend, so max = last nonempty)
  if (t == 0) {
    uint64_t m = 0;
    for (uint32_t i = 0; i < 1024; ++i) {
      const uint64_t x = s_last[i];
      s_last[i] = m;
      m = x > m ? x : m;
    }
  }
  __syncthreads();
  // strip words, with each tile's first fill against the last chunk before it
  uint64_t prev = s_last[t], words = 0;
  for (uint32_t s = s0; s < s1; ++s) {
    const uint64_t w = tile_sum[3 * (uint64_t)s + 0];
    const uint64_t f = tile_sum[3 * (uint64_t)s + 1];
    const uint64_t l = tile_sum[3 * (uint64_t)s + 2];
    // a key's first fill is its chunk (prev = 0 means chunk -1)
    words += w + (f ? (f != prev + 1 ? 1u : 0u) : 0u);
    prev = l > prev ? l : prev;
  }
  s_part[t] = words;
  __syncthreads();
  if (t == 0) {
    uint64_t run = 0;
    for (uint32_t i = 0; i < 1024; ++i) {
      const uint64_t x = s_part[i];
      s_part[i] = run;
      run += x;
    }
    *len = run;
  }
  __syncthreads();
  uint64_t pos = s_part[t];
  prev = s_last[t];
  for (uint32_t s = s0; s < s1; ++s) {
    const uint64_t w = tile_sum[3 * (uint64_t)s + 0];
    const uint64_t f = tile_sum[3 * (uint64_t)s + 1];
    const uint64_t l = tile_sum[3 * (uint64_t)s + 2];
    tile_sum[3 * (uint64_t)s + 0] = pos;   // first word of the tile
    tile_sum[3 * (uint64_t)s + 1] = prev;  // last emitted chunk (+1) before it
    pos += w + (f ? (f != prev + 1 ? 1u : 0u) : 0u);
    prev = l > prev ? l : prev;
  }
}

// Pass 3: each tile writes its words: per emitted chunk c a fill of
// c - prev - 1 chunks when nonzero (prev = the previous emitted chunk, -1 at
// the start), then the literal 0x80000000 | payload.  Words past cap are not
// written (the caller compares L with cap).
__global__ __launch_bounds__(kBlock) void k_q_enc_write(const uint32_t* __restrict__ a,
                                                        const uint32_t* __restrict__ b, int op,
                                                        uint64_t nrows, uint64_t chunks,
                                                        const uint64_t* __restrict__ tile_sum,
                                                        uint32_t* __restrict__ out, uint64_t cap) {
  __shared__ uint64_t s_scan[kWaves];
  const int t = threadIdx.x;
  const uint64_t c0 = (uint64_t)blockIdx.x * kEncTile + (uint64_t)t * kEncItems;
  const ThreadChunks tc = thread_chunks(a, b, op, nrows, chunks, c0);
  uint64_t tot;
  uint64_t prev = block_exclusive<OpMax>(tc.last, s_scan, &tot);
  if (!prev) prev = tile_sum[3 * (uint64_t)blockIdx.x + 1];
  uint32_t w = tc.words;
  if (tc.first) w += (tc.first != prev + 1) ? 1u : 0u;
  uint64_t wsum;
  uint64_t pos = tile_sum[3 * (uint64_t)blockIdx.x + 0] + block_exclusive<OpAdd>(w, s_scan, &wsum);
#pragma unroll
  for (int i = 0; i < kEncItems; ++i) {
    const uint32_t x = tc.bits[i];
    if (x) {
      const uint64_t c1 = c0 + i + 1;  // chunk + 1
      const uint64_t fill = c1 - prev - 1;
      if (fill) {
        if (pos < cap) out[pos] = (uint32_t)fill;
        ++pos;
      }
      if (pos < cap) out[pos] = 0x80000000u | x;
      ++pos;
      prev = c1;
    }
  }
}

// ---- launchers ---------------------------------------------------------------

uint64_t decode_status_words(uint64_t len) { return (len + kDecTile - 1) / kDecTile + 1; }
uint64_t encode_tiles(uint64_t nrows) {
  const uint64_t chunks = (nrows + 30) / 31;
  return (chunks + kEncTile - 1) / kEncTile;
}

void launch_decode(const ScanArgs& s, const uint32_t* words, const uint32_t* heads,
                   uint64_t len, uint32_t* bitmap, uint64_t nrows, uint64_t* last_row,
                   hipStream_t st) {
  const uint32_t tiles = (uint32_t)((len + kDecTile - 1) / kDecTile);
  if (!tiles) return;
  const uint32_t grid = tiles < s.max_resident ? tiles : s.max_resident;
  hipLaunchKernelGGL(k_q_decode, dim3(grid), dim3(kBlock), 0, st, words, heads, len, bitmap,
                     nrows, s.status, tiles, s.tile_counter, tile_lanes(grid), s.fault, last_row);
}

void launch_mark_heads(const uint32_t* offsets, uint64_t a, uint64_t b, uint64_t w0,
                       uint32_t* heads, hipStream_t st) {
  if (b <= a) return;
  uint64_t g = (b - a + kBlock - 1) / kBlock;
  if (g > 4096) g = 4096;
  hipLaunchKernelGGL(k_q_mark_heads, dim3((uint32_t)g), dim3(kBlock), 0, st, offsets, a, b, w0,
                     heads);
}

void launch_key_range(const uint32_t* keys, const uint32_t* offsets, const uint32_t* lengths,
                      uint64_t keycnt, uint32_t lo, uint32_t hi, uint64_t* range, hipStream_t st) {
  hipLaunchKernelGGL(k_q_key_range, dim3(1), dim3(64), 0, st, keys, offsets, lengths, keycnt, lo,
                     hi, range);
}

void launch_encode(const uint32_t* a, const uint32_t* b, int op, uint64_t nrows,
                   uint64_t* tile_sum, uint32_t* out, uint64_t cap, uint64_t* len,
                   hipStream_t st) {
  const uint64_t chunks = (nrows + 30) / 31;
  const uint32_t tiles = (uint32_t)encode_tiles(nrows);
  if (!tiles) return;
  hipLaunchKernelGGL(k_q_enc_count, dim3(tiles), dim3(kBlock), 0, st, a, b, op, nrows, chunks,
                     tile_sum);
  hipLaunchKernelGGL(k_q_enc_scan, dim3(1), dim3(1024), 0, st, tile_sum, tiles, len);
  hipLaunchKernelGGL(k_q_enc_write, dim3(tiles), dim3(kBlock), 0, st, a, b, op, nrows, chunks,
                     tile_sum, out, cap);
}

}  // namespace wah
