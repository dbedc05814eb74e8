// wah_kernels.hip -- gfx950 kernels of the WAH bitmap-index build.
//
// Fused build (what wahgpu_build runs), for a column of n uint32 values:
//   k_fill32      zeroes the sort header (histogram, counters)
//   k_histogram   one read of the values: histograms of all four 8-bit
//                 digits and the maximum, zeroing of look-back state
//   k_plan        digit offsets; passes whose digit is the same for every
//                 value are switched off (on the device, no host sync)
//   k_onesweep x4 stable LSD radix pass per active digit: wave64 ballot
//                 ranking, LDS staging, decoupled look-back over tiles
//   k_tail        sorted (value, rid) -> chunk ids, OR-merged literals, fills,
//                 zero-fill removal, word positions, keys and offsets
//   k_lengths     lengths[i] = offsets[i+1] - offsets[i]
// Staged kernels (what `phases` runs, one per reference stage) and the scan /
// compaction primitives reuse the same look-back scheme.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "wah_device.h"
#include "wah_kernels.h"

namespace wah {

__global__ __launch_bounds__(kBlock) void k_fill32(uint32_t* __restrict__ p, uint64_t words,
                                                   uint32_t v) {
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < words;
       i += (uint64_t)gridDim.x * kBlock)
    p[i] = v;
}

// ---------------------------------------------------------------------------
// Histogram of all four 8-bit digits (reads the values once).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void hist_add(uint32_t* h, uint32_t bin, bool valid) {
  // A wave whose valid lanes all share the bin adds once (the high digits of
  // small-range columns are constant); otherwise one LDS atomic per lane.
  const uint32_t first = __builtin_amdgcn_readfirstlane(bin);
  const uint64_t differ = ballot(valid && bin != first);
  if (differ == 0) {
    const uint64_t cnt = ballot(valid);
    if (lane_id() == 0 && cnt) atomicAdd(&h[first], (uint32_t)__popcll(cnt));
  } else if (valid) {
    atomicAdd(&h[bin], 1u);
  }
}

__global__ __launch_bounds__(kBlock) void k_histogram(
    const uint32_t* __restrict__ values, uint64_t n, int vec4, SortState* __restrict__ st,
    uint32_t* __restrict__ status0, uint64_t status0_words,
    uint64_t* __restrict__ tail_status, uint64_t tail_words,
    wahgpu_counts_dev* __restrict__ counts) {
  __shared__ uint32_t h[4 * 256];
  __shared__ uint32_t s_max[kWaves];
  const int t = threadIdx.x;
  for (int i = t; i < 1024; i += kBlock) h[i] = 0;
  if (blockIdx.x == 0 && t == 0) {
    counts->k = 0; counts->length = 0; counts->keycnt = 0;
    counts->bound = 0; counts->passes = 0; counts->fault = 0; counts->reserved = 0;
  }
  // zero look-back state used later in this build
  const uint64_t gthreads = (uint64_t)gridDim.x * kBlock;
  const uint64_t gt = (uint64_t)blockIdx.x * kBlock + t;
  for (uint64_t i = gt; i < status0_words; i += gthreads) status0[i] = 0;
  for (uint64_t i = gt; i < tail_words; i += gthreads) tail_status[i] = 0;
  __syncthreads();

  uint32_t vmax = 0;
  // Grid-stride over chunks of 4096 values, four 16-byte loads per thread.
  constexpr uint64_t kChunk = 4 * 4 * kBlock;
  uint64_t scalar_lo = 0;
  if (vec4) {
    const uint4* v4 = reinterpret_cast<const uint4*>(values);
    const uint64_t full = n / kChunk;
    for (uint64_t c = blockIdx.x; c < full; c += gridDim.x) {
      uint4 q[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) q[u] = v4[c * (kChunk / 4) + u * kBlock + t];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const uint32_t e[4] = {q[u].x, q[u].y, q[u].z, q[u].w};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          vmax = e[j] > vmax ? e[j] : vmax;
#pragma unroll
          for (int p = 0; p < 4; ++p) hist_add(h, p * 256 + ((e[j] >> (8 * p)) & 255u), true);
        }
      }
    }
    scalar_lo = full * kChunk;
  }
  // remainder (or unaligned input): one value per thread per step
  for (uint64_t b = scalar_lo + (uint64_t)blockIdx.x * kBlock; b < n;
       b += (uint64_t)gridDim.x * kBlock) {
    const uint64_t i = b + t;
    const bool valid = i < n;
    const uint32_t v = valid ? values[i] : 0;
    vmax = valid && v > vmax ? v : vmax;
#pragma unroll
    for (int p = 0; p < 4; ++p) hist_add(h, p * 256 + ((v >> (8 * p)) & 255u), valid);
  }
  for (int d = 32; d >= 1; d >>= 1) {
    const uint32_t o = __shfl_xor(vmax, d, kWave);
    vmax = o > vmax ? o : vmax;
  }
  if ((t % kWave) == 0) s_max[t / kWave] = vmax;
  __syncthreads();
  // fold into one striped copy of the column histogram (zeroed by the
  // previous launch); copies sit on different memory channels
  uint32_t* copy = st->hist[blockIdx.x % kHistCopies];
  for (int i = t; i < 1024; i += kBlock)
    if (h[i]) atomicAdd(&copy[i], h[i]);
  if (t == 0) {
    uint32_t m = 0;
    for (int i = 0; i < kWaves; ++i) m = s_max[i] > m ? s_max[i] : m;
    atomicMax(&st->maxval, m);
  }
}

// One block, after the histogram: digit offsets and the pass plan.  A kernel
// of its own so the histogram blocks need no device-scope fences (on gfx950
// each one writes back and invalidates the L2).
__global__ __launch_bounds__(kBlock) void k_plan(SortState* __restrict__ st, uint64_t n,
                                                 wahgpu_counts_dev* __restrict__ counts) {
  __shared__ uint64_t s_scan[kWaves];
  __shared__ int s_trivial[4];
  const int t = threadIdx.x;
  if (t < 4) s_trivial[t] = 0;
  __syncthreads();
  for (int p = 0; p < 4; ++p) {
    uint32_t c = 0;
    for (uint32_t k = 0; k < kHistCopies; ++k) c += st->hist[k][p * 256 + t];
    uint64_t total;
    const uint64_t ex = block_exclusive<OpAdd>(c, s_scan, &total);
    st->digit_offset[p][t] = (uint32_t)ex;
    if ((uint64_t)c == n) s_trivial[p] = 1;  // every value has digit t
  }
  __syncthreads();
  if (t == 0) {
    uint32_t a = 0;
    for (int p = 0; p < 4; ++p) {
      PassInfo pi;
      pi.active = (n > 0 && !s_trivial[p]) ? 1u : 0u;
      pi.shift = 8u * p;
      pi.dst = (a % 2 == 0) ? kBufA : kBufB;
      pi.src = (a == 0) ? kBufValues : ((a - 1) % 2 == 0 ? kBufA : kBufB);
      pi.status_sel = a % 2;
      pi.ordinal = a;
      st->pass[p] = pi;
      if (pi.active) ++a;
    }
    st->num_active = a;
    st->final_buf = (a == 0) ? kBufValues : ((a - 1) % 2 == 0 ? kBufA : kBufB);
    counts->bound = st->maxval;
    counts->passes = a;
  }
}

// ---------------------------------------------------------------------------
// One stable LSD radix pass over 8-bit digit p (Onesweep-style).
// Tile = THREADS x ITEMS keys; wave w ranks keys [w*64*ITEMS, +64*ITEMS) of
// the tile in (item, lane) order, which is the input order, so equal digits
// keep their relative order (stability).  Keys and then rids are staged in
// one LDS buffer in digit order and written out in runs per digit.
// ---------------------------------------------------------------------------
template <int THREADS, int ITEMS>
__global__ __launch_bounds__(THREADS, WAH_SORT_WAVES_PER_EU) void k_onesweep(
    int p, const uint32_t* __restrict__ values, uint32_t* keysA, uint32_t* ridsA,
    uint32_t* keysB, uint32_t* ridsB, uint64_t n, uint32_t rid_base,
    SortState* __restrict__ st, uint32_t* status_base, uint64_t status_stride,
    uint32_t num_tiles, uint32_t lanes, uint32_t* fault) {
  constexpr int TILE = THREADS * ITEMS;
  constexpr int WAVES = THREADS / kWave;
  static_assert(THREADS >= 256 && THREADS % 256 == 0, "one thread per digit");
  const PassInfo info = st->pass[p];
  if (!info.active) return;
  __shared__ uint32_t s_stage[TILE];
  __shared__ uint8_t s_digit[TILE];
  __shared__ uint32_t s_wcnt[WAVES][256];
  __shared__ uint32_t s_excl[256];
  __shared__ uint32_t s_glob[256];
  __shared__ uint32_t s_dsum[4];
  __shared__ uint32_t s_tile;

  const bool implicit = info.src == kBufValues;
  const uint32_t* src_keys = implicit ? values : (info.src == kBufA ? keysA : keysB);
  const uint32_t* src_rids = info.src == kBufA ? ridsA : ridsB;
  uint32_t* dst_keys = info.dst == kBufA ? keysA : keysB;
  uint32_t* dst_rids = info.dst == kBufA ? ridsA : ridsB;
  uint32_t* status = status_base + (uint64_t)info.status_sel * status_stride;
  uint32_t* status_next = status_base + (uint64_t)(1u - info.status_sel) * status_stride;
  const uint32_t shift = info.shift;
  const int t = threadIdx.x, w = t / kWave, l = t % kWave;
  const uint32_t goff = t < 256 ? st->digit_offset[p][t] : 0u;

  uint32_t iter = 0;
  for (;;) {
    for (int i = t; i < WAVES * 256; i += THREADS) (&s_wcnt[0][0])[i] = 0;
    const uint32_t tile = next_tile(st->tiles[p], lanes, &s_tile, iter);
    if (tile >= num_tiles) break;
    const uint64_t base = (uint64_t)tile * TILE;
    if (t < 256) status_next[(uint64_t)tile * 256 + t] = 0;  // next active pass's state

    // Keys only for now: the rids are loaded once the keys are staged, so
    // they never occupy registers during ranking and look-back.
    uint32_t key[ITEMS], rank[ITEMS];
    const uint64_t wbase = base + (uint64_t)w * kWave * ITEMS + l;
    const bool full = base + TILE <= n;
    const uint32_t nvalid = full ? (uint32_t)ITEMS * kWave : (uint32_t)(n > wbase ? n - wbase : 0);
    {
      const uint32_t* kp = src_keys + wbase;  // one base, immediate offsets
      if (full) {
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) key[i] = kp[i * kWave];
      } else {
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) key[i] = (uint32_t)(i * kWave) < nvalid ? kp[i * kWave] : 0u;
      }
    }
    // Rank: peers = lanes with the same digit (8 ballots); rank = peers
    // below this lane + keys of that digit the wave ranked before.
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      if ((uint32_t)(i * kWave) < nvalid) {
        const uint32_t d = (key[i] >> shift) & 255u;
        uint64_t peers = ballot(true);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
          const bool bit = (d >> b) & 1u;
          const uint64_t bb = ballot(bit);
          peers &= bit ? bb : ~bb;
        }
        const uint32_t cnt = (uint32_t)__popcll(peers);
        const uint32_t below = mbcnt(peers);
        const uint32_t prior = s_wcnt[w][d];
        rank[i] = prior + below;
        if (below == cnt - 1) s_wcnt[w][d] = prior + cnt;
      }
    }
    __syncthreads();
    // Threads 0..255 own one digit each: wave prefixes, tile total,
    // exclusive scan over digits, look-back.
    if (t < 256) {
      uint32_t total = 0;
#pragma unroll
      for (int i = 0; i < WAVES; ++i) {
        const uint32_t c = s_wcnt[i][t];
        s_wcnt[i][t] = total;
        total += c;
      }
      uint32_t* my = &status[(uint64_t)tile * 256 + t];
      st_agent(my, ((tile == 0 ? kFlagPrefix : kFlagAgg) << 30) | total);
      const uint32_t inc = wave_incl_add(total);
      if (l == kWave - 1) s_dsum[w] = inc;
      s_excl[t] = inc - total;  // within this wave's 64 digits
      uint32_t prev = 0;
      if (tile > 0) {
        prev = lookback_digit(status + t, tile, 256, fault);
        st_agent(my, (kFlagPrefix << 30) | (prev + total));
      }
      s_glob[t] = goff + prev;  // the tile-local offset is subtracted below
    }
    __syncthreads();
    if (t < 256) {
      uint32_t before = 0;
#pragma unroll
      for (int i = 0; i < 4; ++i) before += i < w ? s_dsum[i] : 0u;
      const uint32_t ex = s_excl[t] + before;
      s_excl[t] = ex;
      s_glob[t] -= ex;  // global index = s_glob[d] + local index
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      if ((uint32_t)(i * kWave) < nvalid) {
        const uint32_t d = (key[i] >> shift) & 255u;
        const uint32_t pos = s_excl[d] + s_wcnt[w][d] + rank[i];
        s_stage[pos] = key[i];
        s_digit[pos] = (uint8_t)d;
        rank[i] = pos;  // keep the LDS slot for the rids
      }
    }
    // rids: issued now, in flight while the keys go out
    uint32_t rid[ITEMS];
    if (implicit) {
#pragma unroll
      for (int i = 0; i < ITEMS; ++i) rid[i] = rid_base + (uint32_t)wbase + (uint32_t)(i * kWave);
    } else {
      const uint32_t* rp = src_rids + wbase;
#pragma unroll
      for (int i = 0; i < ITEMS; ++i) rid[i] = (uint32_t)(i * kWave) < nvalid ? rp[i * kWave] : 0u;
    }
    __syncthreads();
    const uint32_t tile_n = (n - base) < (uint64_t)TILE ? (uint32_t)(n - base) : (uint32_t)TILE;
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
      const uint32_t q = (uint32_t)j * THREADS + t;
      if (q < tile_n) {
        const uint32_t g = s_glob[s_digit[q]] + q;
        if (g < n) dst_keys[g] = s_stage[q];
        else atomicOr(fault, 2u);
      }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < ITEMS; ++i)
      if ((uint32_t)(i * kWave) < nvalid) s_stage[rank[i]] = rid[i];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
      const uint32_t q = (uint32_t)j * THREADS + t;
      if (q < tile_n) {
        const uint32_t g = s_glob[s_digit[q]] + q;
        if (g < n) dst_rids[g] = s_stage[q];
      }
    }
    __syncthreads();
  }
}

// Selects the sorted arrays the tail reads (after the active passes).
struct SortedView {
  const uint32_t* keys;
  const uint32_t* rids;
  bool implicit;
};
__device__ __forceinline__ SortedView sorted_view(const SortState* st, const uint32_t* values,
                                                  const uint32_t* keysA, const uint32_t* ridsA,
                                                  const uint32_t* keysB, const uint32_t* ridsB) {
  SortedView v;
  const uint32_t f = st->final_buf;
  v.implicit = f == kBufValues;
  v.keys = f == kBufValues ? values : (f == kBufA ? keysA : keysB);
  v.rids = f == kBufA ? ridsA : ridsB;
  return v;
}

// ---------------------------------------------------------------------------
// Fused tail over the sorted (value, rid) pairs.  A tile of THREADS x ITEMS
// pairs is staged in LDS; wave w owns pairs [w*64*ITEMS, +64*ITEMS) in (item,
// lane) order, so lanes read consecutive LDS words.  Per pair: chid = rid/31;
// an entry starts where (value, chid) changes and ORs the literal bits of its
// run (at most 31 pairs, read ahead in LDS); a key starts where value
// changes.  fill = chid for a key start, else chid - chid_prev - 1; an entry
// emits [fill if nonzero, literal].  Word and key positions: DPP wave scans of
// the packed (keys << 16 | words) counts, then a wave-parallel look-back
// across tiles.  Words are staged in LDS and stored contiguously.
// ---------------------------------------------------------------------------
template <int THREADS, int ITEMS>
__global__ __launch_bounds__(THREADS) void k_tail(
    const SortState* __restrict__ st, const uint32_t* values, const uint32_t* keysA,
    const uint32_t* ridsA, const uint32_t* keysB, const uint32_t* ridsB, uint64_t n,
    uint32_t rid_base, uint32_t* __restrict__ out_keys, uint32_t* __restrict__ out_offsets,
    uint32_t* __restrict__ out_words, uint64_t* status, uint32_t num_tiles,
    uint32_t* tile_counter, uint32_t lanes, wahgpu_counts_dev* __restrict__ counts) {
  constexpr int TILE = THREADS * ITEMS;
  constexpr int WAVES = THREADS / kWave;
  constexpr int SUB = kWave * ITEMS;     // pairs per wave
  constexpr int PAD = 4;                 // s_k[PAD - 1] holds pair base - 1
  constexpr int HALO = 32;               // pairs read past the tile by runs
  constexpr int SPAN = TILE + PAD + HALO;
  static_assert(TILE % 4 == 0 && TILE <= 32768, "word counts fit 16 bits");
  __shared__ __align__(16) uint32_t s_buf[2 * SPAN];
  __shared__ uint32_t s_wtot[WAVES];
  __shared__ uint64_t s_prefix;
  __shared__ uint32_t s_tile;
  uint32_t* s_k = s_buf;
  uint32_t* s_r = s_buf + SPAN;
  uint32_t* s_words = s_buf;  // reused once the pairs are consumed (2 * TILE words)
  const SortedView sv = sorted_view(st, values, keysA, ridsA, keysB, ridsB);
  const int t = threadIdx.x, w = t / kWave, l = t % kWave;
  uint64_t entries = 0;

  uint32_t iter = 0;
  for (;;) {
    const uint32_t tile = next_tile(tile_counter, lanes, &s_tile, iter);
    if (tile >= num_tiles) break;
    const uint64_t base = (uint64_t)tile * TILE;
    // stage pairs [base - 1, base + TILE + HALO - 1) with 16-byte loads
    if (base + TILE <= n) {
      const uint4* k4 = reinterpret_cast<const uint4*>(sv.keys + base);
      const uint4* r4 = reinterpret_cast<const uint4*>(sv.rids + base);
      for (int j = t; j < TILE / 4; j += THREADS) {
        reinterpret_cast<uint4*>(s_k + PAD)[j] = k4[j];
        uint4 rr;
        if (sv.implicit) {
          const uint32_t r0 = rid_base + (uint32_t)(base + 4 * (uint64_t)j);
          rr = make_uint4(r0, r0 + 1, r0 + 2, r0 + 3);
        } else {
          rr = r4[j];
        }
        reinterpret_cast<uint4*>(s_r + PAD)[j] = rr;
      }
    } else {
      for (int j = t; j < TILE; j += THREADS) {
        const uint64_t g = base + j;
        if (g < n) {
          s_k[PAD + j] = sv.keys[g];
          s_r[PAD + j] = sv.implicit ? rid_base + (uint32_t)g : sv.rids[g];
        }
      }
    }
    if (t < HALO) {
      const uint64_t g = base + TILE + t;
      if (g < n) {
        s_k[PAD + TILE + t] = sv.keys[g];
        s_r[PAD + TILE + t] = sv.implicit ? rid_base + (uint32_t)g : sv.rids[g];
      }
    }
    if (t == THREADS - 1 && base > 0) {
      s_k[PAD - 1] = sv.keys[base - 1];
      s_r[PAD - 1] = sv.implicit ? rid_base + (uint32_t)(base - 1) : sv.rids[base - 1];
    }
    __syncthreads();

    uint32_t fill[ITEMS], lit[ITEMS], key[ITEMS], ex[ITEMS];
    uint32_t khmask = 0;
    uint32_t carry = 0;  // (keys << 16 | words) of this wave's earlier rows
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      const int j = w * SUB + i * kWave + l;
      const uint64_t g = base + j;
      const int s = PAD + j;
      fill[i] = 0;
      lit[i] = 0;
      key[i] = 0;
      uint32_t p = 0;
      if (g < n) {
        const uint32_t v = s_k[s], r = s_r[s], c = r / 31u;
        const bool first = g == 0;
        const uint32_t cp = first ? 0u : s_r[s - 1] / 31u;
        const bool kh = first || v != s_k[s - 1];
        if (kh || c != cp) {
          uint32_t x = 0x80000000u | (1u << (r % 31u));
          for (int m = 1; m < 31; ++m) {
            if (g + m >= n) break;
            const uint32_t r2 = s_r[s + m];
            if (s_k[s + m] != v || r2 / 31u != c) break;
            x |= 1u << (r2 % 31u);
          }
          lit[i] = x;
          fill[i] = kh ? c : c - cp - 1u;
          p = 1u + (fill[i] != 0u);
          ++entries;
        }
        if (kh) {
          key[i] = v;
          khmask |= 1u << i;
          p |= 1u << 16;
        }
      }
      const uint32_t inc = wave_incl_add(p);
      ex[i] = carry + inc - p;
      carry += readlane(inc, kWave - 1);
    }
    if (l == 0) s_wtot[w] = carry;
    __syncthreads();
    uint32_t woff = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < WAVES; ++i) {
      woff += i < w ? s_wtot[i] : 0u;
      tot += s_wtot[i];
    }
    const uint64_t tile_agg = ((uint64_t)(tot >> 16) << 31) | (tot & 0xFFFFu);
    if (w == 0) {
      const uint64_t pre = lookback_wave_sum(status, tile, tile_agg, &counts->fault);
      if (l == 0) s_prefix = pre;
    }
    __syncthreads();  // every read of the staged pairs is done
    const uint64_t pre = s_prefix;
    const uint32_t Wpre = (uint32_t)(pre & 0x7FFFFFFFu), Kpre = (uint32_t)(pre >> 31);
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      const uint32_t e = woff + ex[i];
      uint32_t pos = e & 0xFFFFu;
      if ((khmask >> i) & 1u) {
        const uint64_t kidx = (uint64_t)Kpre + (e >> 16);
        if (kidx < n) {
          out_keys[kidx] = key[i];
          out_offsets[kidx] = Wpre + pos;
        } else {
          atomicOr(&counts->fault, 2u);
        }
      }
      if (lit[i]) {
        if (fill[i]) s_words[pos++] = fill[i];
        s_words[pos] = lit[i];
      }
    }
    __syncthreads();
    const uint32_t Wt = tot & 0xFFFFu;
    for (uint32_t j = t; j < Wt; j += THREADS) {
      const uint64_t o = (uint64_t)Wpre + j;
      if (o < 2 * n) out_words[o] = s_words[j];
      else atomicOr(&counts->fault, 2u);
    }
    if (tile == num_tiles - 1 && t == 0) {
      counts->length = (uint64_t)Wpre + Wt;
      counts->keycnt = (uint64_t)Kpre + (tot >> 16);
    }
    __syncthreads();
  }
  // one update of k per wave
  entries = wave_sum64(entries);
  if (l == 0 && entries) atomicAdd((unsigned long long*)&counts->k, (unsigned long long)entries);
}

// lengths[i] = offsets[i+1] - offsets[i]; the last key runs to L.
__global__ __launch_bounds__(kBlock) void k_lengths(const uint32_t* __restrict__ offsets,
                                                    uint32_t* __restrict__ lengths,
                                                    const wahgpu_counts_dev* __restrict__ counts,
                                                    uint64_t cap) {
  const uint64_t K = counts->keycnt < cap ? counts->keycnt : cap, L = counts->length;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < K;
       i += (uint64_t)gridDim.x * kBlock) {
    const uint32_t end = (i + 1 < K) ? offsets[i + 1] : (uint32_t)L;
    lengths[i] = end - offsets[i];
  }
}

// ---------------------------------------------------------------------------
// Staged kernels (one per reference stage) and primitives.
// ---------------------------------------------------------------------------
// Copy the sorted result out of the pass buffers (stage API output).
__global__ __launch_bounds__(kBlock) void k_sort_finalize(
    const SortState* __restrict__ st, const uint32_t* values, const uint32_t* keysA,
    const uint32_t* ridsA, const uint32_t* keysB, const uint32_t* ridsB, uint64_t n,
    uint32_t rid_base, uint32_t* out_keys, uint32_t* out_rids) {
  const SortedView sv = sorted_view(st, values, keysA, ridsA, keysB, ridsB);
  if (sv.keys == out_keys && !sv.implicit) return;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * kBlock) {
    out_keys[i] = sv.keys[i];
    out_rids[i] = sv.implicit ? rid_base + (uint32_t)i : sv.rids[i];
  }
}

// include/produce_chunk_id_literals.cl:17-24
__global__ __launch_bounds__(kBlock) void k_chunks(const uint32_t* __restrict__ rids, uint64_t n,
                                                   uint32_t* __restrict__ chids,
                                                   uint32_t* __restrict__ lits) {
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * kBlock) {
    uint32_t r = rids[i];
    lits[i] = 0x80000000u | (1u << (r % 31u));
    chids[i] = r / 31u;
  }
}

// include/produce_fills.cl:24-43
__global__ __launch_bounds__(kBlock) void k_fills(const uint32_t* __restrict__ vals,
                                                  const uint32_t* __restrict__ chids, uint64_t k,
                                                  uint32_t* __restrict__ fills) {
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < k;
       i += (uint64_t)gridDim.x * kBlock) {
    uint32_t c = chids[i];
    fills[i] = (i > 0 && vals[i] == vals[i - 1]) ? c - chids[i - 1] - 1u : c;
  }
}

// Reduce-by-key OR over runs of equal (value, chid), any run length.  The
// last element of each run writes (value, chid, OR) at index heads-1.
template <int ITEMS>
__global__ __launch_bounds__(kBlock) void k_merge(
    const uint32_t* __restrict__ vals, const uint32_t* __restrict__ chids,
    const uint32_t* __restrict__ lits, uint64_t n, uint32_t* __restrict__ out_v,
    uint32_t* __restrict__ out_c, uint32_t* __restrict__ out_l, uint64_t* status,
    uint32_t num_tiles, uint32_t* tile_counter, uint32_t lanes, uint64_t* result,
    uint32_t* fault) {
  constexpr int TILE = kBlock * ITEMS;
  __shared__ uint64_t s_scan[kWaves];
  __shared__ uint64_t s_prefix;
  __shared__ uint32_t s_tile;
  const int t = threadIdx.x;
  uint32_t iter = 0;
  for (;;) {
    const uint32_t tile = next_tile(tile_counter, lanes, &s_tile, iter);
    if (tile >= num_tiles) break;
    const uint64_t g0 = (uint64_t)tile * TILE + (uint64_t)t * ITEMS;
    uint32_t headmask = 0;
    uint64_t agg = OpSegOr::identity();
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      uint64_t g = g0 + i;
      if (g < n) {
        bool head = g == 0 || vals[g] != vals[g - 1] || chids[g] != chids[g - 1];
        if (head) headmask |= 1u << i;
        agg = OpSegOr::combine(agg, ((uint64_t)head << 32) | lits[g]);
      }
    }
    uint64_t tile_agg;
    uint64_t ex = block_exclusive<OpSegOr>(agg, s_scan, &tile_agg);
    if (t == 0) s_prefix = lookback64<OpSegOr>(status, tile, tile_agg, fault);
    __syncthreads();
    uint64_t run = OpSegOr::combine(s_prefix, ex);
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      uint64_t g = g0 + i;
      if (g < n) {
        bool head = (headmask >> i) & 1u;
        run = OpSegOr::combine(run, ((uint64_t)head << 32) | lits[g]);
        bool last = g + 1 == n || vals[g + 1] != vals[g] || chids[g + 1] != chids[g];
        if (last) {
          uint64_t idx = (run >> 32) - 1;
          if (idx < n) {
            out_v[idx] = vals[g];
            out_c[idx] = chids[g];
            out_l[idx] = (uint32_t)run;
          } else {
            atomicOr(fault, 2u);
          }
        }
      }
    }
    if (tile == num_tiles - 1 && t == 0)
      *result = (OpSegOr::combine(s_prefix, tile_agg) >> 32);
    __syncthreads();
  }
}

// Order-preserving compaction.  Mode 0: keep data[i] where valid[i] != 0
// (include/stream_compaction.cl:124-160).  Mode 1: the interleaved words
// (a[i], b[i]) with zero words dropped (fuse_fill_literals).
template <int ITEMS, int MODE>
__global__ __launch_bounds__(kBlock) void k_compact(
    const uint32_t* __restrict__ a, const uint32_t* __restrict__ b, uint64_t n,
    uint32_t* __restrict__ out, uint64_t* status, uint32_t num_tiles,
    uint32_t* tile_counter, uint32_t lanes, uint64_t* result, uint32_t* fault) {
  constexpr int TILE = kBlock * ITEMS;
  const uint64_t cap = MODE == 0 ? n : 2 * n;
  __shared__ uint64_t s_scan[kWaves];
  __shared__ uint64_t s_prefix;
  __shared__ uint32_t s_tile;
  const int t = threadIdx.x;
  uint32_t iter = 0;
  for (;;) {
    const uint32_t tile = next_tile(tile_counter, lanes, &s_tile, iter);
    if (tile >= num_tiles) break;
    const uint64_t g0 = (uint64_t)tile * TILE + (uint64_t)t * ITEMS;
    uint64_t cnt = 0;
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      uint64_t g = g0 + i;
      if (g < n) {
        if (MODE == 0) cnt += b[g] != 0u;
        else cnt += (a[g] != 0u) + (b[g] != 0u);
      }
    }
    uint64_t tile_agg;
    uint64_t ex = block_exclusive<OpAdd>(cnt, s_scan, &tile_agg);
    if (t == 0) s_prefix = lookback64<OpAdd>(status, tile, tile_agg, fault);
    __syncthreads();
    uint64_t pos = s_prefix + ex;
    if (pos + cnt > cap) {  // never for a consistent scan
      atomicOr(fault, 2u);
      cnt = 0;
      pos = 0;
    }
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      uint64_t g = g0 + i;
      if (g < n) {
        if (cnt == 0) break;
        if (MODE == 0) {
          if (b[g] != 0u) out[pos++] = a[g];
        } else {
          uint32_t x = a[g], y = b[g];
          if (x) out[pos++] = x;
          if (y) out[pos++] = y;
        }
      }
    }
    if (tile == num_tiles - 1 && t == 0) *result = s_prefix + tile_agg;
    __syncthreads();
  }
}

// compute_column_length over merged entries (values, fills): keys at value
// changes, offsets = exclusive scan of (1 + (fill != 0)).
template <int ITEMS>
__global__ __launch_bounds__(kBlock) void k_colen(
    const uint32_t* __restrict__ vals, const uint32_t* __restrict__ fills, uint64_t k,
    uint32_t* __restrict__ out_keys, uint32_t* __restrict__ out_offsets, uint64_t* status,
    uint32_t num_tiles, uint32_t* tile_counter, uint32_t lanes,
    wahgpu_counts_dev* __restrict__ counts, uint32_t* fault) {
  constexpr int TILE = kBlock * ITEMS;
  __shared__ uint64_t s_scan[kWaves];
  __shared__ uint64_t s_prefix;
  __shared__ uint32_t s_tile;
  const int t = threadIdx.x;
  uint32_t iter = 0;
  for (;;) {
    const uint32_t tile = next_tile(tile_counter, lanes, &s_tile, iter);
    if (tile >= num_tiles) break;
    const uint64_t g0 = (uint64_t)tile * TILE + (uint64_t)t * ITEMS;
    uint32_t W = 0, KH = 0;
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      uint64_t g = g0 + i;
      if (g < k) {
        W += 1u + (fills[g] != 0u);
        KH += (g == 0 || vals[g] != vals[g - 1]);
      }
    }
    uint64_t tile_agg;
    uint64_t ex = block_exclusive<OpAdd>(((uint64_t)KH << 31) | W, s_scan, &tile_agg);
    if (t == 0) s_prefix = lookback64<OpAdd>(status, tile, tile_agg, fault);
    __syncthreads();
    const uint64_t pre = s_prefix + ex;
    uint32_t pos = (uint32_t)(pre & 0x7FFFFFFFu);
    uint32_t kidx = (uint32_t)(pre >> 31);
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      uint64_t g = g0 + i;
      if (g < k) {
        if (g == 0 || vals[g] != vals[g - 1]) {
          if (kidx < k) {
            out_keys[kidx] = vals[g];
            out_offsets[kidx] = pos;
          } else {
            atomicOr(fault, 2u);
          }
          ++kidx;
        }
        pos += 1u + (fills[g] != 0u);
      }
    }
    if (tile == num_tiles - 1 && t == 0) {
      uint64_t tot = s_prefix + tile_agg;
      counts->length = tot & 0x7FFFFFFFu;
      counts->keycnt = tot >> 31;
    }
    __syncthreads();
  }
}

// Exclusive uint32 sum scan (include/scan.cl) and segmented inclusive sum
// scan (include/segmented_scan.cl), single pass with look-back.
template <int ITEMS, bool SEGMENTED>
__global__ __launch_bounds__(kBlock) void k_scan(
    const uint32_t* __restrict__ in, const uint32_t* __restrict__ heads, uint64_t n,
    uint32_t* __restrict__ out, uint64_t* status, uint32_t num_tiles,
    uint32_t* tile_counter, uint32_t lanes, uint32_t* fault) {
  constexpr int TILE = kBlock * ITEMS;
  __shared__ uint64_t s_scan[kWaves];
  __shared__ uint64_t s_prefix;
  __shared__ uint32_t s_tile;
  using Op = typename std::conditional<SEGMENTED, OpSegAdd, OpAdd32>::type;
  const int t = threadIdx.x;
  uint32_t iter = 0;
  for (;;) {
    const uint32_t tile = next_tile(tile_counter, lanes, &s_tile, iter);
    if (tile >= num_tiles) break;
    const uint64_t g0 = (uint64_t)tile * TILE + (uint64_t)t * ITEMS;
    uint64_t agg = Op::identity();
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      uint64_t g = g0 + i;
      if (g < n) {
        uint64_t e = in[g];
        if (SEGMENTED) e |= (uint64_t)(g == 0 || heads[g] != 0u) << 32;
        agg = Op::combine(agg, e);
      }
    }
    uint64_t tile_agg;
    uint64_t ex = block_exclusive<Op>(agg, s_scan, &tile_agg);
    if (t == 0) s_prefix = lookback64<Op>(status, tile, tile_agg, fault);
    __syncthreads();
    uint64_t run = Op::combine(s_prefix, ex);
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      uint64_t g = g0 + i;
      if (g < n) {
        uint64_t e = in[g];
        if (SEGMENTED) {
          e |= (uint64_t)(g == 0 || heads[g] != 0u) << 32;
          run = Op::combine(run, e);
          out[g] = (uint32_t)run;
        } else {
          out[g] = (uint32_t)run;
          run = Op::combine(run, e);
        }
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
static void launch_onesweep(const BuildArgs& a, int p, hipStream_t s) {
  const uint32_t tiles = (uint32_t)((a.n + kSortTile - 1) / kSortTile);
  const uint32_t cap = WAH_STATIC_TILES ? tiles : a.max_resident * 256 / kSortThreads;
  uint32_t grid = tiles < cap ? tiles : cap;
  if (grid == 0) grid = 1;
  hipLaunchKernelGGL((k_onesweep<kSortThreads, kSortItems>), dim3(grid), dim3(kSortThreads), 0,
                     s, p, a.values, a.keysA, a.ridsA, a.keysB, a.ridsB, a.n, a.rid_base, a.st,
                     a.sort_status, (uint64_t)tiles * 256, tiles, tile_lanes(grid),
                     &a.counts->fault);
}

uint64_t sort_status_words(uint64_t n) {
  const uint64_t tiles = (n + kSortTile - 1) / kSortTile;
  return 2 * tiles * 256;
}
uint64_t tail_status_words(uint64_t n) { return (n + kTailTile - 1) / kTailTile; }
uint64_t scan_status_words(uint64_t n) { return (n + kScanTile - 1) / kScanTile; }

void launch_sort(const BuildArgs& a, hipStream_t s, LaunchHook* hook) {
  uint32_t hb = (uint32_t)((a.n + 16383) / 16384);
  if (hb > kHistMaxBlocks) hb = kHistMaxBlocks;
  if (hb == 0) hb = 1;
  const int vec4 = ((uintptr_t)a.values % 16) == 0;
  const uint64_t tiles = (a.n + kSortTile - 1) / kSortTile;
  hipLaunchKernelGGL(k_fill32, dim3(64), dim3(kBlock), 0, s, reinterpret_cast<uint32_t*>(a.st),
                     kSortHeaderWords, 0u);
  hipLaunchKernelGGL(k_histogram, dim3(hb), dim3(kBlock), 0, s, a.values, a.n, vec4, a.st,
                     a.sort_status, tiles * 256, a.tail_status, tail_status_words(a.n),
                     a.counts);
  hipLaunchKernelGGL(k_plan, dim3(1), dim3(kBlock), 0, s, a.st, a.n, a.counts);
  if (hook) hook->mark(s);
  for (int p = 0; p < 4; ++p) {
    launch_onesweep(a, p, s);
    if (hook) hook->mark(s);
  }
}

void launch_tail(const BuildArgs& a, hipStream_t s, LaunchHook* hook) {
  const uint32_t tiles = (uint32_t)((a.n + kTailTile - 1) / kTailTile);
  const uint32_t cap = WAH_STATIC_TILES ? tiles : a.max_resident * 256 / kTailThreads;
  const uint32_t grid = tiles < cap ? tiles : cap;
  if (grid) {
    hipLaunchKernelGGL((k_tail<kTailThreads, kTailItems>), dim3(grid), dim3(kTailThreads), 0, s,
                       a.st, a.values, a.keysA, a.ridsA, a.keysB, a.ridsB, a.n, a.rid_base,
                       a.out_keys, a.out_offsets, a.out_words, a.tail_status, tiles,
                       a.st->tiles[kCounterTail], tile_lanes(grid), a.counts);
  }
  if (hook) hook->mark(s);
  uint32_t lg = (uint32_t)((a.n + kBlock - 1) / kBlock);
  if (lg > 1024) lg = 1024;
  if (lg == 0) lg = 1;
  hipLaunchKernelGGL(k_lengths, dim3(lg), dim3(kBlock), 0, s, a.out_offsets, a.out_lengths,
                     a.counts, a.n);
}

void launch_sort_finalize(const BuildArgs& a, hipStream_t s) {
  uint32_t g = (uint32_t)((a.n + kBlock - 1) / kBlock);
  if (g > 4096) g = 4096;
  if (g == 0) return;
  hipLaunchKernelGGL(k_sort_finalize, dim3(g), dim3(kBlock), 0, s, a.st, a.values, a.keysA,
                     a.ridsA, a.keysB, a.ridsB, a.n, a.rid_base, a.sorted_keys,
                     a.sorted_rids);
}

static uint32_t grid_for(uint64_t n) {
  uint64_t g = (n + kBlock - 1) / kBlock;
  if (g > 8192) g = 8192;
  return g ? (uint32_t)g : 1u;
}

void launch_fill32(uint32_t* p, uint64_t words, uint32_t v, hipStream_t s) {
  if (!words) return;
  hipLaunchKernelGGL(k_fill32, dim3(grid_for(words)), dim3(kBlock), 0, s, p, words, v);
}

void launch_chunks(const uint32_t* rids, uint64_t n, uint32_t* chids, uint32_t* lits,
                   hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_chunks, dim3(grid_for(n)), dim3(kBlock), 0, s, rids, n, chids, lits);
}

void launch_fills(const uint32_t* vals, const uint32_t* chids, uint64_t k, uint32_t* fills,
                  hipStream_t s) {
  if (!k) return;
  hipLaunchKernelGGL(k_fills, dim3(grid_for(k)), dim3(kBlock), 0, s, vals, chids, k, fills);
}

static uint32_t scan_grid(uint64_t n, uint32_t max_resident, uint32_t* tiles_out) {
  uint32_t tiles = (uint32_t)((n + kScanTile - 1) / kScanTile);
  *tiles_out = tiles;
  return tiles < max_resident ? tiles : max_resident;
}

void launch_merge(const ScanArgs& a, const uint32_t* vals, const uint32_t* chids,
                  const uint32_t* lits, uint64_t n, uint32_t* ov, uint32_t* oc, uint32_t* ol,
                  uint64_t* result, hipStream_t s) {
  uint32_t tiles, grid = scan_grid(n, a.max_resident, &tiles);
  if (!grid) return;
  hipLaunchKernelGGL(k_merge<kScanItems>, dim3(grid), dim3(kBlock), 0, s, vals, chids, lits, n,
                     ov, oc, ol, a.status, tiles, a.tile_counter, tile_lanes(grid), result,
                     a.fault);
}

void launch_compact(const ScanArgs& a, const uint32_t* data, const uint32_t* valid, uint64_t n,
                    int mode, uint32_t* out, uint64_t* result, hipStream_t s) {
  uint32_t tiles, grid = scan_grid(n, a.max_resident, &tiles);
  if (!grid) return;
  if (mode == 0)
    hipLaunchKernelGGL((k_compact<kScanItems, 0>), dim3(grid), dim3(kBlock), 0, s, data, valid,
                       n, out, a.status, tiles, a.tile_counter, tile_lanes(grid), result,
                       a.fault);
  else
    hipLaunchKernelGGL((k_compact<kScanItems, 1>), dim3(grid), dim3(kBlock), 0, s, data, valid,
                       n, out, a.status, tiles, a.tile_counter, tile_lanes(grid), result,
                       a.fault);
}

void launch_colen(const ScanArgs& a, const uint32_t* vals, const uint32_t* fills, uint64_t k,
                  uint32_t* keys, uint32_t* offsets, uint32_t* lengths,
                  wahgpu_counts_dev* counts, hipStream_t s) {
  uint32_t tiles, grid = scan_grid(k, a.max_resident, &tiles);
  if (grid)
    hipLaunchKernelGGL(k_colen<kScanItems>, dim3(grid), dim3(kBlock), 0, s, vals, fills, k, keys,
                       offsets, a.status, tiles, a.tile_counter, tile_lanes(grid), counts,
                       a.fault);
  uint32_t lg = grid_for(k);
  if (lg > 1024) lg = 1024;
  hipLaunchKernelGGL(k_lengths, dim3(lg), dim3(kBlock), 0, s, offsets, lengths, counts, k);
}

void launch_scan(const ScanArgs& a, const uint32_t* in, const uint32_t* heads, uint64_t n,
                 uint32_t* out, bool segmented, hipStream_t s) {
  uint32_t tiles, grid = scan_grid(n, a.max_resident, &tiles);
  if (!grid) return;
  if (segmented)
    hipLaunchKernelGGL((k_scan<kScanItems, true>), dim3(grid), dim3(kBlock), 0, s, in, heads, n,
                       out, a.status, tiles, a.tile_counter, tile_lanes(grid), a.fault);
  else
    hipLaunchKernelGGL((k_scan<kScanItems, false>), dim3(grid), dim3(kBlock), 0, s, in, heads, n,
                       out, a.status, tiles, a.tile_counter, tile_lanes(grid), a.fault);
}

}  // namespace wah
