// wah_kernels.hip -- gfx950 kernels of the WAH bitmap-index build.
//
// Fused build (what wahgpu_build and the batched builds run), for columns of
// uint32 values:
//   k_begin       zeroes the header and per-column state (and writes a single
//                 column's descriptor)
//   k_histogram   one read of the values: histograms of all four 8-bit
//                 digits, repeat rows, run counts per radix tile, the maximum;
//                 zeroing of look-back state
//   k_plan        per column: sort mode (pairs or runs), key width, digit
//                 offsets; passes whose digit is the same for every value are
//                 switched off (on the device, no host sync)
//   k_runcut      run mode: rows -> one (value, rid, literal bits) item per run
//   k_onesweep x4 stable LSD radix pass per active digit: LDS fetch-add (or
//                 ballot) ranking, LDS staging, decoupled look-back over tiles
//   k_tail        sorted items -> chunk ids, OR-merged literals, fills,
//                 zero-fill removal, word positions, keys and offsets
//   k_finish      per column: k, lengths[i] = offsets[i+1] - offsets[i], fault
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

// Zeroes the build header and the per-column histograms; a single-column
// build also gets its descriptor written here (no host-to-device copy).
__global__ __launch_bounds__(kBlock) void k_begin(Header* __restrict__ hdr,
                                                  ColState* __restrict__ cs, uint32_t cols_cap,
                                                  uint32_t* __restrict__ tile_runs,
                                                  uint32_t tiles_cap, DescHead* __restrict__ desc,
                                                  int inline_desc, Dims one_dims, Column one) {
  const uint64_t gt = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  const uint64_t gs = (uint64_t)gridDim.x * kBlock;
  for (uint64_t i = gt; i < tiles_cap; i += gs) tile_runs[i] = 0;
  uint32_t* h = reinterpret_cast<uint32_t*>(hdr);
  for (uint64_t i = gt; i < sizeof(Header) / 4; i += gs) h[i] = 0;
  const uint64_t zw = (uint64_t)cols_cap * kColZeroWords;
  for (uint64_t i = gt; i < zw; i += gs) {
    const uint64_t c = i / kColZeroWords, w = i % kColZeroWords;
    reinterpret_cast<uint32_t*>(&cs[c])[w] = 0;
  }
  if (inline_desc && gt == 0) {
    desc->dims = one_dims;
    desc_cols(desc)[0] = one;
  }
}

// ---------------------------------------------------------------------------
// Histogram of all four 8-bit digits per column (reads the values once).
// Block b takes a contiguous range of 4096-value chunks, accumulates in LDS
// and folds into the column's histogram when the column changes.
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

// Adds a value to the digit histograms, and to the repeat histograms when
// it repeats its predecessor inside one chunk (rep).
__device__ __forceinline__ void hist_value(uint32_t* h, uint32_t* hr, uint32_t v, bool valid,
                                           bool rep) {
  const uint32_t f = __builtin_amdgcn_readfirstlane(v);
  if (ballot(valid && v != f) == 0) {  // one value across the wave (runs)
    const uint32_t cnt = (uint32_t)__popcll(ballot(valid));
    const uint32_t rc = (uint32_t)__popcll(ballot(rep));
    if (lane_id() == 0 && cnt) {
#pragma unroll
      for (int p = 0; p < 4; ++p) {
        atomicAdd(&h[p * 256 + ((f >> (8 * p)) & 255u)], cnt);
        if (rc) atomicAdd(&hr[p * 256 + ((f >> (8 * p)) & 255u)], rc);
      }
    }
    return;
  }
#pragma unroll
  for (int p = 0; p < 4; ++p) hist_add(h, p * 256 + ((v >> (8 * p)) & 255u), valid);
  if (ballot(rep)) {
#pragma unroll
    for (int p = 0; p < 4; ++p) hist_add(hr, p * 256 + ((v >> (8 * p)) & 255u), rep);
  }
}

__device__ __forceinline__ void hist_flush(uint32_t* h, uint32_t* hr, uint32_t* s_max,
                                           uint32_t& vmax, ColState* col) {
  const int t = threadIdx.x;
  for (int d = 32; d >= 1; d >>= 1) {
    const uint32_t o = __shfl_xor(vmax, d, kWave);
    vmax = o > vmax ? o : vmax;
  }
  if ((t % kWave) == 0) s_max[t / kWave] = vmax;
  __syncthreads();
  for (int i = t; i < 1024; i += kBlock) {
    if (h[i]) atomicAdd(&col->hist[0][i], h[i]);
    if (hr[i]) atomicAdd(&col->rep[0][i], hr[i]);
    h[i] = 0;
    hr[i] = 0;
  }
  if (t == 0) {
    uint32_t m = 0;
    for (int i = 0; i < kWaves; ++i) m = s_max[i] > m ? s_max[i] : m;
    atomicMax(&col->maxval, m);
  }
  vmax = 0;
  __syncthreads();
}

// 8 waves per SIMD: the grid is sized for 8 resident blocks per CU
__global__ __launch_bounds__(kBlock, 8) void k_histogram(
    const uint32_t* __restrict__ values, const DescHead* __restrict__ desc,
    ColState* __restrict__ cs, uint32_t* __restrict__ status0,
    uint64_t* __restrict__ tail_status, uint32_t* __restrict__ tile_entries,
    uint32_t* __restrict__ tile_runs, uint32_t sort_tile) {
  __shared__ uint32_t h[4 * 256];
  __shared__ uint32_t hr[4 * 256];
  __shared__ uint32_t s_max[kWaves];
  const int t = threadIdx.x;
  const Dims dims = desc->dims;
  const Column* cols = desc_cols(desc);
  for (int i = t; i < 1024; i += kBlock) h[i] = hr[i] = 0;
  // zero the look-back state the later kernels of the same index build use
  const uint64_t gthreads = (uint64_t)gridDim.x * kBlock;
  const uint64_t gt = (uint64_t)blockIdx.x * kBlock + t;
  const uint64_t s0 = (uint64_t)dims.sort_tiles * 256;
  for (uint64_t i = gt; i < s0; i += gthreads) status0[i] = 0;
  for (uint64_t i = gt; i < dims.tail_tiles; i += gthreads) {
    tail_status[i] = 0;
    tile_entries[i] = 0;
  }
  __syncthreads();

  const uint32_t per = (dims.hist_chunks + gridDim.x - 1) / gridDim.x;
  uint32_t c0 = blockIdx.x * per;
  const uint32_t c1 = c0 + per < dims.hist_chunks ? c0 + per : dims.hist_chunks;
  if (c0 >= c1) return;
  uint32_t col = find_segment(&cols[0].hist_chunk0, sizeof(Column) / 4, dims.ncols, c0);
  uint32_t col_end = col + 1 < dims.ncols ? cols[col + 1].hist_chunk0 : dims.hist_chunks;
  uint32_t vmax = 0;
  for (uint32_t c = c0; c < c1; ++c) {
    if (c >= col_end) {  // next column (block-uniform)
      hist_flush(h, hr, s_max, vmax, &cs[col]);
      col = find_segment(&cols[0].hist_chunk0, sizeof(Column) / 4, dims.ncols, c);
      col_end = col + 1 < dims.ncols ? cols[col + 1].hist_chunk0 : dims.hist_chunks;
    }
    const Column cd = cols[col];
    const uint64_t lo = (uint64_t)(c - cd.hist_chunk0) * kHistChunk;
    const uint32_t cnt = cd.n - lo < kHistChunk ? (uint32_t)(cd.n - lo) : kHistChunk;
    const uint32_t* src = values + cd.off + lo;
    // row of src[i] is rid_base + lo + i; a row repeats its predecessor when
    // it is not the first of the column or of a chunk and has its value
    const uint32_t row0 = cd.rid_base + (uint32_t)lo;
    uint32_t reps = 0;  // rows of this chunk that repeat their predecessor
    if (cnt == kHistChunk && ((uintptr_t)src % 16) == 0) {
      // four 16-byte loads per thread, and per wave the value before each
      // of its four 256-value spans (a uniform load), all issued before any
      // is used: one memory latency per chunk
      const uint4* v4 = reinterpret_cast<const uint4*>(src);
      const uint32_t wv = __builtin_amdgcn_readfirstlane(t / kWave);
      uint4 q[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) q[u] = v4[u * kBlock + t];
      // lane u < 4: the value before span u of the wave (one vector load)
      uint32_t wprev;
      {
        const uint32_t lu = lane_id();
        const uint32_t s0 = (lu * kBlock + wv * kWave) * 4;  // span's first index
        wprev = lu < 4u && lo + s0 > 0 ? src[(int64_t)s0 - 1] : 0u;
      }
      __builtin_amdgcn_sched_barrier(0);
      // The wave's 1024 rows of the chunk (4 spans of 256).  A digit the same
      // in all 1024 rows adds once (the high digits of small values); a digit
      // that varies adds once per span that holds a single value (inside a
      // run), and otherwise one LDS atomic per row.  Lane 0 adds from its own
      // values, so no per-span scalars stay live.  The repeat histograms are
      // added only when some row of the wave repeats its predecessor (a
      // wave-uniform branch), so columns without repeats carry no per-row
      // conditions on their atomics.
      const uint32_t f0 = __builtin_amdgcn_readfirstlane(q[0].x);
      uint32_t x = 0;    // bits in which a value of the lane differs from its span's first
      uint32_t fx = 0;   // bits in which a span's first value differs from f0
      uint32_t one = 0;  // bit u: span u holds a single value
      uint32_t rm = 0;   // repeat rows of the lane: bit 4u + j
      // row % 31 of q[u].x; spans are 1024 rows apart and 1024 = 33 * 31 + 1
      uint32_t m = (row0 + (uint32_t)t * 4u) % 31u;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const uint32_t i0 = (uint32_t)(u * kBlock + t) * 4;  // index of q[u].x in src
        // lane l-1's last value (DPP wave shift); lane 0 the value before the span
        const uint32_t prev = (uint32_t)__builtin_amdgcn_update_dpp(
            (int)readlane(wprev, u), (int)q[u].w, 0x138, 0xF, 0xF, false);
        const uint32_t e[4] = {q[u].x, q[u].y, q[u].z, q[u].w};
        uint32_t mj = m;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          vmax = e[j] > vmax ? e[j] : vmax;
          const uint32_t pv = j == 0 ? prev : e[j - 1];
          const bool rep = lo + i0 + j > 0 && mj != 0u && e[j] == pv;
          rm |= rep ? 1u << (4 * u + j) : 0u;
          mj = mj == 30u ? 0u : mj + 1u;
        }
        m = m == 30u ? 0u : m + 1u;
        const uint32_t fu = __builtin_amdgcn_readfirstlane(e[0]);
        const uint32_t xu = (e[0] ^ fu) | (e[1] ^ fu) | (e[2] ^ fu) | (e[3] ^ fu);
        one |= ballot(xu != 0u) == 0 ? 1u << u : 0u;
        x |= xu;
        fx |= fu ^ f0;
      }
      reps += (uint32_t)__builtin_popcount(rm);
      // a bit varies over the 1024 rows iff it varies inside a span or
      // between the spans' first values
      const uint32_t vary = wave_or(x) | fx;
      if (vary == 0u) {
        // every row of the chunk holds f0 (inside a run): lane p adds the
        // wave's rows, and its repeat rows, to digit p's bin of f0
        const uint32_t rc = readlane(wave_incl_add((uint32_t)__builtin_popcount(rm)), kWave - 1);
        const uint32_t lp = lane_id();
        if (lp < 4u) {
          const uint32_t bin = lp * 256u + ((f0 >> (8u * lp)) & 255u);
          atomicAdd(&h[bin], 16u * kWave);
          if (rc) atomicAdd(&hr[bin], rc);
        }
      } else {
#pragma unroll
        for (int p = 0; p < 4; ++p) {
          if (((vary >> (8 * p)) & 255u) == 0u) {
            if (lane_id() == 0) atomicAdd(&h[p * 256 + ((q[0].x >> (8 * p)) & 255u)], 16u * kWave);
            continue;
          }
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const uint32_t e[4] = {q[u].x, q[u].y, q[u].z, q[u].w};
            if ((one >> u) & 1u) {
              if (lane_id() == 0) atomicAdd(&h[p * 256 + ((e[0] >> (8 * p)) & 255u)], 4u * kWave);
            } else {
#pragma unroll
              for (int j = 0; j < 4; ++j) atomicAdd(&h[p * 256 + ((e[j] >> (8 * p)) & 255u)], 1u);
            }
          }
        }
        if (ballot(rm != 0u)) {
          uint32_t rcu[4], rc = 0;  // repeats of each span and of the wave
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            rcu[u] = 0;
#pragma unroll
            for (int j = 0; j < 4; ++j) rcu[u] += (uint32_t)__popcll(ballot((rm >> (4 * u + j)) & 1u));
            rc += rcu[u];
          }
#pragma unroll
          for (int p = 0; p < 4; ++p) {
            if (((vary >> (8 * p)) & 255u) == 0u) {
              if (lane_id() == 0) atomicAdd(&hr[p * 256 + ((q[0].x >> (8 * p)) & 255u)], rc);
              continue;
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
              const uint32_t e[4] = {q[u].x, q[u].y, q[u].z, q[u].w};
              if ((one >> u) & 1u) {
                if (lane_id() == 0 && rcu[u]) atomicAdd(&hr[p * 256 + ((e[0] >> (8 * p)) & 255u)], rcu[u]);
              } else {
#pragma unroll
                for (int j = 0; j < 4; ++j)
                  if ((rm >> (4 * u + j)) & 1u) atomicAdd(&hr[p * 256 + ((e[j] >> (8 * p)) & 255u)], 1u);
              }
            }
          }
        }
      }
    } else {
      for (uint32_t b = 0; b < cnt; b += kBlock) {
        const uint32_t i = b + t;
        const bool valid = i < cnt;
        const uint32_t v = valid ? src[i] : 0;
        const bool has_prev = valid && lo + i > 0;
        const uint32_t pv = has_prev ? src[(int64_t)i - 1] : 0u;
        vmax = valid && v > vmax ? v : vmax;
        const bool rep = has_prev && (row0 + i) % 31u != 0u && v == pv;
        reps += rep ? 1u : 0u;
        hist_value(h, hr, v, valid, rep);
      }
    }
    // runs of the wave's rows, added to their radix tile (a run-mode
    // column's cut places its items from these counts); no block barrier,
    // so the waves of a block stream their chunks independently
    reps = readlane(wave_incl_add(reps), kWave - 1);  // at most 4096
    if ((t % kWave) == 0) {
      // rows this wave covered: the fast path gives wave w rows 4 x 64 x
      // {u * 256 + w * 64 ..}; the slow path row i to thread i % 256
      const uint32_t w = t / kWave;
      uint32_t rows = 0;
      if (cnt == kHistChunk && ((uintptr_t)src % 16) == 0) {
        rows = kHistChunk / kWaves;
      } else {
        for (uint32_t b = 0; b < cnt; b += kBlock) {
          const uint32_t lo_i = b + w * kWave;
          rows += cnt > lo_i ? (cnt - lo_i < (uint32_t)kWave ? cnt - lo_i : (uint32_t)kWave) : 0u;
        }
      }
      if (rows - reps) atomicAdd(&tile_runs[cd.sort_tile0 + (uint32_t)(lo / sort_tile)], rows - reps);
    }
  }
  hist_flush(h, hr, s_max, vmax, &cs[col]);
}

// After the histogram, per column: digit offsets and the pass plan (digits
// that are the same for every value of the column are skipped).  A kernel of
// its own so the histogram blocks need no device-scope fences (on gfx950
// each one writes back and invalidates the L2).
__global__ __launch_bounds__(kBlock) void k_plan(const DescHead* __restrict__ desc,
                                                 ColState* __restrict__ cs, Header* __restrict__ hdr,
                                                 uint32_t* __restrict__ tile_runs,
                                                 wahgpu_counts_dev* __restrict__ counts,
                                                 int force_pair, uint32_t sort_tile,
                                                 WideState* __restrict__ wide) {
  __shared__ uint64_t s_scan[kWaves];
  __shared__ int s_trivial[4];
  __shared__ uint32_t s_reps;
  __shared__ uint32_t s_wide;  // 1 + wide slot of a column that takes wide digits
  const int t = threadIdx.x;
  const uint32_t ncols = desc->dims.ncols;
  for (uint32_t c = blockIdx.x; c < ncols; c += gridDim.x) {
    const uint64_t n = desc_cols(desc)[c].n;
    ColState* col = &cs[c];
    if (t < 4) s_trivial[t] = 0;
    if (t == 0) s_reps = 0;
    __syncthreads();
    atomicAdd(&s_reps, col->rep[0][t]);
    __syncthreads();
    // Run mode when it sorts under 2/3 of the rows (an item moves 12 bytes
    // per pass, a pair 8).
    const uint64_t runs = n - s_reps;
    const bool run_mode = !force_pair && n > 0 && 3 * runs < 2 * n;
    const uint64_t items = run_mode ? runs : n;
    if (run_mode) {
      // where each radix tile's runs go: exclusive prefix of the runs per tile
      const Column cd = desc_cols(desc)[c];
      uint32_t* tr = tile_runs + cd.sort_tile0;
      const uint32_t tiles = sort_tiles_of(n, sort_tile);
      // thread t scans a strip of consecutive tiles; a strip of up to 64
      // tiles (2^28 rows in 16384-row tiles) stays in registers between the
      // sum and the prefix, longer ones are read twice in batches of 16
      const uint32_t per = (tiles + kBlock - 1) / kBlock;
      const uint32_t s0 = t * per < tiles ? t * per : tiles;
      const uint32_t s1 = s0 + per < tiles ? s0 + per : tiles;
      constexpr uint32_t kStrip = 64;
      if (per <= kStrip) {  // block-uniform
        uint32_t x[kStrip];
        uint64_t sum = 0;
#pragma unroll
        for (uint32_t i = 0; i < kStrip; ++i) x[i] = s0 + i < s1 ? tr[s0 + i] : 0u;
#pragma unroll
        for (uint32_t i = 0; i < kStrip; ++i) sum += x[i];
        uint64_t tot;
        uint64_t run = block_exclusive<OpAdd>(sum, s_scan, &tot);
#pragma unroll
        for (uint32_t i = 0; i < kStrip; ++i) {
          if (s0 + i < s1) tr[s0 + i] = (uint32_t)run;
          run += x[i];
        }
      } else {
        uint64_t sum = 0;
        for (uint32_t b = s0; b < s1; b += 16) {
          uint32_t x[16];
#pragma unroll
          for (int i = 0; i < 16; ++i) x[i] = b + i < s1 ? tr[b + i] : 0u;
#pragma unroll
          for (int i = 0; i < 16; ++i) sum += x[i];
        }
        uint64_t tot;
        uint64_t run = block_exclusive<OpAdd>(sum, s_scan, &tot);
        for (uint32_t b = s0; b < s1; b += 16) {
          uint32_t x[16];
#pragma unroll
          for (int i = 0; i < 16; ++i) x[i] = b + i < s1 ? tr[b + i] : 0u;
#pragma unroll
          for (int i = 0; i < 16; ++i) {
            if (b + i < s1) tr[b + i] = (uint32_t)run;
            run += x[i];
          }
        }
      }
    }
    for (int p = 0; p < 4; ++p) {
      const uint32_t h = col->hist[p][t] - (run_mode ? col->rep[p][t] : 0u);
      uint64_t total;
      const uint64_t ex = block_exclusive<OpAdd>(h, s_scan, &total);
      col->digit_offset[p][t] = (uint32_t)ex;
      if ((uint64_t)h == items) s_trivial[p] = 1;  // every item has digit t
    }
    __syncthreads();
    if (t == 0) {
      uint32_t a = 0;
      for (int p = 0; p < 4; ++p)
        if (n > 0 && !s_trivial[p]) col->shift[a++] = 8u * p;
      for (uint32_t q = a; q < 4; ++q) col->shift[q] = 0;
      // Wide digits when they take fewer passes: the key's significant bits
      // split as evenly as possible into passes of at most kWideBits.
      const Column cw = desc_cols(desc)[c];
      const uint32_t bits = col->maxval ? 32u - (uint32_t)__clz(col->maxval) : 0u;
      const uint32_t aw = (bits + kWideBits - 1) / kWideBits;
      const bool wide = cw.wide_slot != 0u && !run_mode && !force_pair && n >= kWideMinValues &&
                        aw < a;
      col->wide = wide ? 1u : 0u;
      col->wide_slot = wide ? cw.wide_slot - 1u : 0u;
      s_wide = wide ? cw.wide_slot : 0u;
      if (wide) {
        uint32_t at = 0;
        for (uint32_t i = 0; i < (uint32_t)kWidePasses; ++i) {
          const uint32_t b = i < aw ? bits / aw + (i < bits % aw ? 1u : 0u) : 0u;
          col->wbits[i] = b;
          col->wshift[i] = at;
          at += b;
        }
        a = aw;
      }
      col->num_active = a;
      col->mode = run_mode ? kModeRun : kModePair;
      // keys below 2^16 travel as 16-bit words from the first pass's output
      // on (the stage API reads 32-bit keys)
      col->narrow = !force_pair && a > 0 && col->maxval < 65536u;
      col->items = (uint32_t)items;
      // pair mode: values, then A, B, A, B; run mode: the cut writes B
      if (run_mode) col->final_buf = (a % 2) ? kBufA : kBufB;
      else col->final_buf = (a == 0) ? kBufValues : ((a - 1) % 2 == 0 ? kBufA : kBufB);
      if (a) atomicMax(wide ? &hdr->max_wide : &hdr->max_active[run_mode ? 1 : 0], a);
      if (n) atomicAdd(&hdr->mode_cols[run_mode ? 1 : 0], 1u);
      wahgpu_counts_dev z;
      z.k = 0; z.length = 0; z.keycnt = 0;
      z.bound = col->maxval; z.passes = a; z.fault = 0;
      z.mode = run_mode ? kModeRun : wide ? kModeWideReport : kModePair;
      counts[desc_cols(desc)[c].index] = z;
    }
    __syncthreads();
    if (s_wide) {  // the wide histograms start at zero (k_whist adds to them)
      uint32_t* hw = wide[s_wide - 1u].hist[0];
      for (uint32_t i = t; i < (uint32_t)kWidePasses * kWideBins; i += kBlock) hw[i] = 0;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Run mode, before the radix launches: cuts each run-mode column into runs
// (a row starts one unless it repeats its predecessor's value inside the
// same 31-row chunk) and writes one item per run, in row order, to buffer B:
// value, rid of its first row, literal bits of its rows.  An ordered stream
// compaction without look-back: the histogram counted the runs of each
// radix tile and k_plan turned the counts into per-tile offsets.  A block
// takes whole radix tiles (static grid stride) and walks each in steps of
// kBlock x 16 rows with a running offset.  Lane t owns 16 consecutive rows
// (16-byte loads); a run is at most 31 rows, so the run of a lane's last
// start ends within the next two lanes (or the next wave's first two, or
// the rows past the step, exchanged in LDS).  Items are staged in LDS and
// stored contiguously, one array at a time.
// ---------------------------------------------------------------------------
constexpr int kCutRows = 16;                    // rows per lane
constexpr int kCutStep = kBlock * kCutRows;     // rows per block step
__global__ __launch_bounds__(kBlock, 8) void k_runcut(
    const uint32_t* __restrict__ values, uint32_t* __restrict__ keysB,
    uint32_t* __restrict__ ridsB, uint32_t* __restrict__ litsB,
    const DescHead* __restrict__ desc, const ColState* __restrict__ cs,
    Header* __restrict__ hdr, const uint32_t* __restrict__ tile_runs, uint32_t sort_tile) {
  constexpr int R = kCutRows;
  constexpr uint32_t kAll = (1u << R) - 1u;
  __shared__ uint32_t s_stage[kCutStep];
  __shared__ uint32_t s_wsum[kWaves];
  __shared__ uint32_t s_wlead[kWaves + 1];  // rows before the first stop of wave w; [kWaves]: past the step
  const Dims dims = desc->dims;
  const Column* cols = desc_cols(desc);
  const int t = threadIdx.x, l = t % kWave;
  const int w = __builtin_amdgcn_readfirstlane(t / kWave);  // scalar: per-wave LDS addresses
  if (hdr->mode_cols[1] == 0) return;  // no run-mode column
  for (uint32_t tile = blockIdx.x; tile < dims.sort_tiles; tile += gridDim.x) {
    const uint32_t c = find_segment(&cols[0].sort_tile0, sizeof(Column) / 4, dims.ncols, tile);
    const ColState* col = &cs[c];
    if (col->mode != kModeRun) continue;  // block-uniform, no LDS touched
    const Column cd = cols[c];
    // column-local rows fit 32 bits (a column holds fewer than 2^30)
    const uint32_t n = (uint32_t)cd.n;
    const uint32_t t0 = (tile - cd.sort_tile0) * sort_tile;
    const uint32_t t1 = t0 + sort_tile < n ? t0 + sort_tile : n;
    const uint32_t* vp = values + cd.off;
    const bool aligned = ((uintptr_t)vp % 16) == 0;
    const uint32_t items = col->items;
    uint32_t off = tile_runs[tile];  // runs of the column's earlier rows
    for (uint32_t lo = t0; lo < t1; lo += kCutStep) {
      const uint32_t j0 = lo + (uint32_t)t * R;  // the lane's first row
      const uint32_t nv = t1 > j0 ? (t1 - j0 < (uint32_t)R ? t1 - j0 : (uint32_t)R) : 0u;
      uint32_t v[R];
      if (lo + kCutStep <= t1 && aligned) {
#pragma unroll
        for (int q = 0; q < R / 4; ++q) {
          const uint4 x = reinterpret_cast<const uint4*>(vp + j0)[q];
          v[4 * q] = x.x; v[4 * q + 1] = x.y; v[4 * q + 2] = x.z; v[4 * q + 3] = x.w;
        }
      } else {
#pragma unroll
        for (int i = 0; i < R; ++i) v[i] = (uint32_t)i < nv ? vp[j0 + i] : 0u;
      }
      // predecessor of the lane's first row: lane l-1's last, or for lane 0
      // the row before the wave (one load per wave)
      uint32_t pv;
      {
        const uint32_t gw = lo + (uint32_t)(w * kWave * R);
        const uint32_t p0 = gw > 0 && gw <= n ? vp[gw - 1] : 0u;
        pv = (uint32_t)__builtin_amdgcn_update_dpp((int)p0, (int)v[R - 1], 0x138, 0xF, 0xF, false);
      }
      const uint32_t m0 = (cd.rid_base + j0) % 31u;  // row % 31 of the lane's first row
      // Run starts, without branches: a row starts a run when it differs
      // from its predecessor, begins a chunk (row % 31 == 0: at most one of
      // the lane's R < 31 rows, the (31 - m0) % 31-th) or is the column's
      // first row; rows past the column's end start nothing.
      uint32_t sm;
      {
        const uint32_t ci = m0 == 0u ? 0u : 31u - m0;
        sm = (ci < (uint32_t)R ? 1u << ci : 0u) | (j0 == 0u ? 1u : 0u);
#pragma unroll
        for (int i = 0; i < R; ++i) {
          const uint32_t prev = i ? v[i - 1] : pv;
          sm |= (v[i] != prev ? 1u : 0u) << i;
        }
        sm &= nv >= (uint32_t)R ? kAll : (1u << nv) - 1u;
      }
      const uint32_t stops = sm | (~0u << nv & kAll);
      const uint32_t lead = stops ? (uint32_t)__builtin_ctz(stops) : (uint32_t)R;
      const uint32_t cnt = (uint32_t)__builtin_popcount(sm);
      const uint32_t inc = wave_incl_add(cnt);
      if (l == kWave - 1) s_wsum[w] = inc;
      const uint64_t hsm = ballot(stops != 0u);
      const uint32_t L1 = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)lead, 0x130, 0xF, 0xF, false);
      const uint32_t L2 = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)L1, 0x130, 0xF, 0xF, false);
      {
        // rows before the wave's first stop (within its first two lanes)
        const uint32_t l0 = readlane(lead, 0), l1 = readlane(lead, 1);
        if (l == 0) s_wlead[w] = (hsm & 1ull) ? l0 : (uint32_t)R + l1;
      }
      if (w == kWaves - 1) {  // rows past the step: where its last run ends
        const uint32_t g = lo + (uint32_t)kCutStep + l;
        const bool valid = l < 32 && g < n;
        const uint32_t x = valid ? vp[g] : 0u;
        const uint32_t last = readlane(v[R - 1], kWave - 1);
        const uint32_t px = (uint32_t)__builtin_amdgcn_update_dpp((int)last, (int)x, 0x138, 0xF, 0xF, false);
        const bool st = !valid | ((cd.rid_base + g) % 31u == 0u) | (x != px);  // no branches
        const uint64_t sb = ballot(st);  // lanes 32.. always stop
        if (l == 0) s_wlead[kWaves] = (uint32_t)__builtin_ctzll(sb);
      }
      __syncthreads();
      uint32_t woff, tot;
      {
        const uint32_t x = l < kWaves ? s_wsum[l] : 0u;
        const uint32_t xi = wave_incl_add(x);
        woff = readlane(xi - x, __builtin_amdgcn_readfirstlane(w));
        tot = readlane(xi, kWaves - 1);
      }
      // rows after the lane that the run of its last start covers
      const uint32_t nw = s_wlead[w + 1];
      uint32_t cont;
      if (l == kWave - 1) cont = nw;
      else if ((hsm >> (l + 1)) & 1ull) cont = L1;
      else if (l == kWave - 2) cont = (uint32_t)R + nw;
      else cont = (uint32_t)R + L2;
      if (nv < (uint32_t)R) cont = 0;  // the run ends with the rows
      const uint32_t k0 = woff + inc - cnt;  // items of the step before the lane
      if (off + tot > items) {
        if (t == 0) atomicOr(&hdr->fault, 2u);
        return;  // block-uniform: tile_runs disagree with the items
      }
      // the item of start i: round 0 its value, 1 its rid, 2 its literal bits
      auto item = [&](int round, int i) -> uint32_t {
        if (round == 0) return v[i];
        if (round == 1) return cd.rid_base + j0 + i;
        const uint32_t rest = stops >> (i + 1);
        const uint32_t len = rest ? (uint32_t)__builtin_ctz(rest) + 1u : (uint32_t)(R - i) + cont;
        const uint32_t mi = m0 + i >= 31u ? m0 + i - 31u : m0 + i;
        return ((len >= 32u ? 0u : (1u << len)) - 1u) << mi;
      };
      if (3 * tot <= (uint32_t)kCutStep) {
        // few runs: the three arrays staged together, one round
        uint32_t k = k0;
#pragma unroll
        for (int i = 0; i < R; ++i) {
          if ((sm >> i) & 1u) {
            s_stage[k] = item(0, i);
            s_stage[tot + k] = item(1, i);
            s_stage[2 * tot + k] = item(2, i);
            ++k;
          }
        }
        __syncthreads();
        for (uint32_t j = t; j < 3 * tot; j += kBlock) {
          const uint32_t r = j / tot, q = j - r * tot;  // tot > 0 here
          uint32_t* dst = (r == 0 ? keysB : r == 1 ? ridsB : litsB) + cd.off + off;
          dst[q] = s_stage[j];
        }
        __syncthreads();
      } else {
        // three rounds through LDS: values, rids, literal bits
#pragma unroll 1
        for (int round = 0; round < 3; ++round) {
          uint32_t k = k0;
#pragma unroll
          for (int i = 0; i < R; ++i)
            if ((sm >> i) & 1u) s_stage[k++] = item(round, i);
          __syncthreads();
          uint32_t* dst = (round == 0 ? keysB : round == 1 ? ridsB : litsB) + cd.off + off;
          for (uint32_t j = t; j < tot; j += kBlock) dst[j] = s_stage[j];
          __syncthreads();
        }
      }
      off += tot;
    }
  }
}

// Ranking of a radix pass.  kRankMatch: peers = lanes with the same digit
// from 8 ballots, rank = peers below + earlier count (stable by
// construction, ~50 VALU per key).  kRankAtomic: one LDS fetch-add per key on
// the wave's digit counter; the returned counts are a stable rank because the
// LDS serializes the lanes of one instruction that hit one address in lane
// order -- not a documented guarantee, so a context takes this path only
// after k_rank_selftest has checked it on the device.
constexpr int kRankMatch = 0;
constexpr int kRankAtomic = 1;

// Checks on this device that LDS fetch-adds of one wave instruction return,
// for lanes on the same address, counts in increasing lane order: many
// address patterns (1 to 64 lanes per address, every lane order), repeated.
// *ok = 1 if every check held.
__global__ __launch_bounds__(256) void k_rank_selftest(uint32_t* ok) {
  __shared__ uint32_t cnt[4][256];
  const int t = threadIdx.x, w = t / kWave, l = t % kWave;
  uint32_t good = 1;
  for (uint32_t round = 0; round < 512; ++round) {
    for (int i = t; i < 4 * 256; i += 256) (&cnt[0][0])[i] = 0;
    __syncthreads();
    // address pattern of this round: few to many distinct digits
    const uint32_t h = (uint32_t)l * 2654435761u ^ (round * 40503u + (uint32_t)w * 977u);
    const uint32_t span = 1u + (round % 64u);
    const uint32_t d = (h >> 7) % span;
    uint32_t prev = 0;
    for (int rep = 0; rep < 4; ++rep) {
      const uint32_t r = atomicAdd(&cnt[w][d], 1u);
      uint64_t peers = ballot(true);
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        const bool bit = (d >> b) & 1u;
        const uint64_t bb = ballot(bit);
        peers &= bit ? bb : ~bb;
      }
      const uint32_t expect = (uint32_t)rep * (uint32_t)__popcll(peers) + mbcnt(peers);
      if (r != expect || (rep > 0 && r <= prev)) good = 0;
      prev = r;
    }
    __syncthreads();
  }
  if (!good) atomicAnd(ok, 0u);
}

// ---------------------------------------------------------------------------
// Launch `a` of the stable LSD radix sort runs each column's a-th active pass
// (Onesweep-style): launch 0 reads the values (rids implicit) into buffer A,
// then A -> B, B -> A, A -> B.  Tile = THREADS x ITEMS keys of one column;
// wave w ranks keys [w*64*ITEMS, +64*ITEMS) of the tile in (item, lane)
// order, which is the input order, so equal digits keep their relative order
// (stability).  Keys and then rids are staged in one LDS buffer in digit
// order and written out in runs per digit.  Run mode moves (key, rid,
// literal bits) items the run cut wrote to B: launch 0 reads B.
// ---------------------------------------------------------------------------
template <int THREADS, int ITEMS, bool RUN, int RANK>
__global__ __launch_bounds__(THREADS, sort_waves_per_eu(ITEMS)) void k_onesweep(
    uint32_t a, const uint32_t* __restrict__ values, uint32_t* keysA, uint32_t* ridsA,
    uint32_t* keysB, uint32_t* ridsB, uint32_t* litsA, uint32_t* litsB,
    const DescHead* __restrict__ desc, const ColState* __restrict__ cs, Header* __restrict__ hdr,
    uint32_t* status_base, uint64_t status_stride, uint32_t lanes) {
  constexpr int TILE = THREADS * ITEMS;
  constexpr int WAVES = THREADS / kWave;
  static_assert(THREADS >= 256 && THREADS % 256 == 0, "one thread per digit");
  static_assert(ITEMS <= 32, "participation bits");
  static_assert(TILE <= 65536, "LDS slots packed in 16 bits");
  __shared__ uint32_t s_stage[TILE];
  __shared__ uint32_t s_wcnt[WAVES][256];
  __shared__ uint32_t s_excl[256];
  __shared__ uint32_t s_glob[256];
  __shared__ uint32_t s_dsum[4];
  __shared__ uint32_t s_tilen;
  __shared__ uint32_t s_tile;

  // pair mode: launch 0 reads the values (rids implicit); run mode: the run
  // cut wrote the items to B.  Then A -> B, B -> A, ...
  const bool implicit = !RUN && a == 0;
  const uint32_t* src_keys = implicit ? values : ((a % 2) ? keysA : keysB);
  const uint32_t* src_rids = (a % 2) ? ridsA : ridsB;
  const uint32_t* src_lits = (a % 2) ? litsA : litsB;
  uint32_t* dst_keys = (a % 2) ? keysB : keysA;
  uint32_t* dst_rids = (a % 2) ? ridsB : ridsA;
  uint32_t* dst_lits = (a % 2) ? litsB : litsA;
  uint32_t* status = status_base + (uint64_t)(a % 2) * status_stride;
  uint32_t* status_next = status_base + (uint64_t)(1u - a % 2) * status_stride;
  uint32_t* fault = &hdr->fault;
  const Dims dims = desc->dims;
  const Column* cols = desc_cols(desc);
  const int t = threadIdx.x, l = t % kWave;
  const int w = __builtin_amdgcn_readfirstlane(t / kWave);  // scalar: per-wave LDS addresses
  if (hdr->max_active[RUN ? 1 : 0] <= a) return;  // no column has an a-th pass
  uint32_t* counter = hdr->tiles[(RUN ? kCounterRun : 0) + a];

  uint32_t iter = 0;
  for (;;) {
    for (int i = t; i < WAVES * 256; i += THREADS) (&s_wcnt[0][0])[i] = 0;
    const uint32_t tile = __builtin_amdgcn_readfirstlane(
        next_tile(counter, lanes, &s_tile, iter, WAH_SORT_TILE_GROUP));
    if (tile >= dims.sort_tiles) break;
    const uint32_t c = find_segment(&cols[0].sort_tile0, sizeof(Column) / 4, dims.ncols, tile);
    const ColState* col = &cs[c];
    const Column cd = cols[c];
    const uint32_t first = cd.sort_tile0;
    const uint64_t base = (uint64_t)(tile - first) * TILE;  // column-local
    // elements: rows, or the run items of a run-mode column
    const uint64_t n = RUN ? (uint64_t)col->items : cd.n;
    const uint64_t n_out = n;
    if (col->mode != (RUN ? kModeRun : kModePair) || col->wide || col->num_active <= a ||
        base >= n) {
      // other mode, already sorted, or past the run items; a block draws its
      // tiles in increasing order, so past the last column's items it is done
      if (RUN && c + 1 == dims.ncols && col->mode == kModeRun && base >= n) break;
      __syncthreads();  // s_tile is read by every wave before it is rewritten
      continue;
    }
    const uint32_t shift = col->shift[a];
    // 16-bit keys: every pass writes them; the first reads the values (or
    // the run cut's items) as 32-bit words
    const bool narrow = col->narrow != 0u, src_narrow = narrow && !implicit && !(RUN && a == 0);
    if (t < 256) status_next[(uint64_t)tile * 256 + t] = 0;  // next launch's state
    const uint64_t wbase = base + (uint64_t)w * kWave * ITEMS + l;
    // A full tile runs a straight-line body (no per-item guards), so the
    // compiler can wait for the rid loads without draining the key stores.
    // PACKED: launch 0 of a pair-mode column with keys < 2^16.  Its rids are
    // implicit (rid_base + column row), so one LDS word carries both the key
    // (high half) and the row's tile-local index (low half): one staging
    // round moves keys and rids.
    auto sweep = [&](auto full_tag, auto packed_tag) {
      constexpr bool FULL = decltype(full_tag)::value;
      constexpr bool PACKED = decltype(packed_tag)::value;
      // Keys only for now: the rids are loaded once the keys are staged, so
      // they never occupy registers during ranking and look-back.
      uint32_t key[ITEMS], rank[ITEMS];
      const uint32_t nvalid =
          FULL ? (uint32_t)ITEMS * kWave : (uint32_t)(n > wbase ? n - wbase : 0);
      if (src_narrow) {
        const uint16_t* kp = reinterpret_cast<const uint16_t*>(src_keys + cd.off) + wbase;
#pragma unroll
        for (int i = 0; i < ITEMS; ++i)
          key[i] = FULL || (uint32_t)(i * kWave) < nvalid ? (uint32_t)kp[i * kWave] : 0u;
      } else {
        const uint32_t* kp = src_keys + cd.off + wbase;  // one base, immediate offsets
#pragma unroll
        for (int i = 0; i < ITEMS; ++i)
          key[i] = FULL || (uint32_t)(i * kWave) < nvalid ? kp[i * kWave] : 0u;
      }
      uint32_t part = 0;  // items this lane ranks
#pragma unroll
      for (int i = 0; i < ITEMS; ++i)
        part |= (FULL || (uint32_t)(i * kWave) < nvalid ? 1u : 0u) << i;
      // Rank: peers = lanes with the same digit (8 ballots); rank = peers
      // below this lane + keys of that digit the wave ranked before.
#pragma unroll
      for (int i = 0; i < ITEMS; ++i) {
        if (FULL || ((part >> i) & 1u)) {
          const uint32_t d = (key[i] >> shift) & 255u;
          if (RANK == kRankAtomic) {
            rank[i] = atomicAdd(&s_wcnt[w][d], 1u);
          } else {
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
      }
      __syncthreads();
      // Threads 0..255 own one digit each: wave prefixes, tile total,
      // exclusive scan over digits, look-back.
      uint32_t total = 0;
      if (t < 256) {
        const uint32_t goff = col->digit_offset[shift / 8][t];
        // all the wave counts first (independent reads), then the prefixes:
        // interleaved, each read would wait for the write before it
        uint32_t xw[WAVES];
#pragma unroll
        for (int i = 0; i < WAVES; ++i) xw[i] = s_wcnt[i][t];
#pragma unroll
        for (int i = 0; i < WAVES; ++i) {
          s_wcnt[i][t] = total;
          total += xw[i];
        }
        uint32_t* my = &status[(uint64_t)tile * 256 + t];
        st_agent(my, ((tile == first ? kFlagPrefix : kFlagAgg) << 30) | total);
        const uint32_t inc = wave_incl_add(total);
        if (l == kWave - 1) s_dsum[w] = inc;
        s_excl[t] = inc - total;  // within this wave's 64 digits
        uint32_t prev = 0;
        if (tile > first) {
          prev = lookback_digit(status + t, tile, first, 256, fault);
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
        s_glob[t] -= ex;  // column index = s_glob[d] + local index
        if (t == 255) s_tilen = ex + total;
        // each wave's first slot for digit t: one LDS read per key below
        uint32_t xw[WAVES];
#pragma unroll
        for (int i = 0; i < WAVES; ++i) xw[i] = s_wcnt[i][t];
#pragma unroll
        for (int i = 0; i < WAVES; ++i) s_wcnt[i][t] = xw[i] + ex;
      }
      __syncthreads();
      if constexpr (PACKED) {
        const uint32_t row0 = (uint32_t)(wbase - base);  // tile-local index of item 0
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
          if (FULL || ((part >> i) & 1u)) {
            const uint32_t d = (key[i] >> shift) & 255u;
            s_stage[s_wcnt[w][d] + rank[i]] = key[i] << 16 | (row0 + (uint32_t)(i * kWave));
          }
        }
        __syncthreads();
        const uint32_t tile_n = FULL ? (uint32_t)TILE : s_tilen;
        uint16_t* dk16 = reinterpret_cast<uint16_t*>(dst_keys + cd.off);
        uint32_t* dr = dst_rids + cd.off;
        const uint32_t last = (uint32_t)(n_out - 1);  // stores stay in the column
        const uint32_t rid0 = cd.rid_base + (uint32_t)base;
        bool bad = false;
        // the staged words, then their digits' offsets, then the stores: each
        // batch of LDS reads issues back to back (slots past tile_n read
        // in-bounds stale words and store nothing)
        uint32_t xs[ITEMS], gs[ITEMS];
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) xs[j] = s_stage[(uint32_t)j * THREADS + t];
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) gs[j] = s_glob[(xs[j] >> (16 + shift)) & 255u];
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
          const uint32_t q = (uint32_t)j * THREADS + t;
          if (FULL || q < tile_n) {
            const uint32_t g = gs[j] + q;
            bad |= g > last;
            const uint32_t gc = g < last ? g : last;
            dk16[gc] = (uint16_t)(xs[j] >> 16);
            dr[gc] = rid0 + (xs[j] & 0xFFFFu);
          }
        }
        if (ballot(bad) && l == 0) atomicOr(fault, 2u);  // a scatter index past the column
        return;
      }
#pragma unroll
      for (int i = 0; i < ITEMS; ++i) {
        if (FULL || ((part >> i) & 1u)) {
          const uint32_t d = (key[i] >> shift) & 255u;
          const uint32_t pos = s_wcnt[w][d] + rank[i];
          s_stage[pos] = key[i];
          rank[i] = pos;  // keep the LDS slot for the rids
        }
      }
      // LDS slots, two per register (a tile has at most 2^16 items)
      uint32_t slot2[(ITEMS + 1) / 2];
#pragma unroll
      for (int i = 0; i < ITEMS; i += 2)
        slot2[i / 2] = (rank[i] & 0xFFFFu) | (i + 1 < ITEMS ? rank[i + 1 < ITEMS ? i + 1 : i] << 16 : 0u);
      // rids: issued now, in flight while the keys go out
      uint32_t rid[ITEMS];
      if (implicit) {
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) rid[i] = cd.rid_base + (uint32_t)wbase + (uint32_t)(i * kWave);
      } else {
        const uint32_t* rp = src_rids + cd.off + wbase;
#pragma unroll
        for (int i = 0; i < ITEMS; ++i)
          rid[i] = FULL || (uint32_t)(i * kWave) < nvalid ? rp[i * kWave] : 0u;
      }
      __syncthreads();
      const uint32_t tile_n = FULL ? (uint32_t)TILE : s_tilen;
      uint32_t* dk = dst_keys + cd.off;
      uint32_t* dr = dst_rids + cd.off;
      const uint32_t last = (uint32_t)(n_out - 1);  // stores stay in the column
      bool bad = false;
      // destination of each slot this thread stores, kept for the rids (and
      // literal bits): the digit comes from the staged key
      uint32_t gq[ITEMS], kq[ITEMS];
      uint16_t* dk16 = reinterpret_cast<uint16_t*>(dk);
      // the staged keys, then their digits' offsets, then the stores: each
      // batch of LDS reads issues back to back (slots past tile_n read
      // in-bounds stale keys and store nothing)
#pragma unroll
      for (int j = 0; j < ITEMS; ++j) kq[j] = s_stage[(uint32_t)j * THREADS + t];
#pragma unroll
      for (int j = 0; j < ITEMS; ++j) gq[j] = s_glob[(kq[j] >> shift) & 255u];
#pragma unroll
      for (int j = 0; j < ITEMS; ++j) {
        const uint32_t q = (uint32_t)j * THREADS + t;
        const uint32_t g = gq[j] + q;
        const bool in = FULL || q < tile_n;
        bad |= in && g > last;
        gq[j] = in && g < last ? g : last;
      }
      if (narrow) {
#pragma unroll
        for (int j = 0; j < ITEMS; ++j)
          if (FULL || (uint32_t)j * THREADS + t < tile_n) dk16[gq[j]] = (uint16_t)kq[j];
      } else {
#pragma unroll
        for (int j = 0; j < ITEMS; ++j)
          if (FULL || (uint32_t)j * THREADS + t < tile_n) dk[gq[j]] = kq[j];
      }
      __syncthreads();
#pragma unroll
      for (int i = 0; i < ITEMS; ++i)
        if (FULL || ((part >> i) & 1u)) s_stage[(slot2[i / 2] >> (16 * (i % 2))) & 0xFFFFu] = rid[i];
      __syncthreads();
#pragma unroll
      for (int j = 0; j < ITEMS; ++j) kq[j] = s_stage[(uint32_t)j * THREADS + t];
#pragma unroll
      for (int j = 0; j < ITEMS; ++j)
        if (FULL || (uint32_t)j * THREADS + t < tile_n) dr[gq[j]] = kq[j];
      if (RUN) {
        // literal bits of each run item, moved along with its key and rid
        __syncthreads();
        const uint32_t* lp = src_lits + cd.off + wbase;  // one base, immediate offsets
#pragma unroll
        for (int i = 0; i < ITEMS; ++i)
          if (FULL || ((part >> i) & 1u)) s_stage[(slot2[i / 2] >> (16 * (i % 2))) & 0xFFFFu] = lp[i * kWave];
        __syncthreads();
        uint32_t* dl = dst_lits + cd.off;
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) kq[j] = s_stage[(uint32_t)j * THREADS + t];
#pragma unroll
        for (int j = 0; j < ITEMS; ++j)
          if (FULL || (uint32_t)j * THREADS + t < tile_n) dl[gq[j]] = kq[j];
      }
      if (ballot(bad) && l == 0) atomicOr(fault, 2u);  // a scatter index past the column
    };
    auto dispatch = [&](auto packed_tag) {
      if (base + TILE <= n) sweep(std::true_type{}, packed_tag);
      else sweep(std::false_type{}, packed_tag);
    };
    if constexpr (!RUN) {
      if (implicit && narrow) dispatch(std::true_type{});
      else dispatch(std::false_type{});
    } else {
      dispatch(std::false_type{});
    }
    __syncthreads();
  }
}

#ifndef WAH_WIDE_LOOKAHEAD  // predecessors per look-back round trip of the wide passes
#define WAH_WIDE_LOOKAHEAD 8
#endif

// ---------------------------------------------------------------------------
// Wide digits (pair-mode columns that k_plan marked wide).
//
// k_whist: histograms of the column's wide digits over its values (the
// 8-bit ones of k_histogram do not give them), in LDS per block, added to
// the column's WideState; and zero look-back state for the first wide pass.
// Blocks take contiguous ranges of the 4096-value histogram chunks, as in
// k_histogram, and skip the chunks of other columns.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_whist(const uint32_t* __restrict__ values,
                                                  const DescHead* __restrict__ desc,
                                                  const ColState* __restrict__ cs,
                                                  WideState* __restrict__ ws,
                                                  uint32_t* __restrict__ wstatus0) {
  __shared__ uint32_t h[kWidePasses * kWideBins];
  const int t = threadIdx.x;
  const Dims dims = desc->dims;
  const Column* cols = desc_cols(desc);
  for (uint32_t tile = blockIdx.x; tile < dims.wide_tiles; tile += gridDim.x) {
    const uint32_t c = find_segment(&cols[0].wide_tile0, sizeof(Column) / 4, dims.ncols, tile);
    if (!cs[c].wide) continue;  // block-uniform
    uint32_t* st = wstatus0 + (uint64_t)tile * kWideBins;
    for (uint32_t i = t; i < kWideBins; i += kBlock) st[i] = 0;
  }
  for (uint32_t i = t; i < (uint32_t)kWidePasses * kWideBins; i += kBlock) h[i] = 0;
  __syncthreads();
  const uint32_t per = (dims.hist_chunks + gridDim.x - 1) / gridDim.x;
  const uint32_t c0 = blockIdx.x * per;
  const uint32_t c1 = c0 + per < dims.hist_chunks ? c0 + per : dims.hist_chunks;
  uint32_t col = c0 < c1 ? find_segment(&cols[0].hist_chunk0, sizeof(Column) / 4, dims.ncols, c0)
                         : 0u;
  bool counted = false;  // h holds counts of column `col`
  auto flush = [&]() {
    __syncthreads();
    uint32_t* dst = ws[cs[col].wide_slot].hist[0];
    for (uint32_t i = t; i < (uint32_t)kWidePasses * kWideBins; i += kBlock) {
      if (h[i]) atomicAdd(&dst[i], h[i]);
      h[i] = 0;
    }
    __syncthreads();
    counted = false;
  };
  for (uint32_t c = c0; c < c1;) {
    const uint32_t col_end = col + 1 < dims.ncols ? cols[col + 1].hist_chunk0 : dims.hist_chunks;
    if (c >= col_end) {  // next column (block-uniform)
      if (counted) flush();
      col = find_segment(&cols[0].hist_chunk0, sizeof(Column) / 4, dims.ncols, c);
      continue;
    }
    const ColState* st = &cs[col];
    if (!st->wide) {  // skip the rest of this column's chunks
      c = col_end;
      continue;
    }
    counted = true;
    const Column cd = cols[col];
    const uint64_t lo = (uint64_t)(c - cd.hist_chunk0) * kHistChunk;
    const uint32_t cnt = cd.n - lo < kHistChunk ? (uint32_t)(cd.n - lo) : kHistChunk;
    const uint32_t* src = values + cd.off + lo;
    const uint32_t np = st->num_active;
    const uint32_t sh0 = st->wshift[0], sh1 = st->wshift[1], sh2 = st->wshift[2];
    const uint32_t m0 = (1u << st->wbits[0]) - 1u, m1 = (1u << st->wbits[1]) - 1u,
                   m2 = (1u << st->wbits[2]) - 1u;
    auto add = [&](uint32_t v) {
      atomicAdd(&h[(v >> sh0) & m0], 1u);
      if (np > 1) atomicAdd(&h[kWideBins + ((v >> sh1) & m1)], 1u);
      if (np > 2) atomicAdd(&h[2 * kWideBins + ((v >> sh2) & m2)], 1u);
    };
    if (cnt == kHistChunk && ((uintptr_t)src % 16) == 0) {
      // four 16-byte loads per thread, issued before any is counted
      const uint4* v4 = reinterpret_cast<const uint4*>(src);
      uint4 q[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) q[u] = v4[u * kBlock + t];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        add(q[u].x);
        add(q[u].y);
        add(q[u].z);
        add(q[u].w);
      }
    } else {
      for (uint32_t i = t; i < cnt; i += kBlock) add(src[i]);
    }
    ++c;
  }
  if (counted) flush();
}

// k_wplan: each wide column's digit histograms -> exclusive prefixes (the
// first row of each digit's keys in the column), in place.
__global__ __launch_bounds__(kBlock) void k_wplan(const DescHead* __restrict__ desc,
                                                  const ColState* __restrict__ cs,
                                                  WideState* __restrict__ ws) {
  __shared__ uint64_t s_scan[kWaves];
  constexpr uint32_t kPer = kWideBins / kBlock;
  const int t = threadIdx.x;
  const uint32_t ncols = desc->dims.ncols;
  for (uint32_t c = blockIdx.x; c < ncols; c += gridDim.x) {
    const ColState* col = &cs[c];
    if (!col->wide) continue;  // block-uniform
    for (uint32_t p = 0; p < col->num_active; ++p) {
      uint32_t* hp = ws[col->wide_slot].hist[p] + (uint32_t)t * kPer;
      uint32_t x[kPer];
      uint64_t sum = 0;
#pragma unroll
      for (uint32_t k = 0; k < kPer; ++k) {
        x[k] = hp[k];
        sum += x[k];
      }
      uint64_t total;
      uint64_t ex = block_exclusive<OpAdd>(sum, s_scan, &total);
#pragma unroll
      for (uint32_t k = 0; k < kPer; ++k) {
        hp[k] = (uint32_t)ex;
        ex += x[k];
      }
    }
  }
}

// k_wsweep: launch `a` sorts the a-th wide digit of every wide column, one
// tile of kWideTile pairs at a time, as k_onesweep does with 8-bit digits:
// rank in LDS, publish the tile's digit counts, look back for the counts of
// the tiles before it, scatter through LDS.  The 2048 digits take two 16-bit
// counters per LDS word (a wave ranks at most 64 x kWideItems keys and a
// tile holds 8192), and thread t owns digits 4t .. 4t + 3.
template <int RANK>
__global__ __launch_bounds__(kWideThreads, sort_waves_per_eu(kWideItems)) void k_wsweep(
    uint32_t a, const uint32_t* __restrict__ values, uint32_t* keysA, uint32_t* ridsA,
    uint32_t* keysB, uint32_t* ridsB, const DescHead* __restrict__ desc,
    const ColState* __restrict__ cs, const WideState* __restrict__ ws, Header* __restrict__ hdr,
    uint32_t* status_base, uint64_t status_stride, uint32_t lanes) {
  constexpr int THREADS = kWideThreads, ITEMS = kWideItems, TILE = (int)kWideTile;
  constexpr int WAVES = THREADS / kWave;
  constexpr int DPT = (int)kWideBins / THREADS;  // digits per thread
  static_assert(DPT == 4, "two counter words per thread and wave");
  static_assert(TILE <= 65535, "16-bit counters and slots");
  __shared__ uint32_t s_stage[TILE];
  __shared__ uint32_t s_cnt[WAVES][kWideBins / 2];
  __shared__ uint32_t s_glob[kWideBins];
  __shared__ uint32_t s_wsum[WAVES];
  __shared__ uint32_t s_tilen;
  __shared__ uint32_t s_tile;

  const bool implicit = a == 0;
  const uint32_t* src_keys = implicit ? values : ((a % 2) ? keysA : keysB);
  const uint32_t* src_rids = (a % 2) ? ridsA : ridsB;
  uint32_t* dst_keys = (a % 2) ? keysB : keysA;
  uint32_t* dst_rids = (a % 2) ? ridsB : ridsA;
  uint32_t* status = status_base + (uint64_t)(a % 2) * status_stride;
  uint32_t* status_next = status_base + (uint64_t)(1u - a % 2) * status_stride;
  uint32_t* fault = &hdr->fault;
  const Dims dims = desc->dims;
  const Column* cols = desc_cols(desc);
  const int t = threadIdx.x, l = t % kWave;
  const int w = __builtin_amdgcn_readfirstlane(t / kWave);
  if (hdr->max_wide <= a) return;  // no column has an a-th wide pass
  uint32_t* counter = hdr->tiles[kCounterWide + a];

  uint32_t iter = 0;
  for (;;) {
    for (int i = t; i < WAVES * (int)kWideBins / 2; i += THREADS) (&s_cnt[0][0])[i] = 0;
    const uint32_t tile = __builtin_amdgcn_readfirstlane(
        next_tile(counter, lanes, &s_tile, iter, WAH_SORT_TILE_GROUP));
    if (tile >= dims.wide_tiles) break;
    const uint32_t c = find_segment(&cols[0].wide_tile0, sizeof(Column) / 4, dims.ncols, tile);
    const ColState* col = &cs[c];
    const Column cd = cols[c];
    const uint32_t first = cd.wide_tile0;
    const uint64_t base = (uint64_t)(tile - first) * TILE;  // column-local
    const uint64_t n = cd.n;
    if (!col->wide || col->num_active <= a || base >= n) {
      __syncthreads();  // s_tile is read by every wave before it is rewritten
      continue;
    }
    const uint32_t shift = col->wshift[a];
    const uint32_t dmask = (1u << col->wbits[a]) - 1u;
    const uint32_t* goff = ws[col->wide_slot].hist[a];
    const bool narrow = col->narrow != 0u, src_narrow = narrow && !implicit;
#pragma unroll
    for (int k = 0; k < DPT; ++k) status_next[(uint64_t)tile * kWideBins + DPT * t + k] = 0;
    const uint64_t wbase = base + (uint64_t)w * kWave * ITEMS + l;
    auto sweep = [&](auto full_tag) {
      constexpr bool FULL = decltype(full_tag)::value;
      uint32_t key[ITEMS], rank[ITEMS];
      const uint32_t nvalid =
          FULL ? (uint32_t)ITEMS * kWave : (uint32_t)(n > wbase ? n - wbase : 0);
      if (src_narrow) {
        const uint16_t* kp = reinterpret_cast<const uint16_t*>(src_keys + cd.off) + wbase;
#pragma unroll
        for (int i = 0; i < ITEMS; ++i)
          key[i] = FULL || (uint32_t)(i * kWave) < nvalid ? (uint32_t)kp[i * kWave] : 0u;
      } else {
        const uint32_t* kp = src_keys + cd.off + wbase;
#pragma unroll
        for (int i = 0; i < ITEMS; ++i)
          key[i] = FULL || (uint32_t)(i * kWave) < nvalid ? kp[i * kWave] : 0u;
      }
      uint32_t part = 0;
#pragma unroll
      for (int i = 0; i < ITEMS; ++i)
        part |= (FULL || (uint32_t)(i * kWave) < nvalid ? 1u : 0u) << i;
#pragma unroll
      for (int i = 0; i < ITEMS; ++i) {
        if (FULL || ((part >> i) & 1u)) {
          const uint32_t d = (key[i] >> shift) & dmask;
          const uint32_t hs = 16u * (d & 1u);
          if (RANK == kRankAtomic) {
            rank[i] = (atomicAdd(&s_cnt[w][d >> 1], 1u << hs) >> hs) & 0xFFFFu;
          } else {
            uint64_t peers = ballot(true);
#pragma unroll
            for (int b = 0; b < kWideBits; ++b) {
              const bool bit = (d >> b) & 1u;
              const uint64_t bb = ballot(bit);
              peers &= bit ? bb : ~bb;
            }
            const uint32_t cnt = (uint32_t)__popcll(peers);
            const uint32_t below = mbcnt(peers);
            const uint32_t prior = (s_cnt[w][d >> 1] >> hs) & 0xFFFFu;
            rank[i] = prior + below;
            // two digits share a word: add, do not store
            if (below == cnt - 1) atomicAdd(&s_cnt[w][d >> 1], cnt << hs);
          }
        }
      }
      __syncthreads();
      // digits 4t .. 4t + 3: wave prefixes and tile totals
      uint32_t tot[DPT] = {0u, 0u, 0u, 0u};
#pragma unroll
      for (int i = 0; i < WAVES; ++i) {
        const uint32_t w0 = s_cnt[i][2 * t], w1 = s_cnt[i][2 * t + 1];
        s_cnt[i][2 * t] = tot[0] | (tot[1] << 16);
        s_cnt[i][2 * t + 1] = tot[2] | (tot[3] << 16);
        tot[0] += w0 & 0xFFFFu;
        tot[1] += w0 >> 16;
        tot[2] += w1 & 0xFFFFu;
        tot[3] += w1 >> 16;
      }
      uint32_t* my = &status[(uint64_t)tile * kWideBins + DPT * t];
#pragma unroll
      for (int k = 0; k < DPT; ++k)
        st_agent(&my[k], ((tile == first ? kFlagPrefix : kFlagAgg) << 30) | tot[k]);
      const uint32_t sum4 = tot[0] + tot[1] + tot[2] + tot[3];
      const uint32_t inc = wave_incl_add(sum4);
      if (l == kWave - 1) s_wsum[w] = inc;
      __syncthreads();
      uint32_t ex = inc - sum4;  // tile-local first slot of digit 4t
#pragma unroll
      for (int i = 0; i < WAVES; ++i) ex += i < w ? s_wsum[i] : 0u;
      uint32_t exk[DPT];
#pragma unroll
      for (int k = 0; k < DPT; ++k) {
        exk[k] = ex;
        ex += tot[k];
      }
      if (t == THREADS - 1) s_tilen = ex;
#pragma unroll
      for (int i = 0; i < WAVES; ++i) {
        s_cnt[i][2 * t] += exk[0] | (exk[1] << 16);
        s_cnt[i][2 * t + 1] += exk[2] | (exk[3] << 16);
      }
      __syncthreads();
      // Stage the keys in digit order before the look-back, which then
      // holds few registers (the tile-local slots do not depend on it).
#pragma unroll
      for (int i = 0; i < ITEMS; ++i) {
        if (FULL || ((part >> i) & 1u)) {
          const uint32_t d = (key[i] >> shift) & dmask;
          const uint32_t pos = ((s_cnt[w][d >> 1] >> (16u * (d & 1u))) & 0xFFFFu) + rank[i];
          s_stage[pos] = key[i];
          rank[i] = pos;  // keep the LDS slot for the rids
        }
      }
      uint32_t slot2[(ITEMS + 1) / 2];
#pragma unroll
      for (int i = 0; i < ITEMS; i += 2)
        slot2[i / 2] = (rank[i] & 0xFFFFu) | (i + 1 < ITEMS ? rank[i + 1 < ITEMS ? i + 1 : i] << 16 : 0u);
      uint32_t prev[DPT] = {0u, 0u, 0u, 0u};
      if (tile > first) {
        lookback_digits4<WAH_WIDE_LOOKAHEAD>(status + DPT * t, tile, first, kWideBins, fault, prev);
#pragma unroll
        for (int k = 0; k < DPT; ++k) st_agent(&my[k], (kFlagPrefix << 30) | (prev[k] + tot[k]));
      }
#pragma unroll
      for (int k = 0; k < DPT; ++k)
        s_glob[DPT * t + k] = goff[DPT * t + k] + prev[k] - exk[k];  // + slot = column index
      uint32_t rid[ITEMS];
      if (implicit) {
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) rid[i] = cd.rid_base + (uint32_t)wbase + (uint32_t)(i * kWave);
      } else {
        const uint32_t* rp = src_rids + cd.off + wbase;
#pragma unroll
        for (int i = 0; i < ITEMS; ++i)
          rid[i] = FULL || (uint32_t)(i * kWave) < nvalid ? rp[i * kWave] : 0u;
      }
      __syncthreads();
      const uint32_t tile_n = FULL ? (uint32_t)TILE : s_tilen;
      uint32_t* dk = dst_keys + cd.off;
      uint32_t* dr = dst_rids + cd.off;
      uint16_t* dk16 = reinterpret_cast<uint16_t*>(dk);
      const uint32_t last = (uint32_t)(n - 1);  // stores stay in the column
      bool bad = false;
      uint32_t gq[ITEMS];
#pragma unroll
      for (int j = 0; j < ITEMS; ++j) {
        const uint32_t q = (uint32_t)j * THREADS + t;
        gq[j] = last;
        if (FULL || q < tile_n) {
          const uint32_t k = s_stage[q];
          const uint32_t g = s_glob[(k >> shift) & dmask] + q;
          bad |= g > last;
          gq[j] = g < last ? g : last;
          if (narrow) dk16[gq[j]] = (uint16_t)k;
          else dk[gq[j]] = k;
        }
      }
      __syncthreads();
#pragma unroll
      for (int i = 0; i < ITEMS; ++i)
        if (FULL || ((part >> i) & 1u)) s_stage[(slot2[i / 2] >> (16 * (i % 2))) & 0xFFFFu] = rid[i];
      __syncthreads();
#pragma unroll
      for (int j = 0; j < ITEMS; ++j) {
        const uint32_t q = (uint32_t)j * THREADS + t;
        if (FULL || q < tile_n) dr[gq[j]] = s_stage[q];
      }
      if (ballot(bad) && l == 0) atomicOr(fault, 2u);  // a scatter index past the column
    };
    if (base + TILE <= n) sweep(std::true_type{});
    else sweep(std::false_type{});
    __syncthreads();
  }
}

// Selects the sorted arrays of one column (after its active passes); the
// pointers address the column's first pair.
struct SortedView {
  const uint32_t* keys;  // 32-bit keys, or 16-bit ones at the same start (narrow)
  const uint32_t* rids;
  const uint32_t* lits;  // run mode
  bool implicit;
  bool narrow;
  __device__ __forceinline__ uint32_t key(uint32_t i) const {
    return narrow ? (uint32_t)reinterpret_cast<const uint16_t*>(keys)[i] : keys[i];
  }
};
__device__ __forceinline__ SortedView sorted_view(const ColState* col, const Column& cd,
                                                  const uint32_t* values,
                                                  const uint32_t* keysA, const uint32_t* ridsA,
                                                  const uint32_t* keysB, const uint32_t* ridsB,
                                                  const uint32_t* litsA = nullptr,
                                                  const uint32_t* litsB = nullptr) {
  SortedView v;
  const uint32_t f = col->final_buf;
  v.implicit = f == kBufValues;
  v.narrow = col->narrow != 0u && f != kBufValues;
  v.keys = (f == kBufValues ? values : (f == kBufA ? keysA : keysB)) + cd.off;
  v.rids = (f == kBufA ? ridsA : ridsB) + cd.off;
  v.lits = (f == kBufA ? litsA : litsB) + cd.off;
  return v;
}

// ---------------------------------------------------------------------------
// Fused tail over the sorted items: (value, rid) pairs, or in run mode
// (value, rid, literal bits) runs.  Per item: chid = rid/31; an entry starts
// where (value, chid) changes and ORs the literal bits of the items up to the
// next entry start (at most 31 rows, one chunk); a key starts where value
// changes.  fill = chid for a key start, else chid - chid_prev - 1; an entry
// emits [fill if nonzero, literal].
//
// Blocked layout: thread t owns the ITEMS consecutive items [t*ITEMS, +ITEMS)
// of a tile, loaded from HBM straight into registers with 16-byte loads.  An
// item's predecessor is the lane's previous register, lane l-1's last item
// (DPP wave shift) or, for lane 0, one load per wave.  Word and key counts
// are summed along the lane, then one DPP wave scan and the wave totals give
// each lane its tile-local offsets; a wave-parallel look-back across tiles
// gives the tile's.  The run of a lane's last entry start continues into at
// most four following lanes (31 items, ITEMS >= 8 per lane): their lead bits
// (before their first entry start) arrive by DPP lane shifts; a run crossing
// into the next wave or tile takes that wave's or tile's lead bits, which
// are exchanged in LDS.  Words are staged in LDS at tile-local positions and
// stored contiguously.
// ---------------------------------------------------------------------------
template <int THREADS, int ITEMS, bool RUN>
__global__ __launch_bounds__(THREADS, tail_waves_per_eu(ITEMS)) void k_tail(
    const DescHead* __restrict__ desc, const ColState* __restrict__ cs, Header* __restrict__ hdr,
    const uint32_t* values, const uint32_t* keysA, const uint32_t* ridsA,
    const uint32_t* keysB, const uint32_t* ridsB, const uint32_t* litsA, const uint32_t* litsB,
    uint32_t* __restrict__ keys_out, uint32_t* __restrict__ offsets_out,
    uint32_t* __restrict__ words_out, uint64_t* status, uint32_t* __restrict__ tile_entries,
    uint32_t lanes, wahgpu_counts_dev* __restrict__ counts) {
  constexpr int TILE = THREADS * ITEMS;
  constexpr int WAVES = THREADS / kWave;
  static_assert(ITEMS % 8 == 0 && ITEMS <= 16, "16-byte loads (8 narrow keys); runs span <= 4 lanes");
  static_assert(2 * TILE < 65536, "tile word and key counts fit 16 bits");
  constexpr uint32_t kAll = (1u << ITEMS) - 1u;
  // staged words, one pad word per 16: lanes that write ~16 words apart
  // (two per item) land in different banks
  __shared__ uint32_t s_words[2 * TILE + 2 * TILE / 16];
  __shared__ uint32_t s_wtot[WAVES];
  __shared__ uint32_t s_went[WAVES];
  __shared__ uint32_t s_wlead[WAVES + 1];  // [w]: lead bits of wave w; [WAVES]: of the next tile
  __shared__ uint64_t s_prefix;
  __shared__ uint32_t s_tile;
  const Dims dims = desc->dims;
  const Column* cols = desc_cols(desc);
  // the wave index in a scalar register: per-wave LDS addresses stay out of
  // the 64 vector registers
  const int t = threadIdx.x, l = t % kWave;
  const int w = __builtin_amdgcn_readfirstlane(t / kWave);
  uint32_t* fault = &hdr->fault;
  if (hdr->mode_cols[RUN ? 1 : 0] == 0) return;  // no column in this mode

  uint32_t iter = 0;
  for (;;) {
    const uint32_t tile = __builtin_amdgcn_readfirstlane(next_tile(
        &hdr->tiles[RUN ? kCounterRunTail : kCounterTail][0], lanes, &s_tile, iter,
        WAH_TAIL_TILE_GROUP));
    if (tile >= dims.tail_tiles) break;
    const uint32_t c = find_segment(&cols[0].tail_tile0, sizeof(Column) / 4, dims.ncols, tile);
    const Column cd = cols[c];
    const uint32_t first_tile = cd.tail_tile0;
    // column-local indices fit 32 bits: a column holds fewer than 2^30 items
    const uint32_t base = (tile - first_tile) * (uint32_t)TILE;
    // items: rows, or the runs of a run-mode column (fewer than its tiles)
    const uint32_t n = RUN ? (uint32_t)cs[c].items : (uint32_t)cd.n;
    if (cs[c].mode != (RUN ? kModeRun : kModePair) || base >= n) {
      if (RUN && c + 1 == dims.ncols && cs[c].mode == kModeRun) break;  // past the last items
      __syncthreads();  // s_tile is read by every wave before it is rewritten
      continue;
    }
    // the tile's body, compiled once per key width (16-bit keys load 8 per
    // 16 bytes); the column's width picks one
    // and once for whole aligned tiles (16-byte loads, no per-item checks,
    // and no control-flow merge with the per-item loads of partial tiles)
    const SortedView sv0 =
        sorted_view(&cs[c], cd, values, keysA, ridsA, keysB, ridsB, litsA, litsB);
    const bool aligned = (((uintptr_t)sv0.keys | (uintptr_t)sv0.rids |
                           (RUN ? (uintptr_t)sv0.lits : (uintptr_t)0)) % 16) == 0;
    auto body = [&](auto narrow_tag, auto full_tag) {
    constexpr bool NARROW = decltype(narrow_tag)::value;
    constexpr bool FULL = decltype(full_tag)::value;
    const uint32_t rid_base = cd.rid_base;
    SortedView sv = sv0;
    sv.narrow = NARROW;
    const uint32_t j0 = base + (uint32_t)t * ITEMS;  // the lane's first item
    const uint32_t nv =
        FULL ? (uint32_t)ITEMS
             : (n > j0 ? (n - j0 < (uint32_t)ITEMS ? n - j0 : (uint32_t)ITEMS) : 0u);

    // Loads the tile needs besides its own items, issued before them so that
    // one memory latency covers all: the item before the wave (lane 0's
    // predecessor; wave-uniform) and, for the last wave, the 32 items after
    // the tile (the run that crosses its end).
    // Branch-free: out-of-range indices read item 0 and are masked later
    // (an implicit-rid column's rids array is the unused ping-pong buffer).
    const uint32_t gw = base + (uint32_t)(w * kWave * ITEMS);
    const bool have_pred = gw > 0 && gw <= n;
    const uint32_t gp = have_pred ? gw - 1 : 0u;
    uint32_t k0 = sv.key(gp), r0 = sv.rids[gp];
    const uint32_t gn0 = base + (uint32_t)TILE + (uint32_t)l;
    const bool nv_next = w == WAVES - 1 && l < 32 && gn0 < n;
    const uint32_t gn = nv_next ? gn0 : 0u;
    uint32_t kk = 0, rr = 0, ll = 0;
    if (w == WAVES - 1) {
      kk = sv.key(gn);
      rr = sv.rids[gn];
      if (RUN) ll = sv.lits[gn];
    }

    // ---- items of the lane: key, chid, literal bits (0 past the column)
    uint32_t key[ITEMS], ch[ITEMS], b[ITEMS];
    if constexpr (FULL) {
      if (NARROW) {  // 8 keys per 16-byte load
#pragma unroll
        for (int q = 0; q < ITEMS / 8; ++q) {
          const uint4 k8 = reinterpret_cast<const uint4*>(reinterpret_cast<const uint16_t*>(sv.keys) + j0)[q];
          const uint32_t w4[4] = {k8.x, k8.y, k8.z, k8.w};
#pragma unroll
          for (int h = 0; h < 4; ++h) {
            key[8 * q + 2 * h] = w4[h] & 0xFFFFu;
            key[8 * q + 2 * h + 1] = w4[h] >> 16;
          }
        }
      }
#pragma unroll
      for (int q = 0; q < ITEMS / 4; ++q) {
        if (!NARROW) {
          const uint4 k4 = reinterpret_cast<const uint4*>(sv.keys + j0)[q];
          key[4 * q] = k4.x; key[4 * q + 1] = k4.y; key[4 * q + 2] = k4.z; key[4 * q + 3] = k4.w;
        }
        // loaded either way (an implicit-rid column reads the unused
        // buffer), so the loads are not split by a branch
        const uint4 r4 = reinterpret_cast<const uint4*>(sv.rids + j0)[q];
        ch[4 * q] = r4.x; ch[4 * q + 1] = r4.y; ch[4 * q + 2] = r4.z; ch[4 * q + 3] = r4.w;
        if (RUN) {
          const uint4 l4 = reinterpret_cast<const uint4*>(sv.lits + j0)[q];
          b[4 * q] = l4.x; b[4 * q + 1] = l4.y; b[4 * q + 2] = l4.z; b[4 * q + 3] = l4.w;
        }
      }
    } else {
#pragma unroll
      for (int i = 0; i < ITEMS; ++i) {
        const bool v = (uint32_t)i < nv;
        const uint32_t g = j0 + i;
        key[i] = v ? sv.key(g) : 0u;
        ch[i] = v ? (sv.implicit ? rid_base + g : sv.rids[g]) : 0u;
        if (RUN) b[i] = v ? sv.lits[g] : 0u;
      }
    }
    if (FULL && sv.implicit) {
#pragma unroll
      for (int i = 0; i < ITEMS; ++i) ch[i] = rid_base + j0 + (uint32_t)i;
    }
    // ch[] held rids so far: chunk id q = rid / 31 and the literal bit
    // rid % 31 = (rid + q) & 31 (rid + q = 32q + rid % 31).  Rows below 2^31
    // (every column with rid_base + n <= 2^31) divide by a multiply-high,
    // exact there (it first fails at 2369637147); the others by the
    // compiler's full 32-bit division.
    auto chunks = [&](auto fast_tag) {
      constexpr bool FAST = decltype(fast_tag)::value;
#pragma unroll
      for (int i = 0; i < ITEMS; ++i) {
        const uint32_t r = ch[i];
        const uint32_t q = FAST ? __umulhi(r, 0x84210843u) >> 4 : r / 31u;
        ch[i] = q;
        if (!RUN) b[i] = (uint32_t)i < nv ? 1u << ((r + q) & 31u) : 0u;
      }
    };
    if ((uint64_t)rid_base + cd.n <= (1ull << 31)) chunks(std::true_type{});
    else chunks(std::false_type{});
    // predecessor of the lane's first item: lane l-1's last, or for lane 0
    // the item before the wave (one load per wave)
    uint32_t pk, pc;
    {
      // first use of the side loads, tied to the tile's own (so that no
      // arithmetic on them is placed, with its wait, before those loads)
      k0 = __builtin_amdgcn_readfirstlane(k0);
      r0 = __builtin_amdgcn_readfirstlane(r0);
      asm volatile("" : "+s"(k0), "+s"(r0) : "v"(key[ITEMS - 1]), "v"(ch[ITEMS - 1]));
      if (!have_pred) k0 = r0 = 0u;
      else if (sv.implicit) r0 = rid_base + gp;
      const uint32_t c0 = r0 / 31u;
      pk = (uint32_t)__builtin_amdgcn_update_dpp((int)k0, (int)key[ITEMS - 1], 0x138, 0xF, 0xF, false);
      pc = (uint32_t)__builtin_amdgcn_update_dpp((int)c0, (int)ch[ITEMS - 1], 0x138, 0xF, 0xF, false);
    }
    // ---- entry and key starts, fills, counts
    const bool colfirst = base == 0 && t == 0;
    uint32_t f[ITEMS];
    uint32_t stm = 0, khm = 0, fm = 0;  // per item: entry start, key start, nonzero fill
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      const uint32_t kp = i ? key[i - 1] : pk, cp = i ? ch[i - 1] : pc;
      const bool v = (uint32_t)i < nv;
      const bool kh = v && (key[i] != kp || (i == 0 && colfirst));
      const bool st = kh || (v && ch[i] != cp);
      f[i] = st ? (kh ? ch[i] : ch[i] - cp - 1u) : 0u;
      stm |= st ? 1u << i : 0u;
      khm |= kh ? 1u << i : 0u;
      fm |= f[i] != 0u ? 1u << i : 0u;
    }
    // From here on an item's flags are read back from these bits: the
    // compiler would otherwise keep each item's compare masks (64-bit scalar
    // pairs) alive to the staging and spill them to VGPR lanes.
    stm = opaque(stm);
    khm = opaque(khm);
    fm = opaque(fm);
    const uint32_t nwords = (uint32_t)(__builtin_popcount(stm) + __builtin_popcount(fm));
    // lead bits: items before the lane's first stop (entry start or end of items)
    const uint32_t stops = stm | (~0u << nv & kAll);
    const uint32_t fs = stops ? (uint32_t)__builtin_ctz(stops) : (uint32_t)ITEMS;
    uint32_t lead = 0;
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) lead |= (uint32_t)i < fs ? b[i] : 0u;
    // continuation of the lane's last run inside the wave: leads of lanes
    // l+1 .. the next lane with a stop (at most 4 lanes on)
    const uint64_t hsm = ballot(stops != 0u);
    const uint64_t above = l == kWave - 1 ? 0ull : hsm >> (l + 1);
    const uint32_t dist = above ? (uint32_t)__builtin_ctzll(above) + 1u : 64u;
    uint32_t cont = 0;
    {
      uint32_t sh = lead;
#pragma unroll
      for (uint32_t s = 1; s <= 4; ++s) {
        sh = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)sh, 0x130, 0xF, 0xF, false);  // wave_shl:1
        cont |= s <= dist ? sh : 0u;
      }
    }
    // wave lead: lane 0's items before its first stop and, without one, the
    // lanes after it
    {
      const uint32_t wl = readlane(stops ? lead : lead | cont, 0);
      if (l == 0) s_wlead[w] = wl;
    }
    // next tile's lead: the run that crosses the tile end (last wave, 32 items)
    if (w == WAVES - 1) {
      const bool v = nv_next;
      if (sv.implicit) rr = rid_base + gn;
      uint32_t cc = 0, bb = 0;
      if (!v) kk = 0u;
      if (v) {
        cc = rr / 31u;
        bb = RUN ? ll : 1u << (rr - cc * 31u);
      }
      const uint32_t kl = readlane(key[ITEMS - 1], kWave - 1), cl = readlane(ch[ITEMS - 1], kWave - 1);
      const uint32_t kq = (uint32_t)__builtin_amdgcn_update_dpp((int)kl, (int)kk, 0x138, 0xF, 0xF, false);
      const uint32_t cq = (uint32_t)__builtin_amdgcn_update_dpp((int)cl, (int)cc, 0x138, 0xF, 0xF, false);
      const uint64_t sm = ballot(!v || kk != kq || cc != cq);  // lanes 32.. always stop
      const int fh = __builtin_ctzll(sm);
      const uint32_t tl = wave_or(l < fh ? bb : 0u);
      if (l == 0) s_wlead[WAVES] = tl;
    }
    // tile-local word and key offsets: one wave scan of (keys << 16 | words)
    const uint32_t packed = ((uint32_t)__builtin_popcount(khm) << 16) | nwords;
    const uint32_t inc = wave_incl_add(packed);
    if (l == kWave - 1) s_wtot[w] = inc;
    {
      // entries of the tile (k of the column is their sum, formed by k_finish)
      const uint32_t e = readlane(wave_incl_add((uint32_t)__builtin_popcount(stm)), kWave - 1);
      if (l == 0) s_went[w] = e;
    }
    __syncthreads();

    // ---- literals: reverse sweep along the lane, seeded with the bits the
    // lane's last run picks up from the lanes (wave, tile) after it
    if (above == 0) cont |= s_wlead[w + 1];
    if (nv < (uint32_t)ITEMS) cont = 0;  // the run ends with the items
    uint32_t woff, tot;
    {
      // wave totals: one LDS read per lane and a wave scan
      const uint32_t lo = opaque((uint32_t)l);
      const uint32_t x = lo < WAVES ? s_wtot[lo] : 0u;
      const uint32_t xi = wave_incl_add(x);
      woff = readlane(xi - x, __builtin_amdgcn_readfirstlane(w));
      tot = readlane(xi, WAVES - 1);
    }
    const uint32_t lpos = woff + inc - packed;  // (keys << 16 | words) before the lane
    if (w == 0) {
      const uint64_t tile_agg = ((uint64_t)(tot >> 16) << 31) | (tot & 0xFFFFu);
      const uint64_t pre = lookback_wave_sum<WAH_TAIL_LOOKBACK_SLOTS>(status, tile, first_tile,
                                                                      tile_agg, fault);
      const uint32_t lo = opaque((uint32_t)l);
      const uint32_t e = readlane(wave_incl_add(lo < WAVES ? s_went[lo] : 0u), kWave - 1);
      if (l == 0) {
        s_prefix = pre;
        tile_entries[tile] = e;
      }
    }
    {
      // from the lane's last word back: [fill,] literal of each entry start
      uint32_t carry = cont;
      uint32_t pos = (lpos & 0xFFFFu) + nwords;
#pragma unroll
      for (int i = ITEMS - 1; i >= 0; --i) {
        const bool st = (stm >> i) & 1u, fw = (fm >> i) & 1u;
        carry |= b[i];
        pos -= st ? 1u : 0u;
        if (st) s_words[pos + (pos >> 4)] = 0x80000000u | carry;
        pos -= fw ? 1u : 0u;
        if (fw) s_words[pos + (pos >> 4)] = f[i];
        carry = st ? 0u : carry;
      }
    }
    uint32_t kv[ITEMS];
    if (khm) {
      // key starts: the values (the lane's keys read again, all at once:
      // the lines are still in L2) and the position of each key's first word
      if (nv == (uint32_t)ITEMS && aligned && !NARROW) {
#pragma unroll
        for (int q = 0; q < ITEMS / 4; ++q) {
          const uint4 k4 = reinterpret_cast<const uint4*>(sv.keys + j0)[q];
          kv[4 * q] = k4.x; kv[4 * q + 1] = k4.y; kv[4 * q + 2] = k4.z; kv[4 * q + 3] = k4.w;
        }
      } else if (nv == (uint32_t)ITEMS && aligned) {
#pragma unroll
        for (int q = 0; q < ITEMS / 8; ++q) {
          const uint4 k8 = reinterpret_cast<const uint4*>(reinterpret_cast<const uint16_t*>(sv.keys) + j0)[q];
          const uint32_t w4[4] = {k8.x, k8.y, k8.z, k8.w};
#pragma unroll
          for (int h = 0; h < 4; ++h) {
            kv[8 * q + 2 * h] = w4[h] & 0xFFFFu;
            kv[8 * q + 2 * h + 1] = w4[h] >> 16;
          }
        }
      } else {
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) kv[i] = (uint32_t)i < nv ? sv.key(j0 + i) : 0u;
      }
    }
    __syncthreads();  // words staged, prefix published
    const uint64_t pre = s_prefix;
    const uint32_t Wpre = (uint32_t)(pre & 0x7FFFFFFFu), Kpre = (uint32_t)(pre >> 31);
    uint32_t* out_keys = keys_out + cd.off;
    uint32_t* out_offsets = offsets_out + cd.off;
    uint32_t* out_words = words_out + 2 * cd.off;
    if (khm) {
      bool bad = false;
#pragma unroll
      for (int i = 0; i < ITEMS; ++i) {
        if ((khm >> i) & 1u) {
          const uint32_t below = (1u << i) - 1u;
          const uint32_t kg = Kpre + (lpos >> 16) + __builtin_popcount(khm & below);
          const uint32_t wp = (lpos & 0xFFFFu) + __builtin_popcount(stm & below) +
                              __builtin_popcount(fm & below);
          bad |= kg >= n;
          const uint32_t kc = kg < n ? kg : n - 1u;  // stores stay in the column
          out_keys[kc] = kv[i];
          out_offsets[kc] = Wpre + wp;
        }
      }
      if (bad) atomicOr(fault, 2u);
    }
    const uint32_t Wt = tot & 0xFFFFu;
    if ((uint64_t)Wpre + Wt <= 2ull * n) {
      // single words up to the first 16-byte boundary, then 4 words per
      // 16-byte store, then the rest
      uint32_t* ow = out_words + Wpre;
      const uint32_t head0 = (4u - (uint32_t)(((uintptr_t)ow >> 2) & 3u)) & 3u;
      const uint32_t head = head0 < Wt ? head0 : Wt;
      const uint32_t quads = (Wt - head) / 4u;
      const uint32_t tt = opaque((uint32_t)t);
      if (tt < head) ow[tt] = s_words[tt + (tt >> 4)];
      for (uint32_t qd = tt; qd < quads; qd += THREADS) {
        const uint32_t j = head + 4u * qd;
        uint4 v;
        v.x = s_words[j + (j >> 4)];
        v.y = s_words[j + 1u + ((j + 1u) >> 4)];
        v.z = s_words[j + 2u + ((j + 2u) >> 4)];
        v.w = s_words[j + 3u + ((j + 3u) >> 4)];
        *reinterpret_cast<uint4*>(ow + j) = v;  // 16-byte aligned: j - head is a multiple of 4
      }
      for (uint32_t j = head + 4u * quads + tt; j < Wt; j += THREADS) ow[j] = s_words[j + (j >> 4)];
    } else {
      for (uint32_t j = opaque((uint32_t)t); j < Wt; j += THREADS) {
        const uint32_t o = Wpre + j;
        if (o < 2 * n) out_words[o] = s_words[j + (j >> 4)];
        else atomicOr(fault, 2u);
      }
    }
    if (base + TILE >= n && t == 0) {  // the column's last tile
      counts[cd.index].length = (uint64_t)Wpre + Wt;
      counts[cd.index].keycnt = (uint64_t)Kpre + (tot >> 16);
    }
    };
    const bool full = base + TILE <= n && aligned;
    if (cs[c].narrow) {
      if (full) body(std::true_type{}, std::true_type{});
      else body(std::true_type{}, std::false_type{});
    } else {
      if (full) body(std::false_type{}, std::true_type{});
      else body(std::false_type{}, std::false_type{});
    }
    // the next tile's first barrier (in next_tile) orders these LDS reads
    // before its writes
  }
}

// Per column: k = the sum of its tail tiles' entry counts; lengths[i] =
// offsets[i+1] - offsets[i], the last key running to L; the build's fault
// word is copied into every column's counts.
__global__ __launch_bounds__(kBlock) void k_finish(const DescHead* __restrict__ desc,
                                                   const Header* __restrict__ hdr,
                                                   const uint32_t* __restrict__ tile_entries,
                                                   const uint32_t* __restrict__ offsets,
                                                   uint32_t* __restrict__ lengths,
                                                   wahgpu_counts_dev* __restrict__ counts) {
  const uint32_t ncols = desc->dims.ncols;
  const Column* cols = desc_cols(desc);
  const uint32_t fault = hdr->fault;
  const uint64_t gt = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  const uint64_t gs = (uint64_t)gridDim.x * kBlock;
  {
    // k per column = the entry counts of its tail tiles (zeroed by the
    // histogram, written by the tiles that held items).  Each wave takes 64
    // consecutive tiles, sums them per column (segmented wave scan) and adds
    // each column's part once.
    const uint32_t lane = threadIdx.x % kWave;
    const uint32_t ntiles = desc->dims.tail_tiles;
    for (uint64_t w0 = gt - lane; w0 < ntiles; w0 += gs) {
      const uint32_t tile = (uint32_t)w0 + lane;
      uint32_t c = find_segment(&cols[0].tail_tile0, sizeof(Column) / 4, ncols, (uint32_t)w0);
      uint32_t v = 0;
      if (tile < ntiles) {
        while (c + 1 < ncols && cols[c + 1].tail_tile0 <= tile) ++c;
        v = tile_entries[tile];
      }
      const uint32_t cp = (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)c, 0x138, 0xF, 0xF, false);
      const uint64_t hm = ballot(c != cp);  // first lane of each column
      const uint32_t inc = wave_incl_add(v);
      const uint64_t upto = lane == kWave - 1 ? ~0ull : ((2ull << lane) - 1ull);
      const int h = 63 - __builtin_clzll(hm & upto);  // this lane's first lane of its column
      const uint32_t seg = inc - (uint32_t)__shfl((int)(inc - v), h, kWave);
      const bool last = lane == kWave - 1 || ((hm >> (lane + 1)) & 1ull);
      if (last && seg) atomicAdd((unsigned long long*)&counts[cols[c].index].k, (unsigned long long)seg);
    }
  }
  for (uint32_t c = 0; c < ncols; ++c) {
    const Column cd = cols[c];
    const uint64_t kc = counts[cd.index].keycnt;
    const uint64_t K = kc < cd.n ? kc : cd.n, L = counts[cd.index].length;
    const uint32_t* o = offsets + cd.off;
    uint32_t* len = lengths + cd.off;
    for (uint64_t i = gt; i < K; i += gs) {
      const uint32_t end = (i + 1 < K) ? o[i + 1] : (uint32_t)L;
      len[i] = end - o[i];
    }
    if (gt == 0 && fault) counts[cd.index].fault = fault;
  }
}

// ---------------------------------------------------------------------------
// Staged kernels (one per reference stage) and primitives.
// ---------------------------------------------------------------------------
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

// Copy the sorted result of column 0 out of the pass buffers (stage API).
__global__ __launch_bounds__(kBlock) void k_sort_finalize(
    const DescHead* __restrict__ desc, const ColState* __restrict__ cs, const uint32_t* values,
    const uint32_t* keysA, const uint32_t* ridsA, const uint32_t* keysB, const uint32_t* ridsB,
    uint32_t* out_keys, uint32_t* out_rids) {
  const Column cd = desc_cols(desc)[0];
  const SortedView sv = sorted_view(&cs[0], cd, values, keysA, ridsA, keysB, ridsB);
  if (sv.keys == out_keys && !sv.implicit) return;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < cd.n;
       i += (uint64_t)gridDim.x * kBlock) {
    out_keys[i] = sv.key((uint32_t)i);
    out_rids[i] = sv.implicit ? cd.rid_base + (uint32_t)i : sv.rids[i];
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
static uint32_t cap_grid(uint32_t work, uint32_t cap) {
  const uint32_t g = work < cap ? work : cap;
  return g ? g : 1u;
}

// Radix blocks resident per CU: the waves the register budget allows
// (sort_waves_per_eu per SIMD) and the LDS of k_onesweep (staged keys,
// per-wave digit counts, digit tables; 160 KiB per CU).
template <int T, int I>
constexpr uint32_t sort_blocks_per_cu() {
  constexpr uint32_t by_waves = sort_waves_per_eu(I) * 4 / (T / kWave);
  constexpr uint32_t lds = T * I * 4 + (T / kWave) * 1024 + 2 * 1024 + 64;
  constexpr uint32_t by_lds = 160u * 1024u / lds;
  constexpr uint32_t b = by_waves < by_lds ? by_waves : by_lds;
  return b ? b : 1u;
}

// Wide-digit blocks per CU: two 512-thread blocks (4 waves per SIMD) with
// ~72 KiB of LDS each (staged keys, 16-bit counters, digit offsets).
constexpr uint32_t kWideBlocksPerCU = 2;

// The run cut and the four radix launches for one tile shape (with the
// wide-digit launches when a column may take them).
template <int T, int I>
static void launch_radix(const BuildArgs& a, hipStream_t s, LaunchHook* hook) {
  const uint32_t grid = cap_grid(a.sort_tiles_cap, WAH_STATIC_TILES
                                                       ? a.sort_tiles_cap
                                                       : a.max_resident / 4 * sort_blocks_per_cu<T, I>());
  hipLaunchKernelGGL(k_runcut, dim3(cap_grid(a.sort_tiles_cap, 2 * a.max_resident)), dim3(kBlock),
                     0, s, a.values, a.keysB, a.ridsB, a.litsB, a.desc, a.cs, a.hdr, a.tile_runs,
                     (uint32_t)(T * I));
  // wide digits: the histograms of the wide columns' digits and their offsets
  const uint32_t wgrid = cap_grid(a.wide_tiles_cap, WAH_STATIC_TILES ? a.wide_tiles_cap
                                                                    : a.max_resident / 4 * kWideBlocksPerCU);
  if (a.wide_tiles_cap) {
    uint64_t hb = (a.n_cap + 4 * kHistChunk - 1) / (4 * kHistChunk);
    if (hb > a.hist_chunks_cap) hb = a.hist_chunks_cap;
    hipLaunchKernelGGL(k_whist, dim3(cap_grid((uint32_t)hb, kHistMaxBlocks)), dim3(kBlock), 0, s,
                       a.values, a.desc, a.cs, a.wide, a.wide_status);
    hipLaunchKernelGGL(k_wplan, dim3(cap_grid(a.cols_cap, 1024)), dim3(kBlock), 0, s, a.desc,
                       a.cs, a.wide);
  }
  if (hook) hook->mark(s);
  // each launch runs the a-th pass of pair-mode columns, then of run-mode
  // columns; a kernel with no such column returns at once
  for (uint32_t p = 0; p < 4; ++p) {
    if (a.atomic_rank) {
      hipLaunchKernelGGL((k_onesweep<T, I, false, kRankAtomic>), dim3(grid), dim3(T), 0, s, p,
                         a.values, a.keysA, a.ridsA, a.keysB, a.ridsB, a.litsA, a.litsB, a.desc,
                         a.cs, a.hdr, a.sort_status, a.sort_status_stride, tile_lanes(grid));
      hipLaunchKernelGGL((k_onesweep<T, I, true, kRankAtomic>), dim3(grid), dim3(T), 0, s, p,
                         a.values, a.keysA, a.ridsA, a.keysB, a.ridsB, a.litsA, a.litsB, a.desc,
                         a.cs, a.hdr, a.sort_status, a.sort_status_stride, tile_lanes(grid));
    } else {
      hipLaunchKernelGGL((k_onesweep<T, I, false, kRankMatch>), dim3(grid), dim3(T), 0, s, p,
                         a.values, a.keysA, a.ridsA, a.keysB, a.ridsB, a.litsA, a.litsB, a.desc,
                         a.cs, a.hdr, a.sort_status, a.sort_status_stride, tile_lanes(grid));
      hipLaunchKernelGGL((k_onesweep<T, I, true, kRankMatch>), dim3(grid), dim3(T), 0, s, p,
                         a.values, a.keysA, a.ridsA, a.keysB, a.ridsB, a.litsA, a.litsB, a.desc,
                         a.cs, a.hdr, a.sort_status, a.sort_status_stride, tile_lanes(grid));
    }
    if (a.wide_tiles_cap && p < (uint32_t)kWidePasses) {
      if (a.atomic_rank)
        hipLaunchKernelGGL((k_wsweep<kRankAtomic>), dim3(wgrid), dim3(kWideThreads), 0, s, p,
                           a.values, a.keysA, a.ridsA, a.keysB, a.ridsB, a.desc, a.cs, a.wide,
                           a.hdr, a.wide_status, a.wide_status_stride, tile_lanes(wgrid));
      else
        hipLaunchKernelGGL((k_wsweep<kRankMatch>), dim3(wgrid), dim3(kWideThreads), 0, s, p,
                           a.values, a.keysA, a.ridsA, a.keysB, a.ridsB, a.desc, a.cs, a.wide,
                           a.hdr, a.wide_status, a.wide_status_stride, tile_lanes(wgrid));
    }
    if (hook) hook->mark(s);
  }
}

void launch_sort(const BuildArgs& a, hipStream_t s, LaunchHook* hook) {
  hipLaunchKernelGGL(k_begin, dim3(64), dim3(kBlock), 0, s, a.hdr, a.cs, a.cols_cap,
                     a.tile_runs, a.sort_tiles_cap, a.desc, (int)a.inline_desc, a.one_dims, a.one);
  uint64_t hb = (a.n_cap + 4 * kHistChunk - 1) / (4 * kHistChunk);
  if (hb > a.hist_chunks_cap) hb = a.hist_chunks_cap;
  hipLaunchKernelGGL(k_histogram, dim3(cap_grid((uint32_t)hb, kHistMaxBlocks)), dim3(kBlock), 0,
                     s, a.values, a.desc, a.cs, a.sort_status, a.tail_status, a.tile_entries,
                     a.tile_runs,
                     radix_tile(a.shape));
  hipLaunchKernelGGL(k_plan, dim3(cap_grid(a.cols_cap, 1024)), dim3(kBlock), 0, s, a.desc, a.cs,
                     a.hdr, a.tile_runs, a.counts, (int)a.force_pair, radix_tile(a.shape), a.wide);
  switch (a.shape) {
#define WAH_SHAPE_CASE(sh) \
  case sh: launch_radix<radix_threads(sh), radix_items(sh)>(a, s, hook); break;
    WAH_SHAPE_CASE(1)
    WAH_SHAPE_CASE(2)
    WAH_SHAPE_CASE(3)
    WAH_SHAPE_CASE(4)
    WAH_SHAPE_CASE(5)
#undef WAH_SHAPE_CASE
    default: launch_radix<radix_threads(0), radix_items(0)>(a, s, hook); break;
  }
  static_assert(kRadixShapes == 6, "a case per compiled shape");
}

// Both tails of a long column, on 16384-item tiles.
static void launch_long_tail(const BuildArgs& a, hipStream_t s) {
  const uint32_t grid =
      cap_grid(a.tail_tiles_cap, WAH_STATIC_TILES ? a.tail_tiles_cap
                                                  : a.max_resident / 4 * kLongTailBlocksPerCU);
  const uint32_t lanes = WAH_TAIL_LANES ? WAH_TAIL_LANES : tile_lanes(grid);
  hipLaunchKernelGGL((k_tail<kLongTailThreads, kLongTailItems, false>), dim3(grid),
                     dim3(kLongTailThreads), 0, s, a.desc, a.cs, a.hdr, a.values, a.keysA,
                     a.ridsA, a.keysB, a.ridsB, a.litsA, a.litsB, a.out_keys, a.out_offsets,
                     a.out_words, a.tail_status, a.tile_entries, lanes, a.counts);
  hipLaunchKernelGGL((k_tail<kLongTailThreads, kLongTailItems, true>), dim3(grid),
                     dim3(kLongTailThreads), 0, s, a.desc, a.cs, a.hdr, a.values, a.keysA,
                     a.ridsA, a.keysB, a.ridsB, a.litsA, a.litsB, a.out_keys, a.out_offsets,
                     a.out_words, a.tail_status, a.tile_entries, lanes, a.counts);
}

// Both tails of other columns and of batches, on 8192-item tiles.
static void launch_short_tail(const BuildArgs& a, hipStream_t s) {
  const uint32_t grid =
      cap_grid(a.tail_tiles_cap, WAH_STATIC_TILES ? a.tail_tiles_cap
                                                  : a.max_resident / 4 * kTailBlocksPerCU);
  const uint32_t lanes = WAH_TAIL_LANES ? WAH_TAIL_LANES : tile_lanes(grid);
  hipLaunchKernelGGL((k_tail<kTailThreads, kTailItems, false>), dim3(grid), dim3(kTailThreads),
                     0, s, a.desc, a.cs, a.hdr, a.values, a.keysA, a.ridsA, a.keysB, a.ridsB,
                     a.litsA, a.litsB, a.out_keys, a.out_offsets, a.out_words, a.tail_status,
                     a.tile_entries, lanes, a.counts);
  // run mode: the same tiles as 512 threads x 16 items (fewer, longer lanes
  // suit its sparser items)
  const uint32_t rgrid =
      cap_grid(a.tail_tiles_cap, WAH_STATIC_TILES ? a.tail_tiles_cap
                                                  : a.max_resident / 4 * kRunTailBlocksPerCU);
  const uint32_t rlanes = WAH_TAIL_LANES ? WAH_TAIL_LANES : tile_lanes(rgrid);
  hipLaunchKernelGGL((k_tail<kRunTailThreads, kRunTailItems, true>), dim3(rgrid),
                     dim3(kRunTailThreads), 0, s, a.desc, a.cs, a.hdr, a.values, a.keysA, a.ridsA,
                     a.keysB, a.ridsB, a.litsA, a.litsB, a.out_keys, a.out_offsets, a.out_words,
                     a.tail_status, a.tile_entries, rlanes, a.counts);
}

void launch_tail(const BuildArgs& a, hipStream_t s, LaunchHook* hook) {
  if (a.long_tail)
    launch_long_tail(a, s);
  else
    launch_short_tail(a, s);
  if (hook) hook->mark(s);
  uint64_t lg = (a.n_cap + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(k_finish, dim3(cap_grid(lg > 1024 ? 1024u : (uint32_t)lg, 1024)),
                     dim3(kBlock), 0, s, a.desc, a.hdr, a.tile_entries, a.out_offsets,
                     a.out_lengths, a.counts);
}

void launch_sort_finalize(const BuildArgs& a, hipStream_t s) {
  uint64_t g = (a.n_cap + kBlock - 1) / kBlock;
  if (g > 4096) g = 4096;
  if (g == 0) return;
  hipLaunchKernelGGL(k_sort_finalize, dim3((uint32_t)g), dim3(kBlock), 0, s, a.desc, a.cs,
                     a.values, a.keysA, a.ridsA, a.keysB, a.ridsB, a.sorted_keys, a.sorted_rids);
}

void launch_rank_selftest(uint32_t* ok, hipStream_t s) {
  hipLaunchKernelGGL(k_rank_selftest, dim3(64), dim3(256), 0, s, ok);
}

uint64_t scan_status_words(uint64_t n) { return (n + kScanTile - 1) / kScanTile; }

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
