// wah_device.h -- device-side building blocks for the gfx950 WAH kernels:
// wave64 intrinsics, block scans over 256 threads, and the decoupled
// look-back that chains tiles of a single-pass scan across the whole grid.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace wah {

constexpr int kWave = 64;
constexpr int kBlock = 256;               // every kernel here runs 4 waves
constexpr int kWaves = kBlock / kWave;

// Look-back status flags (top two bits of a status word).
constexpr uint32_t kFlagAgg = 1u;         // tile aggregate published
constexpr uint32_t kFlagPrefix = 2u;      // inclusive prefix published

__device__ __forceinline__ uint32_t lane_id() {
  return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}
// The same value, opaque to the optimiser: an address derived from it is
// recomputed where it is used (one or two VALU) instead of being hoisted out
// of a loop and kept, or spilled, across it.  Emits no instruction.
__device__ __forceinline__ uint32_t opaque(uint32_t x) {
  asm volatile("" : "+v"(x));
  return x;
}
__device__ __forceinline__ uint64_t ballot(bool p) {
  return __builtin_amdgcn_ballot_w64(p);
}
// Number of set bits of m below this lane.
__device__ __forceinline__ uint32_t mbcnt(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Agent-scope relaxed atomics: status words are read by blocks on other XCDs,
// whose L2 is not this block's L2.  The word itself carries flag and value,
// so no ordering with other memory is needed.
__device__ __forceinline__ uint32_t ld_agent(uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t ld_agent(uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint64_t shfl_up64(uint64_t x, int d) {
  uint32_t lo = __shfl_up((uint32_t)x, d, kWave);
  uint32_t hi = __shfl_up((uint32_t)(x >> 32), d, kWave);
  return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint64_t shfl64(uint64_t x, int src) {
  uint32_t lo = __shfl((uint32_t)x, src, kWave);
  uint32_t hi = __shfl((uint32_t)(x >> 32), src, kWave);
  return ((uint64_t)hi << 32) | lo;
}

// Tile scheduling for the single-pass kernels.  Tiles are handed out by
// `lanes` counters (one per 256-byte line, kLaneStride words apart): block b
// draws from counter b % lanes and gets tiles lane, lane + lanes, ... in
// increasing order.  A global atomic on one address costs ~11 ns on MI355X
// and they serialize, so spreading the draws over lanes removes that floor.
// A lane's draws ascend, so the lowest unfinished tile is either held by a
// running block or is the next draw of its lane; grids are sized so that all
// their blocks fit on the device at once, so a block of that lane draws it
// and look-back waits end.  That needs every lane to get its blocks placed:
// a lane whose XCD is held by another launch that is waiting itself would
// stall the rest, so the host never overlaps two launch sequences of these
// kernels on one device.  WAH_STATIC_TILES=1 maps tile = blockIdx.x instead
// (one tile per block), relying on in-order dispatch.
#ifndef WAH_STATIC_TILES
#define WAH_STATIC_TILES 0
#endif
constexpr uint32_t kTileLanes = 8;
constexpr uint32_t kLaneStride = 64;
//
// With group > 1 a lane's j-th draw is tile (j / group) * group * lanes +
// lane * group + j % group: runs of `group` consecutive tiles go to one lane.
// Workgroups are dispatched to the 8 XCDs round robin, so lane = block % 8
// is the XCD, and neighbouring tiles then write the cache lines they share
// (the ends of their output runs) through one L2, which merges them.  A
// lane's draws still ascend, so the argument above holds.
__device__ __forceinline__ uint32_t next_tile(uint32_t* counters, uint32_t lanes,
                                              uint32_t* s_tile, uint32_t& iter,
                                              uint32_t group = 1) {
#if WAH_STATIC_TILES
  (void)counters;
  (void)lanes;
  (void)s_tile;
  (void)group;
  __syncthreads();
  return iter++ == 0 ? blockIdx.x : 0xFFFFFFFFu;
#else
  (void)iter;
  if (threadIdx.x == 0) {
    const uint32_t lane = blockIdx.x % lanes;
    const uint32_t j = atomicAdd(&counters[lane * kLaneStride], 1u);
    *s_tile = group == 1 ? j * lanes + lane
                         : (j / group) * group * lanes + lane * group + j % group;
  }
  __syncthreads();
  return *s_tile;
#endif
}
// Counter lanes for a grid of `grid` blocks.
__host__ __device__ __forceinline__ uint32_t tile_lanes(uint32_t grid) {
  return grid < kTileLanes ? (grid ? grid : 1u) : kTileLanes;
}

// Asynchronous copy of 16 bytes per lane from global memory straight into
// LDS (gfx950 global_load_lds_dwordx4): lane l's bytes land at
// lds_wave_base + 16 * l.  The base must be the same for the whole wave.
// Completion is tracked by vmcnt; see wait_vm().
__device__ __forceinline__ void load_lds16(const void* g, uint32_t* lds_wave_base) {
  __builtin_amdgcn_global_load_lds(g, (__attribute__((address_space(3))) void*)lds_wave_base,
                                   16, 0, 0);
}
// s_waitcnt vmcnt(N) with lgkmcnt/expcnt left alone (gfx9 encoding).
template <int N>
__device__ __forceinline__ void wait_vm() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
}

// Inclusive uint32 prefix sum across the wave with DPP row shifts and row
// broadcasts (rows of 16 lanes, then rows 1/3 and 2/3 pick up the totals).
__device__ __forceinline__ uint32_t wave_incl_add(uint32_t x) {
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xF, 0xF, true);   // row_shr:1
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xF, 0xF, true);   // row_shr:2
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xF, 0xF, true);   // row_shr:4
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xF, 0xF, true);   // row_shr:8
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xA, 0xF, false);  // row_bcast:15
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xC, 0xF, false);  // row_bcast:31
  return x;
}
// OR over the wave (every lane gets it).
__device__ __forceinline__ uint32_t wave_or(uint32_t x) {
  x |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xF, 0xF, true);   // row_shr:1
  x |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xF, 0xF, true);   // row_shr:2
  x |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xF, 0xF, true);   // row_shr:4
  x |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xF, 0xF, true);   // row_shr:8
  x |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xA, 0xF, false);  // row_bcast:15
  x |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xC, 0xF, false);  // row_bcast:31
  return (uint32_t)__builtin_amdgcn_readlane((int)x, kWave - 1);
}
__device__ __forceinline__ uint32_t readlane(uint32_t x, int lane) {
  return (uint32_t)__builtin_amdgcn_readlane((int)x, lane);
}
__device__ __forceinline__ uint64_t wave_sum64(uint64_t v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    uint32_t lo = __shfl_xor((uint32_t)v, d, kWave);
    uint32_t hi = __shfl_xor((uint32_t)(v >> 32), d, kWave);
    v += ((uint64_t)hi << 32) | lo;
  }
  return v;
}

// Index of the last entry of a nondecreasing table `start` (count entries,
// `stride` words apart) with start <= x: the segment that holds x.  Empty
// segments share their start with the next one, which this skips.  Uniform
// arguments keep the search in scalar loads.
__device__ __forceinline__ uint32_t find_segment(const uint32_t* start, uint32_t stride,
                                                 uint32_t count, uint32_t x) {
  uint32_t lo = 0, hi = count;  // answer in [lo, hi)
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (start[(uint64_t)mid * stride] <= x) lo = mid;
    else hi = mid;
  }
  return lo;
}

// ---- associative operators on 62-bit look-back payloads -------------------
// combine(a, b): a covers the earlier elements, b the later ones.

// Two independent 31-bit counters packed as (hi << 31 | lo); per-operand
// values stay below 2^31 because a column holds fewer than 2^30 values.
struct OpAdd {
  static __device__ __forceinline__ uint64_t identity() { return 0; }
  static __device__ __forceinline__ uint64_t combine(uint64_t a, uint64_t b) { return a + b; }
};
// uint32 sum that wraps mod 2^32 (the reference's uint32 scans).
struct OpAdd32 {
  static __device__ __forceinline__ uint64_t identity() { return 0; }
  static __device__ __forceinline__ uint64_t combine(uint64_t a, uint64_t b) {
    return (uint32_t)(a + b);
  }
};
// Segmented OR: payload = heads (30 bits) << 32 | or-since-last-head (32).
struct OpSegOr {
  static __device__ __forceinline__ uint64_t identity() { return 0; }
  static __device__ __forceinline__ uint64_t combine(uint64_t a, uint64_t b) {
    uint64_t h = (a >> 32) + (b >> 32);
    uint32_t x = (b >> 32) ? (uint32_t)b : ((uint32_t)a | (uint32_t)b);
    return (h << 32) | x;
  }
};
// Segmented sum: payload = has-head (bit 32) | sum-since-last-head (32).
struct OpSegAdd {
  static __device__ __forceinline__ uint64_t identity() { return 0; }
  static __device__ __forceinline__ uint64_t combine(uint64_t a, uint64_t b) {
    bool hb = (b >> 32) & 1u;
    uint32_t x = hb ? (uint32_t)b : (uint32_t)a + (uint32_t)b;
    return ((uint64_t)(((a >> 32) | (b >> 32)) & 1u) << 32) | x;
  }
};

// Inclusive scan within a wave.
template <class Op>
__device__ __forceinline__ uint64_t wave_inclusive(uint64_t x) {
  const uint32_t l = lane_id();
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    uint64_t y = shfl_up64(x, d);
    if (l >= (uint32_t)d) x = Op::combine(y, x);
  }
  return x;
}

// Exclusive scan over the 256 threads of a block; *total receives the block
// aggregate.  smem holds kWaves entries.  Contains __syncthreads().
template <class Op>
__device__ __forceinline__ uint64_t block_exclusive(uint64_t x, uint64_t* smem,
                                                    uint64_t* total) {
  const int t = threadIdx.x, w = t / kWave, l = t % kWave;
  uint64_t inc = wave_inclusive<Op>(x);
  uint64_t exc = shfl_up64(inc, 1);
  if (l == 0) exc = Op::identity();
  if (l == kWave - 1) smem[w] = inc;
  __syncthreads();
  uint64_t before = Op::identity();
  uint64_t all = Op::identity();
#pragma unroll
  for (int i = 0; i < kWaves; ++i) {
    if (i < w) before = Op::combine(before, smem[i]);
    all = Op::combine(all, smem[i]);
  }
  __syncthreads();
  *total = all;
  return Op::combine(before, exc);
}

// Bounded wait on a look-back status word: returns the word once its flag is
// set, or 0 after kSpinLimit polls (then *fault is raised and the caller
// proceeds with a zero contribution, so every wave still finishes).
// Once any wait of the launch has given up, the others stop waiting too.
constexpr uint32_t kSpinLimit = 1u << 18;

template <class T, int SHIFT>
__device__ __forceinline__ T wait_status_t(T* p, uint32_t* fault) {
  T s;
  for (uint32_t i = 0; i < kSpinLimit; ++i) {
    if ((s = ld_agent(p)) >> SHIFT) return s;
    if ((i & 1023u) == 1023u && ld_agent(fault)) return 0;
    __builtin_amdgcn_s_sleep(1);
  }
  atomicOr(fault, 1u);
  return 0;
}
__device__ __forceinline__ uint32_t wait_status(uint32_t* p, uint32_t* fault) {
  return wait_status_t<uint32_t, 30>(p, fault);
}
__device__ __forceinline__ uint64_t wait_status(uint64_t* p, uint32_t* fault) {
  return wait_status_t<uint64_t, 62>(p, fault);
}

// Exclusive prefix of one digit over the preceding tiles, reading up to
// WAH_LOOKAHEAD predecessors per round trip (32-bit words: flag << 30 | count).
#ifndef WAH_LOOKAHEAD
#define WAH_LOOKAHEAD 4
#endif
// Tiles below `first` (another column's) count as an empty inclusive prefix.
__device__ __forceinline__ uint32_t lookback_digit(uint32_t* status, uint32_t tile,
                                                   uint32_t first, uint32_t stride,
                                                   uint32_t* fault) {
  constexpr int LA = WAH_LOOKAHEAD;
  uint32_t acc = 0;
  const int64_t lo = first;
  int64_t at = (int64_t)tile - 1;  // newest tile not yet summed
  uint32_t spins = 0;
  while (at >= lo) {
    uint32_t s[LA];
#pragma unroll
    for (int k = 0; k < LA; ++k)
      s[k] = at - k >= lo ? ld_agent(&status[(uint64_t)(at - k) * stride]) : (kFlagPrefix << 30);
    int k = 0;
    bool done = false;
#pragma unroll
    for (int j = 0; j < LA; ++j) {
      if (done || k != j) continue;
      const uint32_t f = s[j] >> 30;
      if (f == 0) continue;              // not published yet: poll again from here
      acc += s[j] & 0x3FFFFFFFu;
      ++k;
      if (f == kFlagPrefix) done = true;
    }
    if (done) break;
    at -= k;
    if (k == 0) {
      if (++spins >= kSpinLimit) {
        atomicOr(fault, 1u);
        break;
      }
      if ((spins & 1023u) == 0 && ld_agent(fault)) break;
      __builtin_amdgcn_s_sleep(1);
    }
  }
  return acc;
}

// Exclusive prefixes of four consecutive digits in one walk (the wide-digit
// radix passes, where a thread owns four digits): each step reads LA
// predecessors' four status words as two 64-bit loads each, advances over
// the predecessors whose digits still walking all carry a flag, and a digit
// stops at its nearest inclusive prefix.  Same result as four lookback_digit
// calls, one chain of round trips instead of four.
template <int LA>
__device__ __forceinline__ void lookback_digits4(uint32_t* status, uint32_t tile, uint32_t first,
                                                 uint32_t stride, uint32_t* fault,
                                                 uint32_t acc[4]) {
  constexpr uint64_t kPrefixPair = ((uint64_t)kFlagPrefix << 62) | ((uint64_t)kFlagPrefix << 30);
  uint32_t open = 0xFu;  // digits still walking
  acc[0] = acc[1] = acc[2] = acc[3] = 0;
  const int64_t lo = first;
  int64_t at = (int64_t)tile - 1;  // newest tile not yet summed
  uint32_t spins = 0;
  while (at >= lo && open) {
    uint64_t s[LA][2];
#pragma unroll
    for (int j = 0; j < LA; ++j)
#pragma unroll
      for (int h = 0; h < 2; ++h)
        s[j][h] = at - j >= lo
                      ? ld_agent(reinterpret_cast<uint64_t*>(&status[(uint64_t)(at - j) * stride + 2 * h]))
                      : kPrefixPair;
    int adv = 0;
    bool stop = false;
#pragma unroll
    for (int j = 0; j < LA; ++j) {
      if (stop) continue;
      const uint32_t w[4] = {(uint32_t)s[j][0], (uint32_t)(s[j][0] >> 32), (uint32_t)s[j][1],
                             (uint32_t)(s[j][1] >> 32)};
      bool ready = true;
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (((open >> k) & 1u) && (w[k] >> 30) == 0u) ready = false;
      if (!ready) {  // not published yet: poll again from here
        stop = true;
        continue;
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        if ((open >> k) & 1u) {
          acc[k] += w[k] & 0x3FFFFFFFu;
          if ((w[k] >> 30) == kFlagPrefix) open &= ~(1u << k);
        }
      }
      ++adv;
      if (!open) stop = true;
    }
    at -= adv;
    if (adv == 0) {
      if (++spins >= kSpinLimit) {
        atomicOr(fault, 1u);
        break;
      }
      if ((spins & 1023u) == 0 && ld_agent(fault)) break;
      __builtin_amdgcn_s_sleep(1);
    }
  }
}

// Decoupled look-back over 64-bit status words (flag << 62 | payload).
// Called by one thread of the tile; publishes the tile's aggregate first so
// later tiles can proceed, then walks back to the nearest inclusive prefix,
// publishes its own prefix and returns the exclusive prefix of the tile.
template <class Op>
__device__ uint64_t lookback64(uint64_t* status, uint32_t tile, uint64_t agg, uint32_t* fault) {
  constexpr uint64_t kMask = (1ull << 62) - 1;
  if (tile == 0) {
    st_agent(&status[0], ((uint64_t)kFlagPrefix << 62) | (agg & kMask));
    return Op::identity();
  }
  st_agent(&status[tile], ((uint64_t)kFlagAgg << 62) | (agg & kMask));
  uint64_t acc = Op::identity();
  int64_t t = (int64_t)tile - 1;
  while (t >= 0) {
    uint64_t s = wait_status(&status[t], fault);
    acc = Op::combine(s & kMask, acc);
    if ((s >> 62) != kFlagAgg) break;  // prefix found (or gave up)
    --t;
  }
  st_agent(&status[tile], ((uint64_t)kFlagPrefix << 62) | (Op::combine(acc, agg) & kMask));
  return acc;
}

// Look-back for commutative sums, run by a whole wave: each round reads the
// statuses of the K x 64 preceding tiles at once (K per lane, nearest first)
// and stops at the nearest inclusive prefix; slots complete up to a tile
// without a status are kept and the round resumes from there.  Returns the
// exclusive prefix in every lane.  Tile `first` opens a segment
// (a column): it publishes its prefix at once and nothing below it is read.
template <int K>
__device__ uint64_t lookback_wave_sum(uint64_t* status, uint32_t tile, uint32_t first,
                                      uint64_t agg, uint32_t* fault) {
  constexpr uint64_t kMask = (1ull << 62) - 1;
  const uint32_t lane = lane_id();
  if (lane == 0)
    st_agent(&status[tile], ((uint64_t)(tile == first ? kFlagPrefix : kFlagAgg) << 62) | (agg & kMask));
  if (tile == first) return 0;
  uint64_t acc = 0;
  int64_t base = (int64_t)tile - 1;  // nearest tile not summed yet
  const int64_t lo = first;
  uint32_t spins = 0;
  for (;;) {
    uint64_t s[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int64_t idx = base - (int64_t)(k * kWave) - (int64_t)lane;
      s[k] = idx >= lo ? ld_agent(&status[idx]) : ((uint64_t)kFlagPrefix << 62);
    }
    uint64_t take = 0;
    int done = 0;  // slots summed in full
    bool found = false, stall = false;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      if (found || stall) continue;  // wave-uniform
      const uint32_t f = (uint32_t)(s[k] >> 62);
      const uint64_t pmask = ballot(f == kFlagPrefix);
      const uint64_t rmask = ballot(f != 0);
      const uint64_t need = pmask ? (((pmask & (~pmask + 1)) << 1) - 1) : ~0ull;
      if ((rmask & need) != need) {
        stall = true;
        continue;
      }
      take += ((need >> lane) & 1u) ? (s[k] & kMask) : 0;
      if (pmask) found = true;
      else ++done;
    }
    acc += wave_sum64(take);
    if (found) break;
    base -= (int64_t)done * kWave;
    if (!stall) continue;
    if (++spins >= kSpinLimit) {
      if (lane == 0) atomicOr(fault, 1u);
      break;
    }
    if ((spins & 1023u) == 0 && ld_agent(fault)) break;
    __builtin_amdgcn_s_sleep(1);
  }
  if (lane == 0) st_agent(&status[tile], ((uint64_t)kFlagPrefix << 62) | ((acc + agg) & kMask));
  return acc;
}

}  // namespace wah
