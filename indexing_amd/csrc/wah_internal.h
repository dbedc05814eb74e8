// wah_internal.h -- library-internal entry points shared by the C ABI
// (wahgpu_api.hip) and the build server (wah_server.cpp); not part of the
// public header.
#pragma once

#include <stdint.h>

#include "wahgpu.h"

#ifdef __cplusplus
extern "C" {
#endif

// Running totals of a prospective batch, for an upper bound of the scratch
// its batched build needs (admission by device footprint).
typedef struct wah_batch_acc {
  uint64_t values;       // sum of the columns' lengths rounded up to 64
  uint32_t cols;
  uint64_t sort_tiles;   // counted with the smallest radix tile
  uint64_t tail_tiles;
  uint64_t hist_chunks;
} wah_batch_acc;
void wah_batch_acc_add(wah_batch_acc* acc, uint64_t n);
// Bytes of scratch that bound the batched build of the accumulated columns.
uint64_t wah_batch_acc_scratch(const wah_batch_acc* acc);

// wahgpu_build_batch_device with its scratch in the caller's device buffer
// (no arena, no graph): asynchronous on the context's stream.
int wah_build_batch_ws(wahgpu_ctx* ctx, const uint32_t* d_values, uint64_t values_len,
                       const wahgpu_column* cols, uint32_t ncols, uint32_t* d_keys,
                       uint32_t* d_offsets, uint32_t* d_lengths, uint32_t* d_words,
                       wahgpu_counts* d_counts, void* d_scratch, uint64_t scratch_bytes);

#ifdef __cplusplus
}
#endif
