// phases -- the reference's multi-kernel device pipeline program
// (src/phases.cpp) on libwahgpu.
//
//   phases -f FILE [-b BOUND] [-d DEVICE] [-p] [--binary] [--fused]
//
// Reads the value file (text as written by `generate`; --binary for raw
// uint32), builds the index on the GPU through the staged kernels, one group
// per reference stage, and prints the reference's nine timing lines
// (src/phases.cpp:725,812,859,907,953,984,1082,1092,1126):
//   Transfer to / Sort / Chunks + literals / Merge literals / Produce fills /
//   Fuse / Column length / Transfer back / Total, each as "<microseconds> us".
// --fused runs the fused build instead; its tail kernel covers chunks through
// column length, so those buckets read 0 and "Column length" holds the tail.
// -p prints the index in the `cpu` program's format (declared but unused in
// the reference).  -d picks a device by name or gfx arch and falls back to
// the first device (src/phases.cpp:376-383).
#include <stdio.h>

#include <string>
#include <vector>

#include "cli_common.hpp"

int main(int argc, char** argv) {
  std::string file, bound_s = "0", device = "";
  bool print = false, binary = false, fused = false;
  std::vector<cli::Opt> opts = {
      {"-f", "--data-file", false, &file, nullptr, "file with test data (one value per line)"},
      {"-b", "--bound", false, &bound_s, nullptr, "maximum value (0 will scan values)"},
      {"-d", "--device", false, &device, nullptr,
       "device for computation (first available device if not found)"},
      {"-p", "--print", true, nullptr, &print, "print resulting bitmap index"},
      {nullptr, "--binary", true, nullptr, &binary, "value file is raw little-endian uint32"},
      {nullptr, "--fused", true, nullptr, &fused, "run the fused build (one tail kernel)"},
  };
  int pr = cli::parse(argc, argv, opts);
  if (pr == 1) return 0;
  if (pr == 2) return 2;
  if (file.empty()) {
    printf("A filename with test data is required, see --help for detailts.\n");
    return 0;
  }
  std::vector<uint32_t> values;
  if (!cli::load_values(file, binary, &values)) {
    fprintf(stderr, "phases: cannot read '%s'\n", file.c_str());
    return 1;
  }
  uint32_t bound = cli::to_u32(bound_s);
  if (bound == 0)
    for (uint32_t v : values) bound = v > bound ? v : bound;
  int dev = cli::pick_device(device, true);
  if (dev < 0) {
    fprintf(stderr, "No device found.\n");
    return 1;
  }
  wahgpu_ctx* ctx = nullptr;
  if (wahgpu_create(&ctx, dev, 0) != WAHGPU_OK) {
    fprintf(stderr, "phases: cannot open device %d\n", dev);
    return 1;
  }
  uint64_t n = values.size(), cap = n ? n : 1;
  std::vector<uint32_t> keys(cap), offsets(cap), lengths(cap), words(2 * cap);
  wahgpu_counts counts{};
  wahgpu_phase_times t{};
  int rc = fused ? wahgpu_build(ctx, values.data(), n, 0, keys.data(), offsets.data(),
                                lengths.data(), cap, words.data(), 2 * cap, &counts, &t)
                 : wahgpu_build_staged(ctx, values.data(), n, keys.data(), offsets.data(),
                                       lengths.data(), cap, words.data(), 2 * cap, &counts, &t);
  if (rc != WAHGPU_OK) {
    fprintf(stderr, "phases: build failed (%d): %s\n", rc, wahgpu_last_error(ctx));
    wahgpu_destroy(ctx);
    return 1;
  }
  const double buckets[9] = {t.transfer_to, t.sort, t.chunks, t.merge, t.fills,
                             t.fuse, t.column_length, t.transfer_back, t.total};
  for (double us : buckets) printf("%lld us\n", (long long)us);
  if (print) cli::print_index(keys, offsets, lengths, words, counts.keycnt);
  (void)bound;
  wahgpu_destroy(ctx);
  return 0;
}
