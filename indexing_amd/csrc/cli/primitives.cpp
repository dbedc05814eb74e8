// primitives.cpp -- the reference's primitive test programs on libwahgpu:
// built three times, as `scan` (src/scan.cpp: exclusive sum scan of 0/1
// flags), `sc` (src/sc.cpp: stream compaction of random values under random
// heads) and `segmented_scan` (src/segmented_scan.cpp: segmented inclusive
// scan of 1, 2, 3, ...).  Each generates its input when no file is given
// (size = U(1,100) * 2^20 + U(1,100), as the reference), runs the HIP
// primitive through the C ABI, checks it against a host loop and prints the
// reference's verdict lines.  Extensions: --seed (the reference is
// unseeded), --size, --time (prints the device time in microseconds).
#include <hip/hip_runtime_api.h>
#include <stdint.h>
#include <stdio.h>

#include <chrono>
#include <random>
#include <string>
#include <vector>

#include "cli_common.hpp"

#if !defined(PRIM_SCAN) && !defined(PRIM_SC) && !defined(PRIM_SEGSCAN)
#error "build with -DPRIM_SCAN, -DPRIM_SC or -DPRIM_SEGSCAN"
#endif

namespace {

struct Rng {  // splitmix64
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  uint64_t below(uint64_t m) { return next() % m; }           // [0, m)
  bool bern(double p) { return (next() >> 11) * 0x1.0p-53 < p; }
};

#define HIPX(call)                                                              \
  do {                                                                          \
    hipError_t e_ = (call);                                                     \
    if (e_ != hipSuccess) {                                                     \
      printf("[!!!] %s: %s\n", #call, hipGetErrorString(e_));                   \
      return 1;                                                                 \
    }                                                                           \
  } while (0)

}  // namespace

int main(int argc, char** argv) {
  std::string file, device, seed_s, size_s, freq_s = "0.1", bound_s, threshold_s, iter_s;
  bool print = false, timed = false;
  std::vector<cli::Opt> opts = {
      {"-f", "--data-file", false, &file, nullptr, "file with test data (one value per line)"},
      {"-b", "--bound", false, &bound_s, nullptr, "maximum value (accepted, unused)"},
      {"-d", "--device", false, &device, nullptr, "device name or gfx arch (default: first)"},
      {"-p", "--print", true, nullptr, &print, "print the results"},
      {"-t", "--threshold", false, &threshold_s, nullptr, "accepted, unused"},
      {"-i", "--iterations", false, &iter_s, nullptr, "accepted, unused"},
#ifdef PRIM_SC
      {"-F", "--frequency", false, &freq_s, nullptr, "frequency of 1 in the heads array (0.1)"},
#endif
      {nullptr, "--seed", false, &seed_s, nullptr, "seed of the generated input"},
      {nullptr, "--size", false, &size_s, nullptr, "number of generated values"},
      {nullptr, "--time", true, nullptr, &timed, "print the device time in us"},
  };
  int prc = cli::parse(argc, argv, opts);
  if (prc == 1) return 0;
  if (prc) return prc;
  Rng rng{seed_s.empty() ? (uint64_t)std::random_device{}() << 32 ^ std::random_device{}()
                         : strtoull(seed_s.c_str(), nullptr, 10)};
  [[maybe_unused]] const double freq = atof(freq_s.c_str());
  std::vector<uint32_t> values;
  if (file.empty()) {
    uint64_t size = size_s.empty() ? (1 + rng.below(100)) * 1048576ull + 1 + rng.below(100)
                                   : strtoull(size_s.c_str(), nullptr, 10);
#if defined(PRIM_SCAN)
    printf("Using %llu values.\n", (unsigned long long)size);
    values.resize(size);
    for (uint64_t i = 0; i < size; ++i) values[i] = i == 0 ? 1u : (rng.bern(0.1) ? 1u : 0u);
#elif defined(PRIM_SC)
    printf("Compacting %llu values.\n", (unsigned long long)size);
    values.resize(size);
    for (uint64_t i = 0; i < size; ++i) values[i] = 1u + (uint32_t)rng.below(65535);
#else
    values.resize(size);
    for (uint64_t i = 0; i < size; ++i) values[i] = (uint32_t)(i + 1);
    printf("Using %llu values.\n", (unsigned long long)size);
#endif
  } else if (!cli::load_values(file, false, &values)) {
    printf("[!!!] could not read '%s'\n", file.c_str());
    return 1;
  }
  const uint64_t n = values.size();
#if defined(PRIM_SC) || defined(PRIM_SEGSCAN)
  std::vector<uint32_t> heads(n);
  for (uint64_t i = 0; i < n; ++i) heads[i] = i == 0 ? 1u : (rng.bern(freq) ? 1u : 0u);
#endif
  const int dev = cli::pick_device(device, true);
  if (dev < 0) {
    printf("No device found.\n");
    return 1;
  }
  wahgpu_ctx* ctx = nullptr;
  if (wahgpu_create(&ctx, dev, 0) != WAHGPU_OK) {
    printf("[!!!] wahgpu_create failed\n");
    return 1;
  }
  const uint64_t nn = n ? n : 1;
  uint32_t *d_in = nullptr, *d_aux = nullptr, *d_out = nullptr;
  HIPX(hipMalloc(&d_in, nn * 4));
  HIPX(hipMalloc(&d_aux, nn * 4));
  HIPX(hipMalloc(&d_out, nn * 4));
  if (n) HIPX(hipMemcpy(d_in, values.data(), n * 4, hipMemcpyHostToDevice));
#if defined(PRIM_SC) || defined(PRIM_SEGSCAN)
  if (n) HIPX(hipMemcpy(d_aux, heads.data(), n * 4, hipMemcpyHostToDevice));
#endif
  auto t0 = std::chrono::steady_clock::now();
  int rc;
  uint64_t count = n;
#if defined(PRIM_SCAN)
  rc = wahgpu_exclusive_scan(ctx, d_in, n, d_out);
#elif defined(PRIM_SC)
  rc = wahgpu_stream_compaction(ctx, d_in, d_aux, n, d_out, &count);
#else
  rc = wahgpu_segmented_inclusive_scan(ctx, d_in, d_aux, n, d_out);
#endif
  if (rc == WAHGPU_OK) rc = wahgpu_synchronize(ctx);
  const double us =
      std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
  if (rc != WAHGPU_OK) {
    printf("[!!!] %s\n", wahgpu_last_error(ctx));
    return 1;
  }
  std::vector<uint32_t> got(count);
  if (count) HIPX(hipMemcpy(got.data(), d_out, count * 4, hipMemcpyDeviceToHost));
  int status = 0;
#if defined(PRIM_SCAN)
  // src/scan.cpp:246-261
  uint32_t curr = 0;
  bool error = false;
  for (uint64_t i = 0; i < n; ++i) {
    if (got[i] != curr) {
      printf("%u --> %u ***** should be %u\n", values[i], got[i], curr);
      error = true;
    }
    curr += values[i];
  }
  if (!error) printf("Success!\n");
  status = error;
#elif defined(PRIM_SC)
  // src/sc.cpp:268-278
  std::vector<uint32_t> expected;
  for (uint64_t i = 0; i < n; ++i)
    if (heads[i]) expected.push_back(values[i]);
  const bool fail = expected != got;
  printf("%s\n", fail ? "FAILURE" : "SUCCESS");
  status = fail;
#else
  // src/segmented_scan.cpp:336-358
  std::vector<uint32_t> expected(n);
  uint32_t run = 0;
  for (uint64_t i = 0; i < n; ++i) {
    run = heads[i] ? values[i] : run + values[i];
    expected[i] = run;
  }
  if (expected != got) {
    printf("Expected different result\n");
    printf("   idx   ||    val   | expected | received |\n");
    uint64_t shown = 0;
    for (uint64_t i = 0; i < n && shown < 32; ++i) {
      if (expected[i] == got[i]) continue;
      printf("%8llu || %8u | %8u | %8u | !!!!\n", (unsigned long long)i, values[i], expected[i],
             got[i]);
      ++shown;
    }
    printf("Failure\n");
    status = 1;
  } else {
    printf("Success\n");
  }
#endif
  if (print)
    for (uint64_t i = 0; i < count; ++i) printf("%u\n", got[i]);
  if (timed) printf("%.0f us\n", us);
  (void)hipFree(d_in);
  (void)hipFree(d_aux);
  (void)hipFree(d_out);
  wahgpu_destroy(ctx);
  return status;
}
