// generate -- the reference's test-data generator (src/generate_test_data.cpp).
//
//   generate [-f FROM] [-t TO] [-a AMOUNT] [-s SEP] [--seed S]
//            [--dist uniform|zipf|zipf-runs] [--cardinality C] [--zipf-s S]
//            [--run-mean M] [--binary] [-o FILE]
//
// Writes AMOUNT decimal uint32 values drawn uniformly from [FROM, TO]
// separated by SEP, then a newline, to stdout (src/generate_test_data.cpp:43-54;
// defaults 0, 65535, 268500000, "\n").  Without --seed the seed comes from
// std::random_device as in the reference; with it the output is reproducible
// on any host (splitmix64 stream, see wahgpu_generate).  --dist selects Zipf
// ranks (value = rank - 1) or Zipf values in geometric runs; --binary writes
// raw little-endian uint32 instead of text.
#include <stdio.h>

#include <random>
#include <string>
#include <vector>

#include "cli_common.hpp"

int main(int argc, char** argv) {
  std::string from = "0", to = "65535", amount = "268500000", sep = "\n", seed_s, dist = "uniform",
              card = "1048576", zs = "1.2", run_mean = "4096", outfile;
  bool binary = false;
  std::vector<cli::Opt> opts = {
      {"-f", "--from", false, &from, nullptr, "set minimum value (default: 0)"},
      {"-t", "--to", false, &to, nullptr, "set maximum value (default: 65535)"},
      {"-a", "--amount", false, &amount, nullptr, "set number of values (default: 268 500 000)"},
      {"-s", "--separator", false, &sep, nullptr, "separator between numbers (default: \\n)"},
      {nullptr, "--seed", false, &seed_s, nullptr, "seed (default: random_device)"},
      {nullptr, "--dist", false, &dist, nullptr, "uniform | zipf | zipf-runs"},
      {nullptr, "--cardinality", false, &card, nullptr, "zipf support size (default 2^20)"},
      {nullptr, "--zipf-s", false, &zs, nullptr, "zipf exponent (default 1.2)"},
      {nullptr, "--run-mean", false, &run_mean, nullptr, "mean run length (default 4096)"},
      {nullptr, "--binary", true, nullptr, &binary, "write raw little-endian uint32"},
      {"-o", "--output", false, &outfile, nullptr, "write to FILE instead of stdout"},
  };
  int pr = cli::parse(argc, argv, opts);
  if (pr == 1) return 0;
  if (pr == 2) return 2;
  wahgpu_gen_params p{};
  p.from = cli::to_u32(from);
  p.to = cli::to_u32(to);
  p.cardinality = cli::to_u32(card);
  p.zipf_s = atof(zs.c_str());
  p.run_mean = atof(run_mean.c_str());
  if (seed_s.empty()) {
    std::random_device rd;
    p.seed = ((uint64_t)rd() << 32) | rd();
  } else {
    p.seed = strtoull(seed_s.c_str(), nullptr, 10);
  }
  if (dist == "uniform") p.dist = WAHGPU_DIST_UNIFORM;
  else if (dist == "zipf") p.dist = WAHGPU_DIST_ZIPF;
  else if (dist == "zipf-runs") p.dist = WAHGPU_DIST_ZIPF_RUNS;
  else {
    fprintf(stderr, "generate: unknown --dist '%s'\n", dist.c_str());
    return 2;
  }
  uint64_t n = strtoull(amount.c_str(), nullptr, 10);
  FILE* out = stdout;
  if (!outfile.empty() && !(out = fopen(outfile.c_str(), "wb"))) {
    fprintf(stderr, "generate: cannot open '%s'\n", outfile.c_str());
    return 1;
  }
  static char obuf[1 << 20];
  setvbuf(out, obuf, _IOFBF, sizeof(obuf));
  // Values are a function of their position, so batches of the range keep
  // memory bounded without changing the output.
  const uint64_t kBatch = 1 << 22;
  std::vector<uint32_t> v;
  std::string text;
  for (uint64_t at = 0; at < n; at += kBatch) {
    uint64_t m = n - at < kBatch ? n - at : kBatch;
    v.resize(m);
    if (wahgpu_generate_range(&p, v.data(), at, m) != WAHGPU_OK) {
      fprintf(stderr, "generate: bad parameters\n");
      return 2;
    }
    if (binary) {
      fwrite(v.data(), 4, m, out);
    } else {
      text.clear();
      char num[16];
      for (uint64_t i = 0; i < m; ++i) {
        int len = snprintf(num, sizeof(num), "%u", v[i]);
        text.append(num, len);
        if (at + i + 1 < n) text.append(sep);
      }
      fwrite(text.data(), 1, text.size(), out);
    }
  }
  if (!binary) fputc('\n', out);
  if (out != stdout) fclose(out);
  else fflush(out);
  return 0;
}
