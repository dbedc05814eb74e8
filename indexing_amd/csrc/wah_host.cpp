// wah_host.cpp -- host-side formats of libwahgpu: value-file readers, the
// WAH bitmap decoder, and the seeded column generator.  No GPU needed.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "wahgpu.h"

namespace {

bool is_space(unsigned char c) {
  // isspace() in the "C" locale
  return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r';
}

bool slurp(const char* path, std::vector<unsigned char>* buf) {
  FILE* f = fopen(path, "rb");
  if (!f) return false;
  if (fseek(f, 0, SEEK_END) != 0) { fclose(f); return false; }
  long sz = ftell(f);
  if (sz < 0) { fclose(f); return false; }
  fseek(f, 0, SEEK_SET);
  buf->resize((size_t)sz);
  size_t got = sz ? fread(buf->data(), 1, (size_t)sz, f) : 0;
  fclose(f);
  buf->resize(got);
  return true;
}

// splitmix64 (Steele, Lea, Flood 2014): the finalizer and the Weyl step.
inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
constexpr uint64_t kGamma = 0x9E3779B97F4A7C15ull;
// The i-th 64-bit draw of stream `seed`.
inline uint64_t draw(uint64_t seed, uint64_t i) { return mix64(mix64(seed) + (i + 1) * kGamma); }
// Uniform integer in [0, range) from the high 32 bits (multiply-shift).
inline uint32_t below(uint64_t r, uint64_t range) {
  return (uint32_t)(((r >> 32) * range) >> 32);
}
// Uniform double in [0, 1).
inline double unit(uint64_t r) { return (double)(r >> 11) * (1.0 / 9007199254740992.0); }

// Inverse-CDF sampler for Zipf(s) over ranks 1..card.
struct Zipf {
  std::vector<double> cdf;
  Zipf(uint32_t card, double s) : cdf(card) {
    double acc = 0.0;
    for (uint32_t k = 0; k < card; ++k) {
      acc += pow((double)(k + 1), -s);
      cdf[k] = acc;
    }
    for (auto& x : cdf) x /= acc;
  }
  uint32_t rank0(double u) const {  // rank - 1
    size_t i = std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin();
    return (uint32_t)std::min(i, cdf.size() - 1);
  }
};

}  // namespace

extern "C" {

// Parses like `std::ifstream >> uint32_t` in the "C" locale (the readers of
// src/cpu.cpp:298-302, src/phases.cpp:363-366, src/gpu.cpp:296-300): skip
// white space; optional '+'/'-'; at least one decimal digit; a value above
// 2^32-1 fails the extraction; '-' negates modulo 2^32.  The first failed
// extraction ends the input (the `while (source >> next)` loop).
int wahgpu_read_values_text(const char* path, uint32_t** values, uint64_t* n) {
  if (!path || !values || !n) return WAHGPU_EINVAL;
  *values = nullptr;
  *n = 0;
  std::vector<unsigned char> buf;
  if (!slurp(path, &buf)) return WAHGPU_EIO;
  size_t cap = buf.size() / 2 + 1;  // every value takes >= 2 bytes but the last
  uint32_t* out = (uint32_t*)malloc(cap * sizeof(uint32_t));
  if (!out) return WAHGPU_ENOMEM;
  const unsigned char* p = buf.data();
  const unsigned char* e = p + buf.size();
  uint64_t cnt = 0;
  for (;;) {
    while (p < e && is_space(*p)) ++p;
    if (p == e) break;
    bool neg = false;
    if (*p == '+' || *p == '-') { neg = *p == '-'; ++p; }
    if (p == e || *p < '0' || *p > '9') break;
    uint64_t acc = 0;
    bool overflow = false;
    while (p < e && *p >= '0' && *p <= '9') {
      acc = acc * 10u + (uint64_t)(*p - '0');
      if (acc > 0xFFFFFFFFull) { overflow = true; acc = 0xFFFFFFFFull; }
      ++p;
    }
    if (overflow) break;
    out[cnt++] = neg ? (uint32_t)(0u - (uint32_t)acc) : (uint32_t)acc;
  }
  *values = out;
  *n = cnt;
  return WAHGPU_OK;
}

int wahgpu_read_values_binary(const char* path, uint32_t** values, uint64_t* n) {
  if (!path || !values || !n) return WAHGPU_EINVAL;
  *values = nullptr;
  *n = 0;
  std::vector<unsigned char> buf;
  if (!slurp(path, &buf)) return WAHGPU_EIO;
  uint64_t cnt = buf.size() / 4;
  uint32_t* out = (uint32_t*)malloc((cnt ? cnt : 1) * sizeof(uint32_t));
  if (!out) return WAHGPU_ENOMEM;
  memcpy(out, buf.data(), cnt * 4);  // little-endian host
  *values = out;
  *n = cnt;
  return WAHGPU_OK;
}

void wahgpu_free(void* p) { free(p); }

// src/cpu.cpp:47-74 (== src/gpu.cpp:65-94): literal words (bit 31 set) give
// their 31 payload bits LSB first; other words repeat bit 30 for 31 x (bits
// 0..29) positions; trailing '0's are trimmed.
uint64_t wahgpu_decode_bitmap(const uint32_t* words, uint64_t len, char* out, uint64_t cap) {
  uint64_t pos = 0, last_one = 0;
  for (uint64_t w = 0; w < len; ++w) {
    uint32_t block = words[w];
    if (block & 0x80000000u) {
      for (int i = 0; i < 31; ++i) {
        bool one = (block >> i) & 1u;
        if (pos < cap) out[pos] = one ? '1' : '0';
        ++pos;
        if (one) last_one = pos;
      }
    } else {
      bool one = (block & 0x40000000u) != 0;
      uint64_t times = (uint64_t)(block & 0x3FFFFFFFu) * 31u;
      if (one) {
        for (uint64_t i = 0; i < times; ++i) {
          if (pos < cap) out[pos] = '1';
          ++pos;
        }
        if (times) last_one = pos;
      } else {
        uint64_t end = pos + times;
        if (pos < cap) memset(out + pos, '0', (size_t)(std::min(end, cap) - pos));
        pos = end;
      }
    }
  }
  return last_one;
}

int wahgpu_generate_range(const wahgpu_gen_params* p, uint32_t* out, uint64_t start,
                          uint64_t count) {
  if (!p || (count && !out)) return WAHGPU_EINVAL;
  switch (p->dist) {
    case WAHGPU_DIST_UNIFORM: {
      if (p->to < p->from) return WAHGPU_EINVAL;
      uint64_t range = (uint64_t)p->to - p->from + 1;
      for (uint64_t i = 0; i < count; ++i)
        out[i] = p->from + below(draw(p->seed, start + i), range);
      return WAHGPU_OK;
    }
    case WAHGPU_DIST_ZIPF: {
      if (p->cardinality == 0) return WAHGPU_EINVAL;
      Zipf z(p->cardinality, p->zipf_s);
      for (uint64_t i = 0; i < count; ++i) out[i] = z.rank0(unit(draw(p->seed, start + i)));
      return WAHGPU_OK;
    }
    case WAHGPU_DIST_ZIPF_RUNS: {
      if (p->cardinality == 0 || p->run_mean < 1.0) return WAHGPU_EINVAL;
      Zipf z(p->cardinality, p->zipf_s);
      // Run j: length from draw 2j (Geometric(1/run_mean) on {1, 2, ...}),
      // value from draw 2j+1.  Runs before `start` are skipped by length.
      const double q = 1.0 - 1.0 / p->run_mean;
      const double lq = q > 0.0 ? log(q) : 0.0;
      const uint64_t end = start + count;
      uint64_t i = 0, j = 0;
      while (i < end) {
        double u = unit(draw(p->seed, 2 * j));
        uint64_t len = 1;
        if (lq < 0.0) len = 1 + (uint64_t)floor(log1p(-u) / lq);
        uint64_t run_end = i + len;
        if (run_end > start) {
          uint32_t v = z.rank0(unit(draw(p->seed, 2 * j + 1)));
          uint64_t a = i > start ? i : start, b = run_end < end ? run_end : end;
          for (uint64_t x = a; x < b; ++x) out[x - start] = v;
        }
        i = run_end;
        ++j;
      }
      return WAHGPU_OK;
    }
    default:
      return WAHGPU_EINVAL;
  }
}

int wahgpu_generate(const wahgpu_gen_params* p, uint32_t* out, uint64_t n) {
  return wahgpu_generate_range(p, out, 0, n);
}

}  // extern "C"
