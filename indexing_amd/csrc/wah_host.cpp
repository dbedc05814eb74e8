// wah_host.cpp -- host-side formats of libwahgpu: value-file readers, the
// WAH bitmap decoder, and the seeded column generator.  No GPU needed.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <thread>
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

// Inverse-CDF sampler for Zipf(s) over ranks 1..card.  rank0(u) is the
// first rank whose cumulative probability exceeds u (std::upper_bound over
// the CDF), found from a guide table (Chen & Asau, 1974): guide[g] is that
// rank for u = g / G, and the answer for u in [g / G, (g + 1) / G) is at or
// after it, usually within a step or two.
struct Zipf {
  std::vector<double> cdf;
  std::vector<uint32_t> guide;
  Zipf(uint32_t card, double s) : cdf(card) {
    double acc = 0.0;
    for (uint32_t k = 0; k < card; ++k) {
      acc += pow((double)(k + 1), -s);
      cdf[k] = acc;
    }
    for (auto& x : cdf) x /= acc;
    const uint32_t G = card;
    guide.resize(G);
    uint32_t i = 0;
    for (uint32_t g = 0; g < G; ++g) {
      const double u = (double)g / (double)G;
      while (i < card && cdf[i] <= u) ++i;
      guide[g] = i;
    }
  }
  uint32_t rank0(double u) const {  // rank - 1
    const size_t G = guide.size();
    size_t g = (size_t)(u * (double)G);
    if (g >= G) g = G - 1;
    size_t i = guide[g];
    while (i < cdf.size() && cdf[i] <= u) ++i;
    return (uint32_t)std::min(i, cdf.size() - 1);
  }
};

// One `ifstream >> uint32_t` extraction loop (src/cpu.cpp:298-302) over
// [p, e): values go to out (room for (e - p + 1) / 2 of them: a value takes
// a digit and, but for the last, a separator); *failed is set when an
// extraction failed before the end (the loop ends there).
uint64_t parse_span(const unsigned char* p, const unsigned char* e, uint32_t* out, bool* failed) {
  static const struct SpaceTable {
    bool t[256];
    SpaceTable() : t() {
      for (int c = 0; c < 256; ++c) t[c] = is_space((unsigned char)c);
    }
  } space;
  uint64_t cnt = 0;
  *failed = false;
  for (;;) {
    while (p < e && space.t[*p]) ++p;
    if (p == e) break;
    bool neg = false;
    if (*p == '+' || *p == '-') { neg = *p == '-'; ++p; }
    if (p == e || (unsigned)(*p - '0') > 9u) { *failed = true; break; }
    // up to 10 digits cannot overflow 64 bits; longer tokens (leading
    // zeros) continue with a saturating accumulator
    const unsigned char* q = p;
    const unsigned char* stop = e - p > 10 ? p + 10 : e;
    uint64_t acc = 0;
    while (q < stop && (unsigned)(*q - '0') <= 9u) acc = acc * 10u + (uint64_t)(*q++ - '0');
    while (q < e && (unsigned)(*q - '0') <= 9u) {
      acc = acc * 10u + (uint64_t)(*q++ - '0');
      if (acc > 0xFFFFFFFFull) acc = 0x100000000ull;
    }
    p = q;
    if (acc > 0xFFFFFFFFull) { *failed = true; break; }
    out[cnt++] = neg ? (uint32_t)(0u - (uint32_t)acc) : (uint32_t)acc;
  }
  return cnt;
}

// A read-only mapping of a whole file (empty files map to nothing).
struct Mapped {
  const unsigned char* p = nullptr;
  size_t size = 0;
  int fd = -1;
  bool open(const char* path) {
    fd = ::open(path, O_RDONLY);
    if (fd < 0) return false;
    struct stat st;
    if (fstat(fd, &st) != 0) return false;
    size = (size_t)st.st_size;
    if (size == 0) return true;
    // populated at once: one pass in the kernel instead of a fault per page
    void* m = mmap(nullptr, size, PROT_READ, MAP_PRIVATE | MAP_POPULATE, fd, 0);
    if (m == MAP_FAILED) return false;
    p = static_cast<const unsigned char*>(m);
    return true;
  }
  ~Mapped() {
    if (p) munmap(const_cast<unsigned char*>(p), size);
    if (fd >= 0) close(fd);
  }
};

// Heap memory (released with free()) backed by 2 MiB pages where the kernel
// allows them: a gigabyte of parsed values then takes hundreds of page
// faults instead of a quarter million.
void* alloc_large(size_t bytes) {
  constexpr size_t kHuge = (size_t)2 << 20;
  if (bytes < kHuge) return malloc(bytes ? bytes : 1);
  const size_t cap = (bytes + kHuge - 1) / kHuge * kHuge;
  void* p = aligned_alloc(kHuge, cap);
  if (p) (void)madvise(p, cap, MADV_HUGEPAGE);
  return p;
}

unsigned auto_threads(int threads) {
  if (threads > 0) return (unsigned)threads;
  unsigned h = std::thread::hardware_concurrency();
  return h ? std::min(h, 16u) : 4u;
}

// Decimal digits of v, written at out; returns their number.
inline int format_u32(uint32_t v, char* out) {
  char tmp[10];
  int n = 0;
  do {
    tmp[n++] = (char)('0' + v % 10u);
    v /= 10u;
  } while (v);
  for (int i = 0; i < n; ++i) out[i] = tmp[n - 1 - i];
  return n;
}

}  // namespace

extern "C" {

// Parses like `std::ifstream >> uint32_t` in the "C" locale (the readers of
// src/cpu.cpp:298-302, src/phases.cpp:363-366, src/gpu.cpp:296-300): skip
// white space; optional '+'/'-'; at least one decimal digit; a value above
// 2^32-1 fails the extraction; '-' negates modulo 2^32.  The first failed
// extraction ends the input (the `while (source >> next)` loop).
//
// In parallel: the mapped file is cut into one span per thread at white
// space, so no token straddles two spans; each thread parses its span as the
// loop would; the values are those of the spans before the first span whose
// parse failed, plus that span's values before its failure.
int wahgpu_read_values_text_threads(const char* path, int threads, uint32_t** values,
                                    uint64_t* n) {
  if (!path || !values || !n || threads < 0) return WAHGPU_EINVAL;
  *values = nullptr;
  *n = 0;
  Mapped f;
  if (!f.open(path)) return WAHGPU_EIO;
  const unsigned char* buf = f.p;
  const size_t size = f.size;
  unsigned T = auto_threads(threads);
  if (threads == 0 && size < ((size_t)1 << 20)) T = 1;  // not worth the threads
  if ((size_t)T > size) T = size ? (unsigned)size : 1u;
  std::vector<size_t> cut(T + 1, size);
  cut[0] = 0;
  for (unsigned i = 1; i < T; ++i) {
    size_t b = std::max(cut[i - 1], size / T * i);
    while (b < size && !is_space(buf[b])) ++b;
    cut[i] = b;
  }
  std::vector<uint32_t*> part(T, nullptr);
  std::vector<uint64_t> cnt(T, 0);
  std::vector<char> failed(T, 0), oom(T, 0);
  auto work = [&](unsigned i) {
    const size_t len = cut[i + 1] - cut[i];
    part[i] = (uint32_t*)alloc_large(((len + 1) / 2 + 1) * sizeof(uint32_t));
    if (!part[i]) { oom[i] = 1; return; }
    bool fl = false;
    cnt[i] = len ? parse_span(buf + cut[i], buf + cut[i + 1], part[i], &fl) : 0;
    failed[i] = fl;
  };
  std::vector<std::thread> th;
  for (unsigned i = 1; i < T; ++i) th.emplace_back(work, i);
  work(0);
  for (auto& t : th) t.join();
  auto free_parts = [&]() {
    for (auto q : part) free(q);
  };
  for (unsigned i = 0; i < T; ++i)
    if (oom[i]) { free_parts(); return WAHGPU_ENOMEM; }
  unsigned last = T;  // spans [0, last) contribute
  for (unsigned i = 0; i < T; ++i)
    if (failed[i]) { last = i + 1; break; }
  std::vector<uint64_t> at(last + 1, 0);
  for (unsigned i = 0; i < last; ++i) at[i + 1] = at[i] + cnt[i];
  uint32_t* out = (uint32_t*)alloc_large((at[last] ? at[last] : 1) * sizeof(uint32_t));
  if (!out) { free_parts(); return WAHGPU_ENOMEM; }
  th.clear();
  for (unsigned i = 1; i < last; ++i)
    th.emplace_back([&, i]() { memcpy(out + at[i], part[i], cnt[i] * sizeof(uint32_t)); });
  if (last) memcpy(out, part[0], cnt[0] * sizeof(uint32_t));
  for (auto& t : th) t.join();
  free_parts();
  *values = out;
  *n = at[last];
  return WAHGPU_OK;
}

int wahgpu_read_values_text(const char* path, uint32_t** values, uint64_t* n) {
  return wahgpu_read_values_text_threads(path, 0, values, n);
}

// The generator's text format (src/generate_test_data.cpp:48-54): the values
// in decimal separated by sep, then a newline.  Threads format spans of
// values into their own buffers, which are written in order.
int wahgpu_write_values_text(const char* path, const uint32_t* values, uint64_t n,
                             const char* sep, int threads) {
  if (!path || (n && !values) || !sep || threads < 0) return WAHGPU_EINVAL;
  const size_t sl = strlen(sep);
  FILE* f = fopen(path, "wb");
  if (!f) return WAHGPU_EIO;
  unsigned T = auto_threads(threads);
  if (threads == 0 && n < ((uint64_t)1 << 16)) T = 1;
  const uint64_t per = (n + T - 1) / (T ? T : 1);
  std::vector<std::vector<char>> text(T);
  auto work = [&](unsigned i) {
    const uint64_t a = std::min<uint64_t>(n, per * i), b = std::min<uint64_t>(n, a + per);
    std::vector<char>& o = text[i];
    o.resize((b - a) * (10 + sl) + 1);
    char* w = o.data();
    for (uint64_t k = a; k < b; ++k) {
      w += format_u32(values[k], w);
      if (k + 1 < n) {
        memcpy(w, sep, sl);
        w += sl;
      }
    }
    o.resize((size_t)(w - o.data()));
  };
  std::vector<std::thread> th;
  for (unsigned i = 1; i < T; ++i) th.emplace_back(work, i);
  work(0);
  for (auto& t : th) t.join();
  bool ok = true;
  for (auto& o : text)
    if (!o.empty()) ok &= fwrite(o.data(), 1, o.size(), f) == o.size();
  ok &= fputc('\n', f) != EOF;
  ok &= fclose(f) == 0;
  return ok ? WAHGPU_OK : WAHGPU_EIO;
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

// Four interleaved multiply-xorshift chains over 64-bit words; each chain
// depends on the order of its words, and the lengths are folded in.
static inline uint64_t chk_mix(uint64_t h, uint64_t v) {
  h ^= v;
  h *= 0x9E3779B97F4A7C15ull;
  return h ^ (h >> 29);
}
static uint64_t chk_array(uint64_t h, const uint32_t* p, uint64_t n) {
  uint64_t a[4] = {h ^ 0x243F6A8885A308D3ull, h ^ 0x13198A2E03707344ull,
                   h ^ 0xA4093822299F31D0ull, h ^ 0x082EFA98EC4E6C89ull};
  uint64_t i = 0;
  for (; i + 8 <= n; i += 8)
    for (int k = 0; k < 4; ++k)
      a[k] = chk_mix(a[k], (uint64_t)p[i + 2 * k] | ((uint64_t)p[i + 2 * k + 1] << 32));
  for (int k = 0; i < n; ++i, k = (k + 1) & 3) a[k] = chk_mix(a[k], p[i]);
  uint64_t r = chk_mix(n, a[0]);
  for (int k = 1; k < 4; ++k) r = chk_mix(r, a[k]);
  return r;
}

uint64_t wahgpu_index_checksum(const uint32_t* keys, const uint32_t* offsets,
                               const uint32_t* lengths, uint64_t keycnt, const uint32_t* words,
                               uint64_t length) {
  uint64_t h = 0x6A09E667F3BCC908ull;
  h = chk_array(h, keys, keys ? keycnt : 0);
  h = chk_array(h, offsets, offsets ? keycnt : 0);
  h = chk_array(h, lengths, lengths ? keycnt : 0);
  return chk_array(h, words, words ? length : 0);
}

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
