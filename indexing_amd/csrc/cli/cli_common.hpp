// cli_common.hpp -- shared pieces of the phases / gpu / generate programs:
// option parsing in the reference's style (short "-f X", long "--data-file X"
// or "--data-file=X"), value-file loading and index printing.
#pragma once

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "wahgpu.h"

namespace cli {

struct Opt {
  const char* shrt;  // "-f" or nullptr
  const char* lng;   // "--data-file"
  bool flag;         // takes no value
  std::string* sval;
  bool* bval;
  const char* help;
};

inline void usage(const char* prog, const std::vector<Opt>& opts) {
  printf("%s options:\n", prog);
  for (auto& o : opts)
    printf("  %-4s %-20s %s\n", o.shrt ? o.shrt : "", o.lng, o.help);
}

// Returns 0 on success, 1 if help was printed, 2 on a bad argument.
inline int parse(int argc, char** argv, const std::vector<Opt>& opts) {
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    if (a == "-h" || a == "--help" || a == "-?") {
      usage(argv[0], opts);
      return 1;
    }
    bool matched = false;
    for (auto& o : opts) {
      std::string lng = o.lng;
      bool eq = a.compare(0, lng.size() + 1, lng + "=") == 0;
      if (a == lng || (o.shrt && a == o.shrt) || eq) {
        matched = true;
        if (o.flag) {
          *o.bval = true;
        } else if (eq) {
          *o.sval = a.substr(lng.size() + 1);
        } else if (i + 1 < argc) {
          *o.sval = argv[++i];
        } else {
          fprintf(stderr, "%s: option %s needs a value\n", argv[0], a.c_str());
          return 2;
        }
        break;
      }
    }
    if (!matched) {
      fprintf(stderr, "%s: unknown option '%s' (see --help)\n", argv[0], a.c_str());
      return 2;
    }
  }
  return 0;
}

inline uint32_t to_u32(const std::string& s) { return (uint32_t)strtoul(s.c_str(), nullptr, 10); }

// Reads a value file: text (the reference format) unless binary is set.
inline bool load_values(const std::string& path, bool binary, std::vector<uint32_t>* out) {
  uint32_t* v = nullptr;
  uint64_t n = 0;
  int rc = binary ? wahgpu_read_values_binary(path.c_str(), &v, &n)
                  : wahgpu_read_values_text(path.c_str(), &v, &n);
  if (rc == WAHGPU_EIO) {  // an unreadable ifstream extracts nothing
    out->clear();
    return true;
  }
  if (rc != WAHGPU_OK) return false;
  out->assign(v, v + n);
  wahgpu_free(v);
  return true;
}

inline std::string decoded(const uint32_t* words, uint64_t len) {
  std::string s;
  uint64_t need = wahgpu_decode_bitmap(words, len, nullptr, 0);
  s.resize(need);
  if (need) wahgpu_decode_bitmap(words, len, &s[0], need);
  return s;
}

// The -p listing of src/cpu.cpp:329-351: every key with its bitmap, and an
// empty line for each missing value below it.
inline void print_index(const std::vector<uint32_t>& keys, const std::vector<uint32_t>& offsets,
                        const std::vector<uint32_t>& lengths, const std::vector<uint32_t>& words,
                        uint64_t keycnt) {
  uint32_t next = 0;
  std::string line;
  for (uint64_t i = 0; i < keycnt; ++i) {
    uint32_t key = keys[i];
    if (key != next)
      for (uint32_t j = next; j < key; ++j) printf("%u\t\n", j);
    line = decoded(words.data() + offsets[i], lengths[i]);
    printf("%u\t%s\n", key, line.c_str());
    next = key + 1;
  }
}

inline int pick_device(const std::string& name, bool fallback_first) {
  int dev = wahgpu_find_device(name.c_str());
  if (dev < 0 && fallback_first) dev = wahgpu_find_device("");
  return dev;
}

}  // namespace cli
