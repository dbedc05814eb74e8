// wah_index.hpp -- exact stitching of partial WAH indexes built over
// consecutive row ranges (the merger side of the indexer contract).
#pragma once

#include <stdint.h>

#include <unordered_map>
#include <vector>

namespace wah {

class Merger {
 public:
  // Appends one partial index (keys ascending, offsets, `length` words).
  // Partials must arrive in row order.  Returns false on malformed input.
  bool add(uint64_t keycnt, uint64_t length, const uint32_t* keys, const uint32_t* offsets,
           const uint32_t* words);
  uint64_t keycnt() const { return keys_.size(); }
  uint64_t length() const { return total_; }
  void export_to(uint32_t* keys, uint32_t* offsets, uint32_t* lengths, uint32_t* words) const;
  const std::vector<uint32_t>* find(uint32_t key) const;

 private:
  struct Key {
    std::vector<uint32_t> words;
    uint32_t last_chunk = 0;  // chunk of the last literal
  };
  bool append(Key& k, bool fresh, const uint32_t* w, uint64_t len);
  std::unordered_map<uint32_t, Key> keys_;
  uint64_t total_ = 0;
};

}  // namespace wah
