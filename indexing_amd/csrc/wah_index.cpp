// wah_index.cpp -- exact stitching of partial WAH indexes (host only).
//
// A partial built with global row ids encodes each key as
//   [head fill = first chunk (dropped if 0)], literal, [gap fill], literal, ...
// so the run's chunks are global.  Appending a run to a key that already ends
// at chunk `last`: if the run starts in `last`, its first literal is OR-ed
// into the key's last literal; otherwise the head fill becomes
// first - last - 1 (dropped when zero).  This is exactly what a single build
// over the concatenated rows produces (src/cpu.cpp:223-257).
#include "wah_index.hpp"

#include <algorithm>

#include "wahgpu.h"

namespace wah {

namespace {
inline bool is_literal(uint32_t w) { return (w & 0x80000000u) != 0; }
}  // namespace

bool Merger::append(Key& k, bool fresh, const uint32_t* w, uint64_t len) {
  if (len == 0) return false;
  uint64_t i = 0;
  uint64_t first = 0;
  if (!is_literal(w[0])) {
    first = w[0] & 0x3FFFFFFFu;
    i = 1;
  }
  if (i >= len || !is_literal(w[i])) return false;
  uint64_t added = 0;
  if (fresh) {
    if (first) { k.words.push_back((uint32_t)first); ++added; }
  } else if (first == k.last_chunk) {
    k.words.back() |= w[i];
    ++i;
    if (i == len) { k.last_chunk = (uint32_t)first; return true; }
    // the remaining run continues after chunk `first`
    uint64_t next = first + 1;
    for (uint64_t j = i; j < len; ++j) {
      k.words.push_back(w[j]);
      ++added;
      if (is_literal(w[j])) ++next; else next += w[j] & 0x3FFFFFFFu;
    }
    k.last_chunk = (uint32_t)(next - 1);
    total_ += added;
    return true;
  } else if (first > k.last_chunk) {
    uint64_t gap = first - k.last_chunk - 1;
    if (gap) { k.words.push_back((uint32_t)gap); ++added; }
  } else {
    return false;  // rows out of order
  }
  uint64_t next = first;
  for (uint64_t j = i; j < len; ++j) {
    k.words.push_back(w[j]);
    ++added;
    if (is_literal(w[j])) ++next; else next += w[j] & 0x3FFFFFFFu;
  }
  k.last_chunk = (uint32_t)(next - 1);
  total_ += added;
  return true;
}

bool Merger::add(uint64_t keycnt, uint64_t length, const uint32_t* keys, const uint32_t* offsets,
                 const uint32_t* words) {
  for (uint64_t i = 0; i < keycnt; ++i) {
    uint64_t off = offsets[i];
    uint64_t end = (i + 1 == keycnt) ? length : offsets[i + 1];
    if (end < off || end > length) return false;
    auto it = keys_.find(keys[i]);
    bool fresh = it == keys_.end();
    Key& k = fresh ? keys_[keys[i]] : it->second;
    if (!append(k, fresh, words + off, end - off)) return false;
  }
  return true;
}

void Merger::export_to(uint32_t* keys, uint32_t* offsets, uint32_t* lengths,
                       uint32_t* words) const {
  std::vector<uint32_t> order;
  order.reserve(keys_.size());
  for (auto& kv : keys_) order.push_back(kv.first);
  std::sort(order.begin(), order.end());
  uint64_t pos = 0;
  for (size_t i = 0; i < order.size(); ++i) {
    const Key& k = keys_.at(order[i]);
    if (keys) keys[i] = order[i];
    if (offsets) offsets[i] = (uint32_t)pos;
    if (lengths) lengths[i] = (uint32_t)k.words.size();
    if (words) std::copy(k.words.begin(), k.words.end(), words + pos);
    pos += k.words.size();
  }
}

const std::vector<uint32_t>* Merger::find(uint32_t key) const {
  auto it = keys_.find(key);
  return it == keys_.end() ? nullptr : &it->second.words;
}

}  // namespace wah

struct wahgpu_merger {
  wah::Merger m;
};

extern "C" {

int wahgpu_merger_create(wahgpu_merger** out) {
  if (!out) return WAHGPU_EINVAL;
  *out = new (std::nothrow) wahgpu_merger();
  return *out ? WAHGPU_OK : WAHGPU_ENOMEM;
}

void wahgpu_merger_destroy(wahgpu_merger* m) { delete m; }

int wahgpu_merger_add(wahgpu_merger* m, uint64_t keycnt, uint64_t length, const uint32_t* keys,
                      const uint32_t* offsets, const uint32_t* words) {
  if (!m || (keycnt && (!keys || !offsets || !words))) return WAHGPU_EINVAL;
  return m->m.add(keycnt, length, keys, offsets, words) ? WAHGPU_OK : WAHGPU_EINVAL;
}

uint64_t wahgpu_merger_keycnt(const wahgpu_merger* m) { return m ? m->m.keycnt() : 0; }
uint64_t wahgpu_merger_length(const wahgpu_merger* m) { return m ? m->m.length() : 0; }

int wahgpu_merger_export(const wahgpu_merger* m, uint32_t* keys, uint32_t* offsets,
                         uint32_t* lengths, uint32_t* words) {
  if (!m) return WAHGPU_EINVAL;
  m->m.export_to(keys, offsets, lengths, words);
  return WAHGPU_OK;
}

int wahgpu_merger_key_words(const wahgpu_merger* m, uint32_t key, const uint32_t** words,
                            uint64_t* len) {
  if (!m || !words || !len) return WAHGPU_EINVAL;
  const std::vector<uint32_t>* v = m->m.find(key);
  *words = v ? v->data() : nullptr;
  *len = v ? v->size() : 0;
  return WAHGPU_OK;
}

}  // extern "C"
