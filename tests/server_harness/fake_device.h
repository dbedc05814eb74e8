// fake_device.h -- what the server harness's driver knows of the fake build
// (fake_device.cpp): the word it writes per distinct value and the marker
// value that makes a column report a kernel fault.
#pragma once

#include <stdint.h>

#include <vector>

constexpr uint32_t kFakeFaultMarker = 0xFFFFFFF0u;

uint32_t fake_word(uint64_t count, uint32_t rid_base);
uint64_t fake_distinct(const uint32_t* values, uint64_t n);
// The keys and words the fake build writes for one column.
void fake_expected(const uint32_t* values, uint64_t n, uint32_t rid_base,
                   std::vector<uint32_t>* keys, std::vector<uint32_t>* words);
