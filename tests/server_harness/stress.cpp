// stress.cpp -- drives wah_server.cpp over the fake device from many threads:
// submitters with ragged and empty columns, waits, polls, a second thread
// waiting on the same request as its owner, columns that report a kernel
// fault, an arena small enough that slots wait for memory, and a server
// destroyed with requests still queued.  Every result is compared with what
// the fake build writes for that request's own values and rid_base.  Built
// with -fsanitize=thread and with -fsanitize=address,undefined by the test
// that runs it; exits nonzero on the first wrong result or broken invariant.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include <atomic>
#include <memory>
#include <mutex>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "fake_device.h"
#include "hip/hip_runtime_api.h"
#include "wahgpu.h"

namespace {

std::atomic<int> g_failures{0};

#define CHECK(cond, ...)                                   \
  do {                                                     \
    if (!(cond)) {                                         \
      fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
      fprintf(stderr, __VA_ARGS__);                        \
      fprintf(stderr, "\n");                               \
      ++g_failures;                                        \
    }                                                      \
  } while (0)

struct Job {
  std::vector<uint32_t> values;
  uint32_t rid_base = 0;
  std::vector<uint32_t> keys, offsets, lengths, words;
  wahgpu_result res{};
  uint64_t id = 0;
  bool fault = false;
};

std::unique_ptr<Job> make_job(std::minstd_rand& rng) {
  auto j = std::make_unique<Job>();
  const uint32_t kind = rng() % 8;
  uint64_t n = kind == 0 ? 0 : kind < 4 ? rng() % 200 : kind < 7 ? rng() % 6000 : 20000 + rng() % 40000;
  const uint32_t range = 1 + rng() % (kind == 7 ? 70000 : 300);
  j->values.resize(n);
  for (auto& v : j->values) v = rng() % range;
  j->fault = n > 0 && rng() % 50 == 0;
  if (j->fault) j->values[0] = kFakeFaultMarker;
  j->rid_base = rng() % 1000;
  const uint64_t cap = n + 1;
  j->keys.assign(cap, 0xCDCDCDCDu);
  j->offsets.assign(cap, 0);
  j->lengths.assign(cap, 0);
  j->words.assign(2 * cap, 0);
  j->res.keys = j->keys.data();
  j->res.offsets = j->offsets.data();
  j->res.lengths = j->lengths.data();
  j->res.keys_cap = cap;
  j->res.words = j->words.data();
  j->res.words_cap = 2 * cap;
  return j;
}

void check_job(const Job& j, int status, const char* where) {
  if (j.fault) {
    CHECK(status == WAHGPU_EHIP, "%s: faulting column got status %d", where, status);
    return;
  }
  CHECK(status == WAHGPU_OK, "%s: status %d", where, status);
  if (status != WAHGPU_OK) return;
  std::vector<uint32_t> keys, words;
  fake_expected(j.values.data(), j.values.size(), j.rid_base, &keys, &words);
  CHECK(j.res.counts.keycnt == keys.size() && j.res.counts.length == keys.size(),
        "%s: counts %llu/%llu, expected %zu", where, (unsigned long long)j.res.counts.keycnt,
        (unsigned long long)j.res.counts.length, keys.size());
  if (j.res.counts.keycnt != keys.size()) return;
  if (keys.empty()) return;
  CHECK(memcmp(j.keys.data(), keys.data(), keys.size() * 4) == 0, "%s: keys differ", where);
  CHECK(memcmp(j.words.data(), words.data(), words.size() * 4) == 0, "%s: words differ", where);
  for (size_t i = 0; i < keys.size(); ++i)
    if (j.offsets[i] != i || j.lengths[i] != 1) {
      CHECK(false, "%s: offsets/lengths differ at %zu", where, i);
      break;
    }
}

// Submitters each keep a window of requests in flight and complete them by
// wait, by a poll loop, or (every 7th) with a helper thread waiting on the
// same id while the owner polls: at least one of the two sees the status.
void run_batched(uint32_t slots, uint64_t arena, uint32_t threads, uint32_t per_thread,
                 uint32_t seed) {
  wahgpu_server_config cfg{};
  cfg.device = 0;
  cfg.slots = slots;
  cfg.arena_bytes = arena;
  cfg.max_batch_cols = 32;
  wahgpu_server* s = nullptr;
  CHECK(wahgpu_server_create(&s, &cfg) == WAHGPU_OK, "server_create");
  if (!s) return;
  std::atomic<uint64_t> done{0}, refused{0};
  std::vector<std::thread> th;
  for (uint32_t t = 0; t < threads; ++t)
    th.emplace_back([&, t] {
      std::minstd_rand rng(seed * 977 + t);
      std::vector<std::unique_ptr<Job>> window;
      auto finish = [&](Job& j, size_t idx) {
        int st;
        if (idx % 7 == 3) {
          std::atomic<int> helper_st{WAHGPU_EINVAL};
          std::thread helper([&] { helper_st = wahgpu_server_wait(s, j.id); });
          do {
            st = wahgpu_server_poll(s, j.id);
            if (st == WAHGPU_EBUSY) usleep(50);
          } while (st == WAHGPU_EBUSY);
          helper.join();
          // the one that erases the request returns its status; the other
          // returns it too if it had found the request, EINVAL if not
          const int h = helper_st.load();
          CHECK(st != WAHGPU_EINVAL || h != WAHGPU_EINVAL, "owner and helper both lost %llu",
                (unsigned long long)j.id);
          CHECK(st == WAHGPU_EINVAL || h == WAHGPU_EINVAL || st == h,
                "owner %d and helper %d disagree", st, h);
          if (st == WAHGPU_EINVAL) st = h;
        } else if (idx % 3 == 0) {
          do {
            st = wahgpu_server_poll(s, j.id);
            if (st == WAHGPU_EBUSY) usleep(20);
          } while (st == WAHGPU_EBUSY);
        } else {
          st = wahgpu_server_wait(s, j.id);
        }
        check_job(j, st, "batched");
        CHECK(wahgpu_server_poll(s, j.id) == WAHGPU_EINVAL, "completed id still known");
        ++done;
      };
      for (uint32_t i = 0; i < per_thread; ++i) {
        auto j = make_job(rng);
        const int rc = wahgpu_server_submit(s, j->values.data(), j->values.size(), j->rid_base,
                                            &j->res, &j->id);
        if (rc == WAHGPU_ERANGE) {
          ++refused;
          continue;
        }
        CHECK(rc == WAHGPU_OK, "submit %d", rc);
        if (rc != WAHGPU_OK) continue;
        window.push_back(std::move(j));
        if (window.size() >= 6) {
          const size_t k = rng() % window.size();
          finish(*window[k], i);
          window.erase(window.begin() + (long)k);
        }
      }
      for (size_t i = 0; i < window.size(); ++i) finish(*window[i], i);
    });
  for (auto& x : th) x.join();
  wahgpu_server_stats st{};
  CHECK(wahgpu_server_get_stats(s, &st) == WAHGPU_OK, "stats");
  CHECK(st.requests == done.load(), "requests %llu, completed %llu",
        (unsigned long long)st.requests, (unsigned long long)done.load());
  CHECK(st.arena_bytes == arena && st.arena_peak <= arena, "arena peak %llu of %llu",
        (unsigned long long)st.arena_peak, (unsigned long long)arena);
  const std::string err = wahgpu_server_last_error(s);
  printf("batched: slots %u, %llu requests, %llu batches, max %llu columns, %llu memory waits, "
         "peak %llu of %llu bytes, %llu refused, last error \"%s\"\n",
         slots, (unsigned long long)st.requests, (unsigned long long)st.batches,
         (unsigned long long)st.max_batch_cols, (unsigned long long)st.memory_waits,
         (unsigned long long)st.arena_peak, (unsigned long long)arena,
         (unsigned long long)refused.load(), err.c_str());
  wahgpu_server_destroy(s);
  CHECK(fake_device_bytes() == 0, "device bytes left after destroy: %zu", fake_device_bytes());
}

// An arena that holds one large request: the others wait for memory, none
// is refused, all complete.
void run_memory_waits(uint32_t slots) {
  wahgpu_server_config cfg{};
  cfg.slots = slots;
  cfg.max_batch_cols = 8;
  const uint64_t n = 50000;
  cfg.arena_bytes = 50000 * 40 + 8192 + 4096;  // io (~32 B/value) + fake scratch (4 B/value)
  wahgpu_server* s = nullptr;
  CHECK(wahgpu_server_create(&s, &cfg) == WAHGPU_OK, "server_create");
  if (!s) return;
  std::minstd_rand rng(5);
  std::vector<std::unique_ptr<Job>> jobs;
  for (int i = 0; i < 6; ++i) {
    auto j = std::make_unique<Job>();
    j->values.resize(n);
    for (auto& v : j->values) v = rng() % 5000;
    j->rid_base = (uint32_t)i;
    j->keys.assign(n, 0);
    j->offsets.assign(n, 0);
    j->lengths.assign(n, 0);
    j->words.assign(2 * n, 0);
    j->res = wahgpu_result{j->keys.data(), j->offsets.data(), j->lengths.data(), n,
                           j->words.data(), 2 * n, {}};
    CHECK(wahgpu_server_submit(s, j->values.data(), n, j->rid_base, &j->res, &j->id) ==
              WAHGPU_OK,
          "submit");
    jobs.push_back(std::move(j));
  }
  for (auto& j : jobs) check_job(*j, wahgpu_server_wait(s, j->id), "memory waits");
  wahgpu_server_stats st{};
  wahgpu_server_get_stats(s, &st);
  CHECK(st.memory_waits >= 1 && st.batches == jobs.size(),
        "memory waits %llu, batches %llu", (unsigned long long)st.memory_waits,
        (unsigned long long)st.batches);
  // a column the empty arena cannot hold is refused at once
  std::vector<uint32_t> big(4 * n, 1);
  wahgpu_result r{};
  uint64_t id = 0;
  CHECK(wahgpu_server_submit(s, big.data(), big.size(), 0, &r, &id) == WAHGPU_ERANGE,
        "oversized column not refused");
  printf("memory waits: slots %u, %llu batches, %llu waits\n", slots,
         (unsigned long long)st.batches, (unsigned long long)st.memory_waits);
  wahgpu_server_destroy(s);
}

// Requests still queued when the server is destroyed are served first.
void run_destroy_with_queue(uint32_t plain) {
  wahgpu_server_config cfg{};
  cfg.slots = 2;
  cfg.arena_bytes = 8 << 20;
  cfg.max_batch_cols = 16;
  cfg.plain = plain;
  wahgpu_server* s = nullptr;
  CHECK(wahgpu_server_create(&s, &cfg) == WAHGPU_OK, "server_create");
  if (!s) return;
  std::minstd_rand rng(9 + plain);
  std::vector<std::unique_ptr<Job>> jobs;
  for (int i = 0; i < 200; ++i) {
    auto j = make_job(rng);
    if (j->values.size() > 10000) j->values.resize(10000);
    j->fault = false;
    if (!j->values.empty() && j->values[0] == kFakeFaultMarker) j->values[0] = 1;
    CHECK(wahgpu_server_submit(s, j->values.data(), j->values.size(), j->rid_base, &j->res,
                               &j->id) == WAHGPU_OK,
          "submit");
    jobs.push_back(std::move(j));
  }
  wahgpu_server_destroy(s);
  for (auto& j : jobs) check_job(*j, WAHGPU_OK, plain ? "destroy (plain)" : "destroy");
  printf("destroy with queue (plain %u): %zu requests served\n", plain, jobs.size());
}

}  // namespace

int main(int argc, char** argv) {
  const uint32_t per_thread = argc > 1 ? (uint32_t)atoi(argv[1]) : 150;
  for (uint32_t slots : {1u, 2u, 4u})
    run_batched(slots, 3u << 20, 6, per_thread, slots);
  run_memory_waits(2);
  run_memory_waits(4);
  run_destroy_with_queue(0);
  run_destroy_with_queue(1);
  if (g_failures) {
    fprintf(stderr, "%d failures\n", g_failures.load());
    return 1;
  }
  printf("ok\n");
  return 0;
}
