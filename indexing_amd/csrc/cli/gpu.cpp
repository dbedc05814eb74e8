// gpu -- the reference's actor-driven GPU indexer (src/gpu.cpp) on libwahgpu.
//
//   gpu [-f FILE] [-b BOUND] [-d DEVICE] [--batch-size B] [-p]
//       [-w WORK_GROUPS] [-j JOBS] [--binary]
//
// Three actors exchange the reference's messages (src/gpu.cpp:103-261):
//   indexer  (init, worker, values, jobs, wg_num, wg_size, print, bound), (index)
//            -> worker (config, values)
//            <- worker (config, keys, index, offsets)
//            -> merger (add, keycnt, length, keys, offsets, index)
//   merger   (init, print, bound), (add, ...), (quit): prints "Time: 'N' ms"
//   worker   one GPU context; config = [n_batch, n_wg, (len_i, 2*len_i)...] in,
//            [..., (keycnt_i, length_i)...] out (include/kernel_wah_bitindex.cl:50-54,88-89)
// CAF is not available, so each actor is a thread draining a FIFO mailbox of
// std::variant messages.  Differences from the reference, all deliberate:
//   - slices are indexed with their global row ids and the merger stitches
//     them exactly (wahgpu_merger), so the result equals the whole-column
//     build for any batch size; the reference restarts rids in every work
//     group and concatenates, which is only right for one slice;
//   - a "work group" slice is built by the whole GPU, so wg_size is
//     min(batch-size, n), not capped at the device's work-group size;
//   - -b is --bound (the reference registers -b twice, src/gpu.cpp:278,280);
//   - -d defaults to the first device; no kernel source file is read.
#include <stdio.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <variant>
#include <vector>

#include "cli_common.hpp"

namespace {

using uvec = std::vector<uint32_t>;

template <class Msg>
class Mailbox {
 public:
  void push(Msg m) {
    {
      std::lock_guard<std::mutex> g(mu_);
      q_.push_back(std::move(m));
    }
    cv_.notify_one();
  }
  Msg pop() {
    std::unique_lock<std::mutex> g(mu_);
    cv_.wait(g, [&] { return !q_.empty(); });
    Msg m = std::move(q_.front());
    q_.pop_front();
    return m;
  }

 private:
  std::mutex mu_;
  std::condition_variable cv_;
  std::deque<Msg> q_;
};

// ---- messages --------------------------------------------------------------
struct InitMerger { bool print; uint32_t bound; };
struct Add { uint32_t keycnt, length; uvec keys, offsets, index; };
struct Quit {};
using MergerMsg = std::variant<InitMerger, Add, Quit>;

struct WorkerReq { uvec config, values; };
struct WorkerReply { uvec config, keys, index, offsets; };
using WorkerMsg = std::variant<WorkerReq, Quit>;

struct InitIndexer {
  uvec input;
  uint32_t jobs, wg_num, wg_size;
  bool print;
  uint32_t bound;
};
struct IndexMsg {};
using IndexerMsg = std::variant<InitIndexer, IndexMsg, WorkerReply>;

// ---- merger (src/gpu.cpp:103-165) -------------------------------------------
void merger(Mailbox<MergerMsg>* box) {
  bool print = false;
  uint32_t bound = 0;
  auto start = std::chrono::steady_clock::now();
  wahgpu_merger* m = nullptr;
  wahgpu_merger_create(&m);
  for (;;) {
    MergerMsg msg = box->pop();
    if (auto* i = std::get_if<InitMerger>(&msg)) {
      start = std::chrono::steady_clock::now();
      print = i->print;
      bound = i->bound;
    } else if (auto* a = std::get_if<Add>(&msg)) {
      if (wahgpu_merger_add(m, a->keycnt, a->length, a->keys.data(), a->offsets.data(),
                            a->index.data()) != WAHGPU_OK)
        fprintf(stderr, "merger: malformed partial index\n");
    } else {
      auto stop = std::chrono::steady_clock::now();
      if (print) {
        for (uint64_t key = 0; key <= bound; ++key) {
          const uint32_t* w;
          uint64_t len;
          wahgpu_merger_key_words(m, (uint32_t)key, &w, &len);
          printf("%u\t%s\n", (uint32_t)key, len ? cli::decoded(w, len).c_str() : "");
        }
      }
      printf("Time: '%lld' ms\n",
             (long long)std::chrono::duration_cast<std::chrono::milliseconds>(stop - start).count());
      fflush(stdout);
      wahgpu_merger_destroy(m);
      return;
    }
  }
}

// ---- worker: the kernel actor (src/gpu.cpp:356-367) --------------------------
void worker(Mailbox<WorkerMsg>* box, Mailbox<IndexerMsg>* reply_to, int device,
            uint32_t wg_size) {
  wahgpu_ctx* ctx = nullptr;
  if (wahgpu_create(&ctx, device, 0) != WAHGPU_OK) {
    fprintf(stderr, "worker: cannot open device %d\n", device);
    exit(1);
  }
  uint64_t row = 0;  // global row of the next batch (requests arrive in order)
  for (;;) {
    WorkerMsg msg = box->pop();
    if (std::holds_alternative<Quit>(msg)) break;
    WorkerReq& r = std::get<WorkerReq>(msg);
    WorkerReply out;
    out.config = r.config;
    const uint32_t batch = r.config[0], wg_num = r.config[1];
    out.keys.assign(batch, 0);
    out.offsets.assign(batch, 0);
    out.index.assign(2 * (uint64_t)batch, 0);
    for (uint32_t i = 0; i < wg_num; ++i) {
      const uint64_t from = (uint64_t)i * wg_size;
      const uint32_t len = r.config[2 * i + 2];
      std::vector<uint32_t> lengths(len ? len : 1);
      wahgpu_counts c{};
      int rc = wahgpu_build(ctx, r.values.data() + from, len, (uint32_t)(row + from),
                            out.keys.data() + from, out.offsets.data() + from, lengths.data(),
                            len, out.index.data() + 2 * from, 2 * (uint64_t)len, &c, nullptr);
      if (rc != WAHGPU_OK) {
        fprintf(stderr, "worker: build failed (%d): %s\n", rc, wahgpu_last_error(ctx));
        exit(1);
      }
      out.config[2 * i + 2] = (uint32_t)c.keycnt;
      out.config[2 * i + 3] = (uint32_t)c.length;
    }
    row += batch;
    reply_to->push(std::move(out));
  }
  wahgpu_destroy(ctx);
}

// ---- indexer (src/gpu.cpp:179-261) -------------------------------------------
void indexer(Mailbox<IndexerMsg>* box, Mailbox<WorkerMsg>* worker_box,
             Mailbox<MergerMsg>* merger_box) {
  uvec input;
  uint64_t pos = 0;  // consumed prefix of input
  uint64_t remaining = 0;
  uint32_t in_progress = 0, jobs = 1, wg_num = 1, wg_size = 1;
  for (;;) {
    IndexerMsg msg = box->pop();
    if (auto* i = std::get_if<InitIndexer>(&msg)) {
      remaining = i->input.size();
      input = std::move(i->input);
      jobs = i->jobs;
      wg_num = i->wg_num;
      wg_size = i->wg_size;
      merger_box->push(InitMerger{i->print, i->bound});
      if (remaining == 0) {  // nothing to index
        merger_box->push(Quit{});
        worker_box->push(Quit{});
        return;
      }
    } else if (std::holds_alternative<IndexMsg>(msg)) {
      if (in_progress >= jobs || pos == input.size()) continue;
      uint64_t left = input.size() - pos;
      uint32_t batch = (uint32_t)std::min<uint64_t>(left, (uint64_t)wg_size * wg_num);
      uvec values(input.begin() + pos, input.begin() + pos + batch);
      uint32_t full = batch / wg_size, partial = batch - full * wg_size;
      uvec config(2 + 2 * (uint64_t)full);
      config[0] = batch;
      config[1] = full;
      for (uint32_t k = 0; k < full; ++k) {
        config[2 * k + 2] = wg_size;
        config[2 * k + 3] = wg_size * 2;
      }
      if (partial > 0) {
        config[1] += 1;
        config.push_back(partial);
        config.push_back(partial * 2);
      }
      worker_box->push(WorkerReq{std::move(config), std::move(values)});
      pos += batch;
      in_progress += 1;
      box->push(IndexMsg{});
    } else {
      WorkerReply& r = std::get<WorkerReply>(msg);
      uint32_t processed = r.config[0];
      const uint32_t n_wg = r.config[1];
      in_progress -= 1;
      remaining -= processed;
      for (uint32_t i = 0; i < n_wg; ++i) {
        uint32_t keycnt = r.config[2 * i + 2], length = r.config[2 * i + 3];
        uint32_t buffer_length = std::min(processed, wg_size);
        uint64_t from = (uint64_t)i * wg_size, to = from + buffer_length;
        Add a;
        a.keycnt = keycnt;
        a.length = length;
        a.keys.assign(r.keys.begin() + from, r.keys.begin() + to);
        a.offsets.assign(r.offsets.begin() + from, r.offsets.begin() + to);
        a.index.assign(r.index.begin() + 2 * from, r.index.begin() + 2 * to);
        merger_box->push(std::move(a));
        processed -= buffer_length;
      }
      if (remaining > 0) {
        box->push(IndexMsg{});
      } else {
        merger_box->push(Quit{});
        worker_box->push(Quit{});
        return;
      }
    }
  }
}

}  // namespace

int main(int argc, char** argv) {
  std::string file, bound_s = "0", device = "", batch_s = "1023", jobs_s = "1", wg_s = "1";
  bool print = false, binary = false;
  std::vector<cli::Opt> opts = {
      {"-f", "--data-file", false, &file, nullptr, "file with test data (one value per line)"},
      {"-b", "--bound", false, &bound_s, nullptr, "maximum value (0 will scan values)"},
      {"-d", "--device", false, &device, nullptr, "device for computation (default: first)"},
      {nullptr, "--batch-size", false, &batch_s, nullptr, "values indexed in one batch (1023)"},
      {"-p", "--print", true, nullptr, &print, "print resulting bitmap index"},
      {"-w", "--work-groups", false, &wg_s, nullptr, "Use work-groups"},
      {"-j", "--jobs", false, &jobs_s, nullptr, "jobs sent to GPU (1; 0 = compute units)"},
      {nullptr, "--binary", true, nullptr, &binary, "value file is raw little-endian uint32"},
  };
  int pr = cli::parse(argc, argv, opts);
  if (pr == 1) return 0;
  if (pr == 2) return 2;
  uvec values;
  if (file.empty()) {
    values = {10, 7, 22, 6, 7, 1, 9, 42, 2, 5, 13, 3, 2, 1, 0, 1, 18, 18, 3, 13,
              5, 9, 0, 3, 2, 19, 5, 23, 22, 10, 6, 22};
  } else {
    printf("Reading data from '%s' ... ", file.c_str());
    fflush(stdout);
    if (!cli::load_values(file, binary, &values)) {
      fprintf(stderr, "gpu: cannot read '%s'\n", file.c_str());
      return 1;
    }
  }
  printf("'%zu' values.\n", values.size());
  uint32_t bound = cli::to_u32(bound_s);
  if (bound == 0)
    for (uint32_t v : values) bound = v > bound ? v : bound;
  printf("Maximum value is '%u'.\n", bound);
  int dev = cli::pick_device(device, false);
  if (dev < 0) {
    fprintf(stderr, "Device %s not found.\n", device.c_str());
    return 1;
  }
  char name[256] = {0};
  wahgpu_device_name(dev, name, sizeof(name));
  printf("Using device named '%s'.\n", name);
  fflush(stdout);
  uint32_t jobs = cli::to_u32(jobs_s);
  if (jobs == 0) {
    int cus = wahgpu_device_compute_units(dev);
    jobs = cus > 0 ? (uint32_t)cus : 1u;
  }
  uint32_t batch_size = std::min<uint64_t>(cli::to_u32(batch_s), values.size());
  uint32_t wg_size = batch_size ? batch_size : 1;
  uint32_t wg_num = std::max<uint32_t>(1, cli::to_u32(wg_s));

  Mailbox<MergerMsg> merger_box;
  Mailbox<WorkerMsg> worker_box;
  Mailbox<IndexerMsg> indexer_box;
  std::thread tm(merger, &merger_box);
  std::thread tw(worker, &worker_box, &indexer_box, dev, wg_size);
  std::thread ti(indexer, &indexer_box, &worker_box, &merger_box);
  indexer_box.push(InitIndexer{std::move(values), jobs, wg_num, wg_size, print, bound});
  indexer_box.push(IndexMsg{});
  ti.join();
  tw.join();
  tm.join();
  return 0;
}
