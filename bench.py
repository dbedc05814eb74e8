"""Benchmark: the single-column WAH index build at 2^28 uint32 values.

Workload: 2^28 values i.i.d. uniform in [0, 65535], seed 1 (the single-column
GPU drop-in configuration), resident in HBM before timing starts.  A step is
one fused build (wahgpu_build_device: histogram, radix passes, tail, lengths)
over the whole column.  W warm-up steps, then K timed steps between
synchronizations; value = n * K / elapsed in Gvalues/s.

Beside it:
  roofline      the dominant kernel (the radix passes, k_onesweep) timed with
                HIP events on the build's stream over the timed steps;
                achieved = its algorithmic bytes per launch / its mean launch
                time, with the canonical pipeline's bytes: (value, rid) pairs
                of 8 bytes, pass 1 reads 4n and writes 8n, later passes read
                and write 8n each.  moved_* is the same for the bytes this
                build moves: keys below 2^16 travel as 2-byte words (kb = 2,
                else 4), so pass 1 writes (kb + 4)n and later passes read and
                write (kb + 4)n each
  build_roofline the same for the whole build: B_alg = 16nP + 8n + 4L + 12K
                (histogram 4n, passes, tail 8n read, words 4L, keys, offsets
                and lengths 12K); moved_bytes = (2kb + 8)nP + 8n + 4L + 12K
                with the tail reading (kb + 4)n (run mode: the k run items
                through the cut, the passes and the tail)
  parity        the index of the last timed step, digested, against the
                oracle's digest for the same column (tests/golden/fullsize.json)
  kernels       per kernel group: mean ms, algorithmic and moved bytes, GB/s
  ingest        the column as a text value file: write time, one serial parse
                and the parallel parse (host threads), both checked equal
  cpu_baseline  the oracle (a C restatement of the reference's serial `cpu`
                program) on the whole 2^28 column, run in a child process
                pinned to one core while the GPU legs run, with per-stage
                times, the CPU model and core count, and the literal bubble
                sort checked against the stable sort on a 2^16 sample; then,
                labelled separately ("threaded"), the same build with its
                sort on up to 16 threads, once the GPU legs are done
  batched       the mixed 1024-column workload (lengths log-uniform in
                [2^12, 2^24], cardinality log-uniform in [2, 2^16], half
                uniform half Zipf(1.0), seed 7), resident in HBM: one batched
                build (graph replay) against a loop of per-column
                wahgpu_build_device calls over the same columns
  serve         requests drawn from the same distributions (seed 11), page-
                locked host memory in and out: the continuous-batching server against the
                plain per-column drop-in (wahgpu_build per request), first as
                a burst (saturation throughput), then as an open-loop Poisson
                stream at half the plain path's saturation rate (p50 / p99)
Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "indexed Gvalues/s + HBM-roofline % (2^28 u32); p50/p99 serve latency"
HBM_PEAK_GBS = 8000.0
N_VALUES = 1 << 28
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_onesweep_uniform_2p28.txt")


def read_traffic():
    """HBM bytes per k_onesweep launch measured with rocprofv3 counters
    (FETCH_SIZE + WRITE_SIZE, calibrated on a copy kernel), if recorded."""
    try:
        for line in open(TRAFFIC_FILE):
            if line.startswith("bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def byte_model(n, k, L, K, P, bound, mode):
    """Bytes of one build of a column: n rows, k (value, chunk) entries, L WAH
    words, K keys, P active radix passes, maximum value `bound`, sort mode
    (0 pairs, 1 runs).

    alg_pass / b_alg are the canonical pipeline's bytes: 8-byte (value, rid)
    pairs, radix pass 1 reads 4n values and writes 8n, each later pass reads
    and writes 8n; the whole build is B_alg = 16nP + 8n + 4L + 12K (histogram
    4n, passes, the tail's 8n read, words 4L, keys, offsets and lengths 12K).

    moved_pass / b_moved are the bytes the fused build moves: keys below 2^16
    travel between passes as 2-byte words (kb = 2, else 4); a run-mode column
    sorts one (key, rid, literal bits) item per run, at least k of them, cut
    from a second read of the values.
    """
    alg_pass = [12.0 * n] + [16.0 * n] * (P - 1) if P else []
    kb = 2.0 if bound < 65536 and P > 0 else 4.0
    if mode == 0:
        moved_pass = [(8.0 + kb) * n] + [(2.0 * kb + 8.0) * n] * (P - 1) if P else []
        b_moved = (2.0 * kb + 8.0) * n * P + 8.0 * n + 4.0 * L + 12.0 * K
    else:
        moved_pass = [(20.0 + kb) * k] + [(2.0 * kb + 16.0) * k] * (P - 1) if P else []
        # histogram 4n, run cut 4n + 12k, passes, tail (kb + 8)k, outputs
        b_moved = 8.0 * n + 12.0 * k + sum(moved_pass) + (kb + 8.0) * k + 4.0 * L + 12.0 * K
    b_alg = 16.0 * n * P + 8.0 * n + 4.0 * L + 12.0 * K
    return {"alg_pass": alg_pass, "moved_pass": moved_pass, "b_alg": b_alg, "b_moved": b_moved,
            "kb": kb}


def progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_info():
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, len(os.sched_getaffinity(0)), os.cpu_count()


def cpu_baseline_worker(n, core):
    """Runs in a child process pinned to one core: the oracle (a C
    restatement of the reference's serial `cpu` program, src/cpu.cpp:176-272)
    on the whole 2^28 column with the stable sort, per-stage times; then the
    literal O(n^2) bubble sort of the reference (src/cpu.cpp:180-197) against
    the stable sort on the first 2^16 values of the 2^20 uniform seed-42
    column, compared by index digest.  Prints one JSON object."""
    os.sched_setaffinity(0, {core})
    import indexing_amd as ia
    from indexing_amd.digest import index_digest
    from oracle import oracle
    values = ia.generate(n, "uniform", seed=1, lo=0, hi=65535)
    t0 = time.perf_counter()
    r = oracle.build(values, timed=True)
    dt = time.perf_counter() - t0
    full = {"n": n, "seconds": round(dt, 3), "gvalues_s": n / dt / 1e9,
            "stage_s": {k: round(v, 4) for k, v in r["stage_s"].items()},
            "k": r["k"], "length": r["length"], "keycnt": r["keycnt"],
            "index": index_digest(r["keys"], r["offsets"], r["lengths"], r["words"])}
    del r, values
    sub = ia.generate(1 << 20, "uniform", seed=42, lo=0, hi=65535)[:1 << 16]
    t0 = time.perf_counter()
    b = oracle.build(sub, sort="bubble")
    tb = time.perf_counter() - t0
    s = oracle.build(sub, sort="packed")  # (value, rid) words sorted: the stable order
    dig = [index_digest(x["keys"], x["offsets"], x["lengths"], x["words"]) for x in (b, s)]
    bubble = {"sample": "first 2^16 values of the 2^20 uniform [0,65535] seed-42 column",
              "bubble_seconds": round(tb, 3), "bubble_digest": dig[0], "stable_digest": dig[1],
              "equal": dig[0] == dig[1]}
    print(json.dumps({"full": full, "bubble_vs_stable": bubble}))


def start_cpu_baseline(n):
    """Starts the CPU baseline in a child process on the last core this
    process may use (the GPU legs run meanwhile on the others)."""
    import subprocess
    core = max(os.sched_getaffinity(0))
    return core, subprocess.Popen([sys.executable, os.path.abspath(__file__),
                                   "--cpu-baseline-worker", str(n), str(core)],
                                  stdout=subprocess.PIPE, text=True)


def finish_cpu_baseline(proc, core, gold):
    out, _ = proc.communicate()
    if proc.returncode != 0:
        raise RuntimeError("cpu baseline worker failed")
    r = json.loads(out.strip().splitlines()[-1])
    full = r["full"]
    model, allowed, total = cpu_info()
    res = {"value": round(full["gvalues_s"], 5), "unit": "Gvalues/s", "cores": 1,
           "kind": "port",
           "sample": f"the whole 2^{full['n'].bit_length() - 1}-value column (uniform [0,65535], "
                     f"seed 1), oracle build with the stable sort on one pinned core "
                     f"(core {core}), {full['seconds']} s, text parsing excluded",
           "cpu_model": model, "nproc": allowed, "host_cpus": total,
           "stage_s": full["stage_s"],
           "bubble_vs_stable": r["bubble_vs_stable"]}
    if gold is not None:
        res["matches_gpu_digest"] = full["index"] == gold
    return res


def threaded_cpu_baseline(n, gold):
    """The oracle build of the same column with its sort on up to 16 threads
    (slices sorted, then merged by rank ranges: the serial build's output),
    the other stages serial; run after the GPU legs, on the cores this
    process may use."""
    import indexing_amd as ia
    from indexing_amd.digest import index_digest
    from oracle import oracle
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    values = ia.generate(n, "uniform", seed=1, lo=0, hi=65535)
    t0 = time.perf_counter()
    r = oracle.build(values, timed=True, threads=threads)
    dt = time.perf_counter() - t0
    res = {"value": round(n / dt / 1e9, 5), "unit": "Gvalues/s", "cores": threads,
           "kind": "port",
           "sample": f"the whole 2^{n.bit_length() - 1}-value column, oracle build with the "
                     f"sort on {threads} threads, {dt:.3f} s, text parsing excluded",
           "stage_s": {k: round(v, 4) for k, v in r["stage_s"].items()}}
    if gold is not None:
        res["matches_gpu_digest"] = index_digest(r["keys"], r["offsets"], r["lengths"],
                                                 r["words"]) == gold
    return res


def bench_ingest(values, threads):
    """Text value file of the column (the generator's format, one value per
    line) written and parsed back: one serial pass (`ifstream >>`
    semantics) and the parallel parse, checked equal to the column."""
    import tempfile

    import numpy as np

    import indexing_amd as ia
    fd, path = tempfile.mkstemp(suffix=".txt")
    os.close(fd)
    try:
        t0 = time.perf_counter()
        ia.write_values_text(path, values, threads=threads)
        tw = time.perf_counter() - t0
        size = os.path.getsize(path)
        res = {"file_bytes": size, "write_s": round(tw, 3), "threads": threads}
        for name, th in (("serial", 1), ("parallel", threads)):
            t0 = time.perf_counter()
            got = ia.read_values_text(path, threads=th)
            dt = time.perf_counter() - t0
            res[f"{name}_parse_s"] = round(dt, 3)
            res[f"{name}_GBs"] = round(size / dt / 1e9, 2)
            res[f"{name}_equal"] = bool(np.array_equal(got, values))
            del got
        res["speedup"] = round(res["serial_parse_s"] / res["parallel_parse_s"], 2)
        return res
    finally:
        os.remove(path)


def bench_batched(ncols, seed=7, reps=3):
    """C4: one batched build of ncols mixed columns vs per-column builds."""
    import numpy as np
    import torch

    import indexing_amd as ia
    from indexing_amd import workloads as W
    specs = W.column_specs(ncols, seed)
    progress(f"batched: generating {ncols} columns")
    t0 = time.perf_counter()
    buf, layout = W.make_columns(specs)
    gen_s = time.perf_counter() - t0
    progress(f"batched: {gen_s:.1f} s to generate, building")
    total = int(sum(n for _, n in layout))
    d_vals = torch.from_numpy(buf.view(np.int32)).to("cuda:0")
    del buf
    ctx = ia.Context(0)
    tuning = ctx.autotune()
    ctx.batch_reserve(d_vals.numel(), ncols)
    out = ctx.build_batch_device(d_vals, layout)          # captures the graph
    torch.cuda.synchronize()
    counts = ctx.read_counts_batch(out[4], ncols)
    assert all(c.fault == 0 for c in counts)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    best_b = 1e30
    for _ in range(reps):
        ev[0].record()
        ctx.build_batch_device(d_vals, layout, out=out)
        ev[1].record()
        torch.cuda.synchronize()
        best_b = min(best_b, ev[0].elapsed_time(ev[1]))
    # plain: one fused build per column on the same device data and buffers
    keys, offs, lens, words, _ = out
    cnt = torch.zeros(ncols * 10, dtype=torch.int32, device="cuda:0")
    best_p = 1e30
    for _ in range(reps):
        ev[0].record()
        for i, (o, n) in enumerate(layout):
            ctx.build_device(d_vals[o:o + n], out=(keys[o:o + n], offs[o:o + n], lens[o:o + n],
                                                   words[2 * o:2 * (o + n)],
                                                   cnt[10 * i:10 * i + 10]))
        ev[1].record()
        torch.cuda.synchronize()
        best_p = min(best_p, ev[0].elapsed_time(ev[1]))
    # same results both ways (counts of every column)
    pc = [ctx.read_counts(cnt[10 * i:10 * i + 10]) for i in range(ncols)]
    assert all((a.length, a.keycnt, a.k) == (b.length, b.keycnt, b.k) for a, b in zip(counts, pc))
    captures, replays = ctx.graph_stats()
    ctx.close()
    del d_vals, out
    torch.cuda.empty_cache()
    return {"workload": f"{ncols} columns, lengths log-uniform [2^12, 2^24], cardinality "
                        "log-uniform [2, 2^16], half uniform half Zipf(1.0), seed 7, resident",
            "columns": ncols, "values": total, "batched_ms": round(best_b, 3),
            "batched_gvalues_s": round(total / best_b / 1e6, 2),
            "per_column_ms": round(best_p, 3),
            "per_column_gvalues_s": round(total / best_p / 1e6, 2),
            "speedup": round(best_p / best_b, 2), "tile_shapes": tuning["shape"],
            "digit_bits": tuning["digit_bits"],
            "wide_digit_columns": sum(int(c.mode) == 2 for c in counts),
            "graph_captures": captures,
            "graph_replays": replays, "host_generation_s": round(gen_s, 1)}


def pcie_link(nbytes=1 << 30, reps=3):
    """Host-link rates with page-locked buffers: host-to-device, device-to-
    host, and both at once on two streams (GB/s, best of reps)."""
    import torch
    h_in = torch.empty(nbytes // 4, dtype=torch.int32).pin_memory()
    h_out = torch.empty(nbytes // 4, dtype=torch.int32).pin_memory()
    d_in = torch.empty(nbytes // 4, dtype=torch.int32, device="cuda:0")
    d_out = torch.empty(nbytes // 4, dtype=torch.int32, device="cuda:0")
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    res = {}
    for name in ("h2d", "d2h", "both"):
        best = 0.0
        for _ in range(reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            if name in ("h2d", "both"):
                with torch.cuda.stream(s1):
                    d_in.copy_(h_in, non_blocking=True)
            if name in ("d2h", "both"):
                with torch.cuda.stream(s2):
                    h_out.copy_(d_out, non_blocking=True)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            best = max(best, (2 if name == "both" else 1) * nbytes / dt / 1e9)
        res[name + "_GBs"] = round(best, 1)
    return res


class _ResultRing:
    """Result buffers carved in order from one page-locked block and freed
    in the same order (requests are retired in submission order)."""

    def __init__(self, words):
        import collections
        import threading

        import indexing_amd as ia
        self.buf = ia.pinned_empty(words)
        self.cap = words
        self.live = collections.deque()  # (start, end) in allocation order
        self.cv = threading.Condition()

    def alloc(self, size):
        with self.cv:
            while True:
                start = None
                if not self.live:
                    start = 0
                else:
                    tail, head = self.live[0][0], self.live[-1][1]
                    if head > tail:      # in use: [tail, head)
                        if head + size <= self.cap:
                            start = head
                        elif size <= tail:
                            start = 0
                    elif head + size <= tail:  # in use: [tail, cap) and [0, head)
                        start = head
                if start is not None:
                    self.live.append((start, start + size))
                    return self.buf[start:start + size]
                self.cv.wait()

    def free_oldest(self):
        with self.cv:
            self.live.popleft()
            self.cv.notify_all()


def serve_run(srv, arrivals, req_col, cols, layout, ring, pool, checksum, alone, record):
    """Submits request i (column req_col[i]) at arrivals[i] seconds (open
    loop) with its result buffers carved from `ring`; a retire thread waits
    for the requests in order, checksums each result on `pool` and frees ring
    space in order as the checksums end (and all of it before it sleeps, so a
    submitter waiting for ring space always gets it).  Results are compared
    with alone[c] (recorded first when `record`).  Returns the server's stats
    and the mismatches."""
    import collections
    import threading
    srv.reset_stats()
    pend = []
    cv = threading.Condition()
    state = {"done": False, "late": 0.0, "out_bytes": 0}
    sums = [None] * len(arrivals)

    def retire():
        fifo = collections.deque()

        def free_front(block):
            while fifo and (block or fifo[0][1].done()):
                i, f = fifo.popleft()
                sums[i] = f.result()
                ring.free_oldest()

        j = 0
        while True:
            with cv:
                ready = j < len(pend) or state["done"]
            if not ready:
                free_front(True)
            with cv:
                while j >= len(pend) and not state["done"]:
                    cv.wait()
                if j >= len(pend):
                    break
                p = pend[j]
            try:
                srv.wait(p)
            except Exception as e:  # a failed request is a mismatch, the run goes on
                state.setdefault("errors", []).append(str(e))
                fifo.append((j, pool.submit(lambda: None)))
                free_front(False)
                j += 1
                continue
            r = p.res.counts
            K, L = int(r.keycnt), int(r.length)
            state["out_bytes"] += 4 * L + 12 * K
            k, o, ln, w = p.bufs
            fifo.append((j, pool.submit(checksum, k[:K], o[:K], ln[:K], w[:L])))
            free_front(False)
            j += 1
        free_front(True)

    th = threading.Thread(target=retire)
    th.start()
    t0 = time.perf_counter()
    for c, at in zip(req_col, arrivals):
        delay = t0 + at - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        else:
            state["late"] = max(state["late"], -delay)
        n = layout[c][1]
        out = ring.alloc(5 * n + 4)
        bufs = (out[:n], out[n:2 * n], out[2 * n:3 * n], out[3 * n:5 * n])
        p = srv.submit(cols[c], out=bufs)
        with cv:
            pend.append(p)
            cv.notify()
    with cv:
        state["done"] = True
        cv.notify()
    th.join()
    bad = 0
    for i, c in enumerate(req_col[:len(arrivals)]):
        if record and alone[c] is None:
            alone[c] = sums[i]
        elif alone[c] != sums[i]:
            bad += 1
    st = srv.stats()
    return {"p50_ms": round(st["p50_us"] / 1e3, 3), "p99_ms": round(st["p99_us"] / 1e3, 3),
            "gvalues_s": round(st["values"] / st["busy_s"] / 1e9, 3) if st["busy_s"] else 0.0,
            "requests": int(st["requests"]), "batches": int(st["batches"]),
            "max_batch_cols": int(st["max_batch_cols"]),
            "memory_waits": int(st["memory_waits"]),
            "arena_peak_gib": round(st["arena_peak"] / 2**30, 2),
            "max_submit_lateness_ms": round(state["late"] * 1e3, 2),
            "out_bytes_per_value": round(state["out_bytes"] / max(1, st["values"]), 3),
            "mismatches": bad, "errors": state.get("errors", [])[:3]}


def bench_serve(nreq, pool_cols=512, seed=11, arena_gib=64, slots=3):
    """C5: continuous-batching server against the plain per-column path.

    Requests: nreq columns cycling over a pool of pool_cols distinct columns
    drawn from the C4 distributions (seed 11), page-locked host memory in and
    out.  Every served result's checksum is compared with the same column
    built alone (the plain path).  Runs: a burst (saturation throughput),
    open-loop Poisson arrivals at a fixed rate below the plain path's
    saturation, and at a rate above it."""
    import concurrent.futures as cf

    import numpy as np

    import indexing_amd as ia
    from indexing_amd import workloads as W
    specs = W.column_specs(pool_cols, seed)
    progress(f"serve: generating a pool of {pool_cols} columns")
    buf, layout = W.make_columns(specs, alloc=ia.pinned_empty)
    cols = [buf[o:o + n] for o, n in layout]
    req_col = [i % pool_cols for i in range(nreq)]
    total = int(sum(layout[c][1] for c in req_col))
    ring = _ResultRing(1 << 30)
    alone = [None] * pool_cols     # checksum of each column built alone (plain)
    pool = cf.ThreadPoolExecutor(12)

    def run(srv, arrivals, record):
        return serve_run(srv, arrivals, req_col, cols, layout, ring, pool, ia.index_checksum,
                         alone, record)

    res = {}
    burst = np.zeros(nreq)
    plain_cfg = dict(slots=1, arena_bytes=8 << 30, plain=True)
    served_cfg = dict(slots=slots, arena_bytes=arena_gib << 30, max_batch_cols=1024,
                      max_batch_values=1 << 24)
    progress(f"serve: plain burst of {nreq}")
    with ia.Server(**plain_cfg) as plain:
        res["plain_burst"] = run(plain, burst, True)
    progress(f"serve: served burst of {nreq}")
    with ia.Server(**served_cfg) as srv:
        res["served_burst"] = run(srv, burst, False)
    rng = np.random.default_rng(seed)
    mean_n = total / nreq
    for name, frac in (("fixed", 0.7), ("overload", 1.25)):
        rate = res["plain_burst"]["gvalues_s"] * 1e9 * frac / mean_n  # requests/s
        gaps = rng.exponential(1.0 / rate, nreq)
        arrivals = np.cumsum(gaps) - gaps[0]
        progress(f"serve: Poisson {name} at {rate:.0f} req/s")
        with ia.Server(**served_cfg) as srv:
            res[f"served_{name}"] = run(srv, arrivals, False)
        with ia.Server(**plain_cfg) as plain:
            res[f"plain_{name}"] = run(plain, arrivals, False)
        res[f"{name}_rate_req_s"] = round(rate, 1)
        res[f"{name}_offered_gvalues_s"] = round(rate * mean_n / 1e9, 3)
    pool.shutdown()
    progress("serve: host link")
    link = pcie_link()
    # the link's ceiling for this mix: 4 bytes in and out_bytes_per_value out
    # per value, each direction at its own rate, both directions at once
    outb = res["plain_burst"]["out_bytes_per_value"]
    # duplex: each direction at its own measured rate at once; shared: both
    # directions through one link of the single-direction rate
    ceil_dir = min(link["h2d_GBs"] / 4.0, link["d2h_GBs"] / outb)
    ceil_shared = max(link["h2d_GBs"], link["d2h_GBs"]) / (4.0 + outb)
    sb = res["served_burst"]["gvalues_s"]
    res.update({"requests": nreq, "values": total, "pool_columns": pool_cols,
                "link": link, "link_bytes_per_value": round(4.0 + outb, 3),
                "link_ceiling_duplex_gvalues_s": round(ceil_dir, 3),
                "link_ceiling_shared_gvalues_s": round(ceil_shared, 3),
                "served_burst_link_GBs": round(sb * (4.0 + outb), 1),
                "served_burst_of_duplex_ceiling": round(sb / ceil_dir, 3),
                "plain_burst_of_duplex_ceiling": round(res["plain_burst"]["gvalues_s"] / ceil_dir,
                                                       3),
                "workload": f"{nreq} requests cycling over {pool_cols} distinct columns drawn "
                            "from the C4 distributions (lengths log-uniform [2^12, 2^24], "
                            "cardinality log-uniform [2, 2^16], half uniform half Zipf(1.0)), "
                            f"seed 11, page-locked host memory in and out, one {arena_gib} GiB "
                            f"device arena shared by {slots} slots, admission by device footprint, "
                            "batches of at most 2^24 values; every result checked against the "
                            "same column built alone (the plain path)",
                "throughput_ratio_burst": round(res["served_burst"]["gvalues_s"] /
                                                res["plain_burst"]["gvalues_s"], 2)})
    return res


GOLD_FILE = os.path.join(ROOT, "tests", "golden", "fullsize.json")


def golden_record(column, n):
    """The oracle's counts and index digest for this configuration, when
    tests/golden/fullsize.json holds it (c1, c2, c3), else None."""
    try:
        gold = json.load(open(GOLD_FILE))
    except OSError:
        return None, None
    for name in ("c1", "c2", "c3"):
        p = gold.get(name, {}).get("params", {})
        dist = "uniform" if column == "uniform" else "zipf-runs"
        if p.get("n") == n and p.get("dist") == dist and p.get("seed") == 1:
            return name, gold[name]
    return None, None


def kernel_table(group_ms, calls, bm, n, k, L, K, mode):
    """Per kernel group: mean ms, algorithmic bytes (the canonical byte model:
    histogram 4n, passes 12n / 16n, tail 8n + 4L + 8K, lengths 4K, which sum
    to B_alg) and the bytes the fused build moves, each as GB/s."""
    kb = bm["kb"]
    names = ["histogram", "pass0", "pass1", "pass2", "pass3", "tail", "lengths"]
    pass_ms = [group_ms[1 + p] / calls for p in range(4)]
    P = len(bm["alg_pass"])
    active = sorted(sorted(range(4), key=lambda p: -pass_ms[p])[:P])
    alg = {"histogram": 4.0 * n, "tail": 8.0 * n + 4.0 * L + 8.0 * K, "lengths": 4.0 * K}
    moved = {"histogram": 4.0 * n if mode == 0 else 8.0 * n + 12.0 * k,
             "tail": ((kb + 4.0) * n if mode == 0 else (kb + 8.0) * k) + 4.0 * L + 8.0 * K,
             "lengths": 4.0 * K}
    for i, p in enumerate(active):
        alg[f"pass{p}"] = bm["alg_pass"][i]
        moved[f"pass{p}"] = bm["moved_pass"][i]
    out = {}
    for name, g in zip(names, group_ms):
        ms = g / calls
        if name.startswith("pass") and name not in alg:
            continue  # an inactive pass (its digit is constant): returns at once
        a, m = alg.get(name, 0.0), moved.get(name, 0.0)
        out[name] = {"ms": round(ms, 4), "alg_bytes": a, "alg_GBs": round(a / ms / 1e6, 1),
                     "moved_bytes": m, "moved_GBs": round(m / ms / 1e6, 1)}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=N_VALUES, help="column length (default 2^28)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default tile shapes (profiled runs: only the measured builds)")
    ap.add_argument("--column", choices=["uniform", "zipf-runs"], default="uniform",
                    help="uniform [0,65535] seed 1 (default), or Zipf(1.2) over 2^20 values in "
                         "runs of mean length 4096, seed 1")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batch-cols", type=int, default=1024,
                    help="columns of the batched workload (0 skips it)")
    ap.add_argument("--serve-requests", type=int, default=20000,
                    help="requests of the serving workload (0 skips it)")
    ap.add_argument("--serve-slots", type=int, default=3, help="server slots (batches in flight)")
    ap.add_argument("--no-ingest", dest="ingest", action="store_false",
                    help="skip the text value-file write / parse timing")
    ap.add_argument("--cpu-baseline-worker", nargs=2, type=int, metavar=("N", "CORE"),
                    help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_baseline_worker:
        cpu_baseline_worker(*args.cpu_baseline_worker)
        return
    if args.gpus != 1:
        print(f"bench.py: runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    # the CPU baseline runs meanwhile in a child pinned to one core (the
    # reference's serial program on the same 2^28 column)
    cpu = None
    if not args.no_cpu_baseline and args.column == "uniform":
        cpu = start_cpu_baseline(args.n)
        progress(f"cpu baseline started on core {cpu[0]}")

    import numpy as np
    import torch

    import indexing_amd as ia
    from indexing_amd.digest import index_digest

    n = args.n
    if args.column == "uniform":
        host = ia.generate(n, "uniform", seed=1, lo=0, hi=65535)
        column = "uniform [0,65535], seed 1"
    else:
        host = ia.generate(n, "zipf-runs", seed=1, cardinality=1 << 20, zipf_s=1.2,
                           run_mean=4096.0)
        column = "Zipf(1.2) over 2^20 values in runs of mean length 4096, seed 1"
    d_values = torch.from_numpy(host.view(np.int32)).to("cuda:0")
    ingest = None
    if args.ingest:
        progress("text ingest")
        # leave the CPU baseline's core alone
        ingest = bench_ingest(host, max(1, min(16, len(os.sched_getaffinity(0)) -
                                               (1 if cpu else 0))))
    del host
    ctx = ia.Context(0)
    # tile shape per length bucket
    tuning = ctx.tuning() if args.no_autotune else ctx.autotune()
    out = ctx.build_device(d_values)                   # allocates outputs + arena
    for _ in range(args.warmup):
        ctx.build_device(d_values, out=out)
    # every timed step writes its own counts, so a fault in any step is seen
    step_counts = [torch.zeros(out[4].numel(), dtype=torch.int32, device="cuda:0")
                   for _ in range(args.steps)]
    torch.cuda.synchronize()

    ctx.timing_begin(args.steps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        ctx.build_device(d_values, out=out[:4] + (step_counts[i],))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    group_ms, calls = ctx.timing_end()
    all_counts = [ctx.read_counts(c) for c in step_counts]
    counts = all_counts[-1]
    rank_mode = ctx.rank_mode
    assert all(c.fault == 0 for c in all_counts), "device reported a fault in a timed step"
    assert len({(c.k, c.length, c.keycnt) for c in all_counts}) == 1, "steps disagree"
    assert calls == args.steps

    ms_per_step = elapsed * 1e3 / args.steps
    value = n * args.steps / elapsed / 1e9

    # parity of the timed output: the index of the last timed step against the
    # oracle's digest for this configuration
    L, K, P = int(counts.length), int(counts.keycnt), int(counts.passes)
    progress("digesting the output")
    keys, offs, lens, words = (x.cpu().numpy().view(np.uint32) for x in out[:4])
    got = index_digest(keys[:K], offs[:K], lens[:K], words[:L])
    del keys, offs, lens, words
    gname, grec = golden_record(args.column, n)
    parity = {"index_digest": got}
    if grec is not None:
        parity.update({"reference": f"oracle digest, tests/golden/fullsize.json {gname}",
                       "equal": got == grec["index"] and
                       (int(counts.k), L, K) == (grec["k"], grec["length"], grec["keycnt"])})
    else:
        parity.update({"reference": "none recorded for this size", "equal": None})

    # sizes for the byte model
    pass_ms = [group_ms[1 + p] / calls for p in range(4)]
    # the P active passes are the P slowest (inactive ones exit at once)
    active = sorted(sorted(range(4), key=lambda p: -pass_ms[p])[:P])
    bm = byte_model(n, int(counts.k), L, K, P, int(counts.bound), int(counts.mode))
    sort_ms = sum(pass_ms[p] for p in active)
    per_launch_alg = sum(bm["alg_pass"]) / P if P else 0.0
    per_launch_bytes = sum(bm["moved_pass"]) / P if P else 0.0
    per_launch_ms = sort_ms / P if P else 0.0
    achieved = per_launch_alg / (per_launch_ms * 1e-3) / 1e9 if P else 0.0
    moved = per_launch_bytes / (per_launch_ms * 1e-3) / 1e9 if P else 0.0
    traffic = read_traffic() if args.column == "uniform" and n == N_VALUES else None
    b_alg, b_moved = bm["b_alg"], bm["b_moved"]
    build_ms_events = sum(group_ms) / calls

    result = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Gvalues/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (seeded splitmix64 column, generated on the host)",
        "parity": parity,
        "config": {"workload": f"single-column WAH build, n=2^{n.bit_length() - 1} uint32 "
                               f"{column}, resident in HBM",
                   "n": n, "k": int(counts.k), "L": L, "K": K, "radix_passes": P,
                   "sort_mode": {0: "pairs", 1: "runs", 2: "pairs, wide digits"}[int(counts.mode)],
                   "radix_ranking": "lds-fetch-add" if rank_mode else "ballot-match",
                   "tuning": tuning},
        "roofline": {"bound": "hbm", "kernel": "k_onesweep (radix pass, mean of active launches)",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "alg_bytes_per_launch": per_launch_alg,
                     "mean_launch_ms": round(per_launch_ms, 4),
                     "moved_bytes_per_launch": per_launch_bytes,
                     "moved_achieved": round(moved, 1),
                     "moved_frac": round(moved / HBM_PEAK_GBS, 4)},
        "build_roofline": {"alg_bytes": b_alg, "ms": round(build_ms_events, 4),
                           "achieved": round(b_alg / (build_ms_events * 1e-3) / 1e9, 1),
                           "frac": round(b_alg / (build_ms_events * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                         4),
                           "moved_bytes": b_moved,
                           "moved_frac": round(b_moved / (build_ms_events * 1e-3) / 1e9
                                               / HBM_PEAK_GBS, 4)},
        "kernels": kernel_table(group_ms, calls, bm, n, int(counts.k), L, K, int(counts.mode)),
    }
    if ingest:
        result["ingest"] = ingest
    ctx.close()
    del d_values, out, step_counts
    torch.cuda.empty_cache()
    if args.batch_cols:
        result["batched"] = bench_batched(args.batch_cols)
    if args.serve_requests:
        result["serve"] = bench_serve(args.serve_requests, slots=args.serve_slots)
    if cpu is not None:
        progress("waiting for the cpu baseline")
        result["cpu_baseline"] = finish_cpu_baseline(cpu[1], cpu[0],
                                                     grec["index"] if grec else None)
        progress("threaded cpu baseline")
        result["cpu_baseline"]["threaded"] = threaded_cpu_baseline(
            args.n, grec["index"] if grec else None)
    print(json.dumps(result))
    if parity["equal"] is False:
        print("bench.py: the timed output differs from the oracle's", file=sys.stderr)
        sys.exit(1)


if __name__ == "__main__":
    main()
