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
  cpu_baseline  the oracle (a C restatement of the reference's serial `cpu`
                program, one thread) on a 2^26-value sample of the same
                distribution, timed on this host
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
CPU_SAMPLE = 1 << 26
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

    moved_pass / b_moved are the bytes this build moves: keys below 2^16
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


def cpu_baseline(sample_n):
    import numpy as np

    import indexing_amd as ia
    from oracle import oracle
    values = ia.generate(sample_n, "uniform", seed=1, lo=0, hi=65535)
    t0 = time.perf_counter()
    r = oracle.build(values)
    dt = time.perf_counter() - t0
    assert r["keycnt"] == np.unique(values).size
    return {"value": sample_n / dt / 1e9, "unit": "Gvalues/s", "cores": 1, "kind": "port",
            "sample": f"first 2^{sample_n.bit_length() - 1} values of the same column "
                      f"(uniform [0,65535], seed 1), oracle build timed single-threaded, "
                      f"{dt:.2f} s"}


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
            "graph_captures": captures,
            "graph_replays": replays, "host_generation_s": round(gen_s, 1)}


POISSON_REPS = 3  # passes over the Poisson arrivals, latencies pooled for p50/p99


def bench_serve(nreq, seed=11, arena_gib=64):
    """C5: continuous-batching server vs plain per-column path on host
    requests; burst throughput, then Poisson arrivals at a fixed rate."""
    import numpy as np

    import indexing_amd as ia
    from indexing_amd import workloads as W
    specs = W.column_specs(nreq, seed)
    progress(f"serve: generating {nreq} requests")
    # requests and results live in page-locked host memory (both paths)
    buf, layout = W.make_columns(specs, alloc=ia.pinned_empty)
    cols = [buf[o:o + n] for o, n in layout]
    total = int(sum(n for _, n in layout))
    span = buf.size
    okeys, ooffs, olens, owords = (ia.pinned_empty(span), ia.pinned_empty(span),
                                   ia.pinned_empty(span), ia.pinned_empty(2 * span))
    outs = [(okeys[o:o + n], ooffs[o:o + n], olens[o:o + n], owords[2 * o:2 * (o + n)])
            for o, n in layout]

    def run(srv, arrivals, reps=1):
        # latencies of `reps` passes over the same arrivals are pooled
        progress(f"serve: {len(arrivals)} requests x {reps}")
        srv.reset_stats()
        for _ in range(reps):
            t0 = time.perf_counter()
            pend = []
            for c, o, at in zip(cols, outs, arrivals):
                delay = t0 + at - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                pend.append(srv.submit(c, out=o))
            for p in pend:
                srv.wait(p)
        st = srv.stats()
        return {"p50_ms": round(st["p50_us"] / 1e3, 3), "p99_ms": round(st["p99_us"] / 1e3, 3),
                "gvalues_s": round(st["values"] / st["busy_s"] / 1e9, 3),
                "batches": int(st["batches"]), "max_batch_cols": int(st["max_batch_cols"])}

    burst = [0.0] * nreq
    res = {}
    with ia.Server(slots=1, arena_bytes=8 << 30, plain=True) as plain:
        run(plain, burst[:8])  # warm-up (allocations)
        res["plain_burst"] = run(plain, burst)
    # batches of at most 2^24 values (one largest request), so that the copies
    # of one slot overlap the build and copies of the other and no request
    # waits behind a long batch
    with ia.Server(slots=2, arena_bytes=arena_gib << 30, max_batch_cols=1024,
                   max_batch_values=1 << 24) as srv:
        run(srv, burst[:8])
        res["served_burst"] = run(srv, burst)
        rate = res["plain_burst"]["gvalues_s"] * 1e9 * 0.5 / (total / nreq)  # requests/s
        gaps = np.random.default_rng(seed).exponential(1.0 / rate, nreq)
        arrivals = np.cumsum(gaps) - gaps[0]
        res["served_poisson"] = run(srv, arrivals, POISSON_REPS)
    with ia.Server(slots=1, arena_bytes=8 << 30, plain=True) as plain:
        run(plain, burst[:8])
        res["plain_poisson"] = run(plain, arrivals, POISSON_REPS)
    res.update({"requests": nreq, "values": total, "rate_req_s": round(rate, 1),
                "workload": f"{nreq} requests, lengths log-uniform [2^12, 2^24], cardinality "
                            "log-uniform [2, 2^16], half uniform half Zipf(1.0), seed 11, "
                            f"page-locked host memory in and out, {arena_gib} GiB arena, 2 slots, "
                            "batches of at most 2^24 values; Poisson p50/p99 over "
                            f"{POISSON_REPS} passes of the arrivals",
                "throughput_ratio_burst": round(res["served_burst"]["gvalues_s"] /
                                                res["plain_burst"]["gvalues_s"], 2)})
    return res


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
    ap.add_argument("--serve-requests", type=int, default=400,
                    help="requests of the serving workload (0 skips it)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import indexing_amd as ia

    n = args.n
    if args.column == "uniform":
        host = ia.generate(n, "uniform", seed=1, lo=0, hi=65535)
        column = "uniform [0,65535], seed 1"
    else:
        host = ia.generate(n, "zipf-runs", seed=1, cardinality=1 << 20, zipf_s=1.2,
                           run_mean=4096.0)
        column = "Zipf(1.2) over 2^20 values in runs of mean length 4096, seed 1"
    d_values = torch.from_numpy(host.view(np.int32)).to("cuda:0")
    ctx = ia.Context(0)
    # tile shape per length bucket
    tuning = ctx.tuning() if args.no_autotune else ctx.autotune()
    out = ctx.build_device(d_values)                   # allocates outputs + arena
    for _ in range(args.warmup):
        ctx.build_device(d_values, out=out)
    torch.cuda.synchronize()

    ctx.timing_begin(args.steps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.build_device(d_values, out=out)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    group_ms, calls = ctx.timing_end()
    counts = ctx.read_counts(out[4])
    rank_mode = ctx.rank_mode
    assert counts.fault == 0, "device reported a fault"
    assert calls == args.steps

    ms_per_step = elapsed * 1e3 / args.steps
    value = n * args.steps / elapsed / 1e9

    # sizes for the byte model
    L, K, P = int(counts.length), int(counts.keycnt), int(counts.passes)
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
        "config": {"workload": f"single-column WAH build, n=2^{n.bit_length() - 1} uint32 "
                               f"{column}, resident in HBM",
                   "n": n, "k": int(counts.k), "L": L, "K": K, "radix_passes": P,
                   "sort_mode": "runs" if counts.mode else "pairs",
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
        "kernel_ms": {name: round(g / calls, 4) for name, g in zip(
            ["histogram", "pass0", "pass1", "pass2", "pass3", "tail", "lengths"], group_ms)},
    }
    ctx.close()
    del d_values, out
    torch.cuda.empty_cache()
    if args.batch_cols:
        result["batched"] = bench_batched(args.batch_cols)
    if args.serve_requests:
        result["serve"] = bench_serve(args.serve_requests)
    if not args.no_cpu_baseline:
        progress("cpu baseline")
        result["cpu_baseline"] = cpu_baseline(min(CPU_SAMPLE, n))
    print(json.dumps(result))


if __name__ == "__main__":
    main()
