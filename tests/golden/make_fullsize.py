"""Writes tests/golden/fullsize.json: oracle digests of the index at the
stated sizes and seeds of the benchmark configurations.

    python tests/golden/make_fullsize.py [--jobs 8]

For each configuration the generated input is digested (so a run elsewhere
can show it indexed the same column) and the oracle (oracle/wah_oracle.c, the
C restatement of the reference's serial build, src/cpu.cpp:176-272) builds
the index, whose counts and digest (indexing_amd.digest.index_digest) are
recorded:

  c1   2^20 values, uniform [0, 65535], seed 42
  c2   2^28 values, uniform [0, 65535], seed 1
  c3   2^28 values, Zipf(1.2) over 2^20 values in runs of mean length 4096,
       seed 1
  c4   1024 columns (indexing_amd.workloads.column_specs(1024, 7)): one
       record per column

The oracle runs single-threaded per column; the 1024 columns of c4 run on a
thread pool (the oracle's ctypes calls release the GIL).  About 45 s for c2,
15 s for c3 and a few minutes for c4 on 8 cores.
"""
import argparse
import concurrent.futures as cf
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import indexing_amd as ia  # noqa: E402
from indexing_amd import workloads as W  # noqa: E402
from indexing_amd.digest import array_digest, index_digest  # noqa: E402
from oracle import oracle  # noqa: E402

OUT = os.path.join(HERE, "fullsize.json")

SINGLE = {
    "c1": dict(n=1 << 20, dist="uniform", seed=42, lo=0, hi=65535),
    "c2": dict(n=1 << 28, dist="uniform", seed=1, lo=0, hi=65535),
    "c3": dict(n=1 << 28, dist="zipf-runs", seed=1, cardinality=1 << 20, zipf_s=1.2,
               run_mean=4096.0),
}


def single_column(params):
    p = dict(params)
    n = p.pop("n")
    dist = p.pop("dist")
    return ia.generate(n, dist, **p)


def record(values):
    r = oracle.build(values)
    return {"n": int(values.size), "input": array_digest(values), "k": r["k"],
            "length": r["length"], "keycnt": r["keycnt"],
            "index": index_digest(r["keys"], r["offsets"], r["lengths"], r["words"])}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=min(8, os.cpu_count() or 1))
    args = ap.parse_args()
    out = {"_comment": "Oracle digests at the benchmark configurations' stated sizes and "
                       "seeds; written by tests/golden/make_fullsize.py (indexing_amd.digest)."}
    for name, params in SINGLE.items():
        t0 = time.perf_counter()
        values = single_column(params)
        rec = record(values)
        rec["params"] = params
        out[name] = rec
        print(f"{name}: k={rec['k']} L={rec['length']} K={rec['keycnt']} "
              f"({time.perf_counter() - t0:.1f} s)", flush=True)
        del values
    t0 = time.perf_counter()
    specs = W.column_specs(1024, 7)

    def one(spec):
        return record(W.make_column(spec))

    with cf.ThreadPoolExecutor(args.jobs) as ex:
        cols = list(ex.map(one, specs))
    out["c4"] = {"columns": 1024, "seed": 7, "per_column": cols}
    print(f"c4: {sum(c['n'] for c in cols)} values ({time.perf_counter() - t0:.1f} s)",
          flush=True)
    with open(OUT, "w") as f:
        json.dump(out, f, indent=0)
        f.write("\n")


if __name__ == "__main__":
    main()
