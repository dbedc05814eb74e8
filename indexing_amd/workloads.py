"""Synthetic multi-column workloads for the batched and served builds.

A column is drawn as: length log-uniform in [2^lo_log, 2^hi_log],
cardinality log-uniform in [2, 2^16], values i.i.d. uniform over the
cardinality for even-numbered columns and Zipf(s=1.0) ranks for odd ones.
The draws come from numpy's PCG64 with the workload seed; each column's
values from the library generator seeded per column, so a column is the same
whether it is generated alone or in a batch.
"""
import concurrent.futures as cf
import os

import numpy as np

from . import generate


def column_specs(count, seed, lo_log=12, hi_log=24, card_hi_log=16):
    """-> list of (n, cardinality, dist, column_seed)."""
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(count):
        n = int(2.0 ** rng.uniform(lo_log, hi_log))
        card = max(2, int(2.0 ** rng.uniform(1, card_hi_log)))
        dist = "uniform" if i % 2 == 0 else "zipf"
        specs.append((n, card, dist, seed * 1_000_003 + i))
    return specs


def make_column(spec, out=None):
    n, card, dist, cseed = spec
    if dist == "uniform":
        return generate(n, "uniform", seed=cseed, lo=0, hi=card - 1, out=out)
    return generate(n, "zipf", seed=cseed, cardinality=card, zipf_s=1.0, out=out)


def make_columns(specs, align=64, threads=None, alloc=None):
    """Generates every column into one buffer at `align`-value boundaries
    (alloc(n) makes the buffer, e.g. pinned_empty; default numpy zeros).
    -> (buffer, [(offset, n)])."""
    layout, pos = [], 0
    for n, *_ in specs:
        layout.append((pos, n))
        pos += (n + align - 1) // align * align
    buf = alloc(max(pos, 1)) if alloc else np.zeros(max(pos, 1), np.uint32)
    threads = threads or min(16, os.cpu_count() or 1)
    with cf.ThreadPoolExecutor(threads) as ex:
        list(ex.map(lambda a: make_column(a[0], out=buf[a[1][0]:a[1][0] + a[1][1]]),
                    zip(specs, layout)))
    return buf, layout
