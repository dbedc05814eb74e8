"""Builds from several contexts on several streams at once.

Each context binds its own torch stream, and the builds are queued round
robin so that, without ordering, their kernels would share the device.  The
single-pass kernels wait on tiles that other workgroups of their launch draw
from per-XCD counters; the library keeps such launches from different
streams from overlapping on one device, so none of them may report a fault
(counts.fault) and every result equals the oracle's.
"""
import numpy as np
import pytest

import indexing_amd as ia
from oracle import oracle
from tests.test_gpu_parity import dev, u32

pytestmark = pytest.mark.gpu


def _check(outs, ref, tag):
    keys, offs, lens, words, counts = outs
    c = ia.Context.read_counts(counts)
    assert c.fault == 0, tag
    K, L = ref["keycnt"], ref["length"]
    assert (c.k, c.length, c.keycnt) == (ref["k"], L, K), tag
    assert np.array_equal(u32(keys)[:K], ref["keys"]), tag
    assert np.array_equal(u32(offs)[:K], ref["offsets"]), tag
    assert np.array_equal(u32(lens)[:K], ref["lengths"]), tag
    assert np.array_equal(u32(words)[:L], ref["words"]), tag


def test_four_contexts_on_four_streams(ctx):
    import torch
    sizes = [1 << 16, 300007, 1 << 19, 1000003, 77777, 1 << 20]
    cols = [ia.generate(n, "uniform", seed=60 + i, lo=0, hi=(1 << (8 + 2 * i)) - 1)
            for i, n in enumerate(sizes)]
    refs = [oracle.build(c) for c in cols]
    ctxs = [ia.Context() for _ in range(4)]
    streams = [torch.cuda.Stream() for _ in range(4)]
    d_cols = [dev(c) for c in cols]
    torch.cuda.synchronize()
    print("[concurrency] 4 contexts x 4 streams, 6 rounds", flush=True)
    pending = []
    for rnd in range(6):
        for j, (cx, st) in enumerate(zip(ctxs, streams)):
            i = (rnd + j) % len(cols)
            with torch.cuda.stream(st):
                pending.append((i, rnd, j, cx.build_device(d_cols[i])))
    torch.cuda.synchronize()
    for i, rnd, j, outs in pending:
        _check(outs, refs[i], f"column {i} round {rnd} context {j}")
    for cx in ctxs:
        cx.close()


def test_batches_and_single_builds_on_two_streams(ctx):
    import torch
    from tests.test_gpu_batch import mixed_columns
    cols = mixed_columns(31, 40, max_log=18)
    big = ia.generate(1 << 21, "uniform", seed=5, lo=0, hi=65535)
    ref_big = oracle.build(big)
    other = ia.Context()
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    d_big = dev(big)
    torch.cuda.synchronize()
    print("[concurrency] batch of 40 on one stream, 2^21 builds on another", flush=True)
    singles = []
    with torch.cuda.stream(s2):  # queued first, still running when the batch starts
        for _ in range(3):
            singles.append(other.build_device(d_big))
    with torch.cuda.stream(s1):
        got = ctx.build_batch(cols)
    torch.cuda.synchronize()
    for i, (c, g) in enumerate(zip(cols, got)):
        ref = oracle.build(c)
        assert np.array_equal(g.words, ref["words"]) and np.array_equal(g.keys, ref["keys"]), i
    for k, outs in enumerate(singles):
        _check(outs, ref_big, f"2^21 build {k}")
    other.close()
