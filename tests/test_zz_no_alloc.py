"""wahgpu_build_device_noalloc (scratch in a static device array, no device
allocation by the library) against wahgpu_build_device on the same columns:
two calls in a row, the second smaller, so the second runs on the first's
leftover scratch contents.  Bit-exact keys, offsets, lengths and words."""
import numpy as np
import pytest

import indexing_amd as ia
from indexing_amd import _native as N
from oracle import oracle

pytestmark = pytest.mark.gpu


def u32(t):
    return t.detach().cpu().numpy().view(np.uint32)


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.uint32).view(np.int32)).to("cuda:0")


def index_of(ctx, out):
    keys, offs, lens, words, counts = out
    c = ctx.read_counts(counts)
    assert c.fault == 0
    K, L = int(c.keycnt), int(c.length)
    return (int(c.k), L, K), [u32(keys)[:K], u32(offs)[:K], u32(lens)[:K], u32(words)[:L]]


def test_noalloc_matches_build_device_twice_second_smaller():
    import torch
    cases = [
        ("uniform 3M", ia.generate(3_000_017, "uniform", seed=42, lo=0, hi=65535)),
        ("zipf-runs 200K", ia.generate(200_003, "zipf-runs", seed=5, cardinality=1 << 20,
                                       zipf_s=1.2, run_mean=300.0)),
    ]
    with ia.Context(0) as ctx:
        for name, values in cases:
            print(f"[noalloc] {name} n={values.size}", flush=True)
            d = dev(values)
            got = ctx.build_device_noalloc(d)
            want = ctx.build_device(d)
            torch.cuda.synchronize()
            gc, ga = index_of(ctx, got)
            wc, wa = index_of(ctx, want)
            assert gc == wc, name
            for g, w, field in zip(ga, wa, ("keys", "offsets", "lengths", "words")):
                assert np.array_equal(g, w), f"{name}: {field}"
            ref = oracle.build(values)
            assert gc == (ref["k"], ref["length"], ref["keycnt"])
            assert np.array_equal(ga[3], ref["words"])


def test_noalloc_limits():
    import torch
    with ia.Context(0) as ctx:
        # the largest column it takes, then an empty one
        values = ia.generate(N.NOALLOC_MAX_VALUES, "uniform", seed=3, lo=0, hi=1 << 20)
        d = dev(values)
        got = ctx.build_device_noalloc(d)
        want = ctx.build_device(d)
        torch.cuda.synchronize()
        assert index_of(ctx, got)[0] == index_of(ctx, want)[0]
        assert all(np.array_equal(a, b) for a, b in zip(index_of(ctx, got)[1],
                                                         index_of(ctx, want)[1]))
        empty = ctx.build_device_noalloc(dev(np.zeros(0, np.uint32)))
        torch.cuda.synchronize()
        assert index_of(ctx, empty)[0] == (0, 0, 0)
        # one value over the limit is refused before any work is queued
        big = torch.zeros(N.NOALLOC_MAX_VALUES + 1, dtype=torch.int32, device="cuda:0")
        with pytest.raises(ia.WahError) as e:
            ctx.build_device_noalloc(big)
        assert e.value.code == N.ERANGE and "exceeds the limit" in str(e.value)
