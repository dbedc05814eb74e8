"""Serving: continuous batching of build requests (wahgpu_server_*).

Every result must be the same alone and in the mix: each request's index is
compared with the oracle (src/cpu.cpp:313-328) and with a single-column build
of the same column, whatever batch it was coalesced into.  Arrivals come in
bursts and from several threads so that requests arrive while others are on
the GPU.
"""
import threading
import time

import numpy as np
import pytest

import indexing_amd as ia
from oracle import oracle
from tests.test_gpu_batch import mixed_columns
from tests.test_gpu_parity import assert_same

pytestmark = pytest.mark.gpu


def test_burst_is_coalesced_and_exact(ctx):
    cols = mixed_columns(21, 60, max_log=16)
    with ia.Server(slots=2, arena_bytes=2 << 30, max_batch_cols=64) as srv:
        print(f"[serve] burst of {len(cols)}", flush=True)
        pend = [srv.submit(c, rid_base=3 * i) for i, c in enumerate(cols)]
        got = [srv.wait(p) for p in pend]
        st = srv.stats()
    assert st["requests"] == len(cols)
    assert st["batches"] < len(cols) and st["max_batch_cols"] > 1, st
    for i, (c, g) in enumerate(zip(cols, got)):
        assert_same(g, oracle.build(c, rid_base=3 * i), f"request {i}")
        alone = ctx.build(c, rid_base=3 * i)
        assert np.array_equal(g.words, alone.words) and np.array_equal(g.keys, alone.keys)


def test_concurrent_submitters_small_slots(ctx):
    """Slots of 2^16 values force many batches; four threads submit with
    pauses so requests keep arriving while batches run."""
    cols = mixed_columns(22, 48, max_log=15)
    results = {}
    with ia.Server(slots=3, arena_bytes=1 << 30, max_batch_cols=16,
                   max_batch_values=1 << 16) as srv:
        def worker(k):
            rng = np.random.default_rng(k)
            mine = []
            for i in range(k, len(cols), 4):
                mine.append((i, srv.submit(cols[i], rid_base=i)))
                time.sleep(float(rng.uniform(0, 0.002)))
            for i, p in mine:
                results[i] = srv.wait(p)
        th = [threading.Thread(target=worker, args=(k,)) for k in range(4)]
        print(f"[serve] 4 submitters, {len(cols)} requests", flush=True)
        for t in th:
            t.start()
        for t in th:
            t.join()
        st = srv.stats()
    assert st["requests"] == len(cols) and st["slot_values"] == 1 << 16
    for i, c in enumerate(cols):
        assert_same(results[i], oracle.build(c, rid_base=i), f"request {i}")


def test_plain_mode_matches(ctx):
    cols = mixed_columns(23, 12, max_log=14)
    with ia.Server(slots=1, arena_bytes=1 << 30, plain=True) as srv:
        pend = [srv.submit(c) for c in cols]
        got = [srv.wait(p) for p in pend]
        st = srv.stats()
    assert st["batches"] == len(cols)
    for i, (c, g) in enumerate(zip(cols, got)):
        assert_same(g, oracle.build(c), f"request {i}")


def test_admission_rejects_oversized_request(ctx):
    with ia.Server(slots=1, arena_bytes=1 << 30, max_batch_values=1 << 12) as srv:
        with pytest.raises(ia.WahError):
            srv.submit(np.zeros((1 << 12) + 1, np.uint32))
        p = srv.submit(np.arange(1 << 12, dtype=np.uint32) % 5)
        assert_same(srv.wait(p), oracle.build(np.arange(1 << 12, dtype=np.uint32) % 5))


def test_poll_until_done(ctx):
    c = np.random.default_rng(3).integers(0, 1000, 200000).astype(np.uint32)
    with ia.Server(slots=1, arena_bytes=1 << 30) as srv:
        p = srv.submit(c)
        t0 = time.time()
        while True:
            g = srv.poll(p)
            if g is not None:
                break
            assert time.time() - t0 < 60
            time.sleep(0.001)
    assert_same(g, oracle.build(c))


def test_request_waits_for_arena_memory_then_runs(ctx):
    """An arena that holds one 2^22-value request's footprint but not two:
    the second and third wait until a running batch frees its block, then
    run; every result is exact and the arena is never overcommitted."""
    n = 1 << 22
    cols = [ia.generate(n, "uniform", seed=40 + i, lo=0, hi=(1 << (10 + 3 * i)) - 1)
            for i in range(3)]
    arena = 300 << 20
    with ia.Server(slots=2, arena_bytes=arena, max_batch_cols=8) as srv:
        print("[serve] three 2^22 requests, arena for one", flush=True)
        pend = [srv.submit(c, rid_base=i) for i, c in enumerate(cols)]
        got = [srv.wait(p) for p in pend]
        st = srv.stats()
    assert st["memory_waits"] >= 1, st
    assert st["arena_bytes"] == arena and 0 < st["arena_peak"] <= arena, st
    assert st["batches"] == 3
    for i, (c, g) in enumerate(zip(cols, got)):
        assert_same(g, oracle.build(c, rid_base=i), f"request {i}")


def test_request_larger_than_the_arena_is_refused(ctx):
    with ia.Server(slots=1, arena_bytes=64 << 20) as srv:
        with pytest.raises(ia.WahError) as e:
            srv.submit(np.zeros(1 << 22, np.uint32))
        assert e.value.code == ia._native.ERANGE
        small = np.arange(100000, dtype=np.uint32) % 77
        assert_same(srv.wait(srv.submit(small)), oracle.build(small))


def test_served_mix_sample_matches_oracle_and_alone(ctx):
    """A seeded sample of the served mix (columns drawn from the batched
    workload's distributions, seed 11) against the oracle and against the
    same column built alone."""
    from indexing_amd import workloads as W
    specs = [s for s in W.column_specs(200, 11) if s[0] <= 1 << 21]
    pick = np.random.default_rng(11).choice(len(specs), 16, replace=False)
    cols = [W.make_column(specs[i]) for i in pick]
    with ia.Server(slots=2, arena_bytes=8 << 30, max_batch_cols=1024,
                   max_batch_values=1 << 24) as srv:
        print(f"[serve] mix sample of {len(cols)}", flush=True)
        pend = [srv.submit(c) for c in cols]
        got = [srv.wait(p) for p in pend]
    for i, (c, g) in enumerate(zip(cols, got)):
        assert_same(g, oracle.build(c), f"sample {i}")
        alone = ctx.build(c)
        assert ia.index_checksum(g.keys, g.offsets, g.lengths, g.words) == \
            ia.index_checksum(alone.keys, alone.offsets, alone.lengths, alone.words)
