"""Wide radix digits (up to 11 bits, 2048 bins) against the oracle.

A pair-mode column of at least 2^14 values whose keys take fewer passes of
up to 11 bits than of 8 bits sorts with wide digits when its length bucket's
tuning allows them (digit_bits 11).  The index must not change: every result
is compared with the oracle (src/cpu.cpp:313-328), and counts.mode / passes
tell which sort ran.  Key widths cover one, two and three wide passes (9-11,
17-22, 25-32 bits) and the widths where wide digits do not save a pass (12,
16, 24 bits: the column keeps 8-bit digits); lengths straddle the 8192-key
wide tiles; batches mix wide, 8-bit, run-mode, empty and short columns.
"""
import numpy as np
import pytest

import indexing_amd as ia
from indexing_amd import _native as N
from oracle import oracle
from tests.test_gpu_batch import mixed_columns, pack, per_column
from tests.test_gpu_parity import assert_same, dev

pytestmark = pytest.mark.gpu

WIDE = [11] * N.TUNE_BUCKETS


def passes8(bits):
    return (bits + 7) // 8


def expect_wide(bits, n):
    """Whether a uniform column of `bits`-bit keys (top bit set somewhere)
    takes wide digits, and with how many passes."""
    aw = (bits + 10) // 11
    return n >= (1 << 14) and aw < passes8(bits), aw


def keys_of(bits, n, seed):
    rng = np.random.default_rng(seed)
    v = rng.integers(0, 1 << bits, n, dtype=np.uint64).astype(np.uint32)
    v[n // 2] = (1 << bits) - 1  # the top bit is set somewhere
    return v


@pytest.fixture(scope="module")
def wctx():
    c = ia.Context(0)
    c.set_tuning([0] * N.TUNE_BUCKETS, WIDE)
    assert c.tuning()["digit_bits"] == WIDE
    yield c
    c.close()


def counts_of(ctx, v, rid_base=0):
    """Build on the device, return (index, counts) for one column."""
    d = dev(v)
    out = ctx.build_device(d, rid_base=rid_base)
    c = ctx.read_counts(out[4])
    got = per_column(out, [c], [(0, v.size)])[0]
    return got, c


@pytest.mark.parametrize("bits", [9, 10, 11, 12, 16, 17, 20, 22, 24, 25, 30, 32])
@pytest.mark.parametrize("n", [1 << 14, 100003, (1 << 20) + 7])
def test_single_column_key_widths(wctx, bits, n):
    v = keys_of(bits, n, 1000 * bits + n % 997)
    print(f"[wide] bits={bits} n={n}", flush=True)
    got, c = counts_of(wctx, v, rid_base=17)
    assert c.fault == 0
    wide, aw = expect_wide(bits, n)
    assert c.mode == (2 if wide else 0), (bits, n, c.mode)
    assert c.passes == (aw if wide else passes8(bits)), (bits, n, c.passes)
    assert_same(got, oracle.build(v, rid_base=17), f"bits {bits} n {n}")


def test_short_and_run_mode_columns_keep_8_bit_digits(wctx):
    short = keys_of(20, (1 << 14) - 1, 5)  # below the wide minimum
    got, c = counts_of(wctx, short)
    assert c.mode == 0 and c.passes == 3
    assert_same(got, oracle.build(short), "short")
    runs = ia.generate(1 << 18, "zipf-runs", seed=4, cardinality=1 << 20, zipf_s=1.2,
                       run_mean=200)
    got, c = counts_of(wctx, runs)
    assert c.mode == 1
    assert_same(got, oracle.build(runs), "runs")


def test_batch_mixes_wide_and_8_bit_columns(wctx):
    rng = np.random.default_rng(9)
    cols = mixed_columns(91, 24, max_log=18)
    for bits in (10, 20, 31, 13):
        cols.append(keys_of(bits, int(rng.integers(1 << 14, 1 << 18)), bits))
    cols.append(ia.generate(1 << 17, "zipf-runs", seed=2, cardinality=1 << 20, zipf_s=1.2,
                            run_mean=100))
    buf, layout = pack(cols, gap=3)
    print(f"[wide] batch of {len(cols)}", flush=True)
    out = wctx.build_batch_device(dev(buf), layout)
    counts = wctx.read_counts_batch(out[4], len(cols))
    got = per_column(out, counts, layout)
    modes = [int(c.mode) for c in counts]
    assert modes.count(2) >= 3, modes
    for i, col in enumerate(cols):
        assert counts[i].fault == 0
        assert_same(got[i], oracle.build(col), f"column {i} (mode {modes[i]})")


def test_match_ranking_wide(monkeypatch):
    monkeypatch.setenv("WAHGPU_RANK", "match")
    m = ia.Context(0)
    try:
        assert m.rank_mode == 0
        m.set_tuning([0] * N.TUNE_BUCKETS, WIDE)
        for bits in (10, 21, 32):
            v = keys_of(bits, 300001, bits)
            got, c = counts_of(m, v)
            assert c.mode == 2
            assert_same(got, oracle.build(v), f"match bits {bits}")
    finally:
        m.close()


def test_graph_replay_and_noalloc_with_wide_digits(wctx):
    import torch
    cols = [keys_of(b, n, b) for b, n in ((10, 40000), (20, 70001), (16, 50000), (27, 90000))]
    buf, layout = pack(cols)
    refs = [oracle.build(c) for c in cols]
    g = ia.Context(0)
    try:
        g.set_tuning([0] * N.TUNE_BUCKETS, WIDE)
        g.batch_reserve(buf.size, len(cols))
        d = dev(buf)
        outs = (g._empty(buf.size), g._empty(buf.size), g._empty(buf.size),
                g._empty(2 * buf.size), torch.zeros(len(cols) * 10, dtype=torch.int32,
                                                     device="cuda:0"))
        for rep in range(3):
            g.build_batch_device(d, layout, out=outs)
            counts = g.read_counts_batch(outs[4], len(cols))
            got = per_column(outs, counts, layout)
            assert [int(c.mode) for c in counts] == [2, 2, 0, 2]
            for i in range(len(cols)):
                assert_same(got[i], refs[i], f"replay {rep} column {i}")
        assert g.graph_stats() == (1, 3)
    finally:
        g.close()
    v = keys_of(20, 1 << 20, 3)
    out = wctx.build_device_noalloc(dev(v))
    c = wctx.read_counts(out[4])
    assert c.mode == 2 and c.fault == 0
    assert_same(per_column(out, [c], [(0, v.size)])[0], oracle.build(v), "noalloc")


def test_autotune_reports_digit_widths():
    c = ia.Context(0)
    try:
        t = c.autotune()
        assert all(d in (8, 11) for d in t["digit_bits"])
        assert t["digit_bits"][0] == 8 and t["digit_ms"][0] == [0.0, 0.0]
        for b in range(1, N.TUNE_BUCKETS):
            ms8, ms11 = t["digit_ms"][b]
            assert ms8 > 0 and ms11 > 0
            assert t["digit_bits"][b] == (11 if ms11 < ms8 else 8)
        print(f"[wide] tuning {t['shape']} {t['digit_bits']} {t['digit_ms']}", flush=True)
    finally:
        c.close()
