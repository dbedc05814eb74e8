"""Run mode: columns whose rows repeat their predecessor are sorted as one item
per run of equal values inside a 31-row chunk (value, first rid, literal
bits).  The index must be the oracle's (src/cpu.cpp:313-328) bit for bit,
whatever the runs look like at chunk, wave, tile and column boundaries.
"""
import numpy as np
import pytest

import indexing_amd as ia
from oracle import oracle
from tests.test_gpu_batch import pack, per_column
from tests.test_gpu_parity import assert_same, dev

pytestmark = pytest.mark.gpu


def runs_column(n, mean, card, seed, dtype_max=1 << 16):
    """Runs of geometric length with the given mean over `card` values."""
    rng = np.random.default_rng(seed)
    out = np.empty(n, np.uint32)
    i = 0
    while i < n:
        ln = int(rng.geometric(1.0 / mean))
        out[i:i + ln] = rng.integers(0, card) % dtype_max
        i += ln
    return out


def cases():
    rng = np.random.default_rng(5)
    out = [
        ("const_1", np.full(1, 4, np.uint32), 0),
        ("const_31", np.full(31, 4, np.uint32), 0),
        ("const_big", np.full(100000, 123456, np.uint32), 0),
        ("two_vals", rng.integers(0, 2, 70000).astype(np.uint32), 0),
        ("sorted", np.sort(rng.integers(0, 3000, 50001)).astype(np.uint32), 0),
        ("runs_1", runs_column(40000, 1.5, 50, 1), 0),
        ("runs_8", runs_column(16383, 8, 1000, 2), 0),
        ("runs_31", runs_column(16384, 31, 1000, 3), 0),
        ("runs_64", runs_column(16385, 64, 1 << 20, 4, 1 << 20), 0),
        ("runs_1000", runs_column(300000, 1000, 1 << 30, 5, 1 << 31), 0),
        ("runs_tiles", runs_column(3 * 16384 + 77, 40, 500, 6), 0),
        ("rid5", runs_column(50000, 20, 300, 7), 5),
        ("rid30", runs_column(50000, 20, 300, 8), 30),
        ("rid31", runs_column(50000, 20, 300, 9), 31),
        ("rid_big", runs_column(70000, 50, 300, 10), 1_000_003),
        # rows past 2^31 (the tail's full division; rows from 2369637147 on
        # are where a multiply-high by 2^36 / 31 would be off by one)
        ("rid_3e9", runs_column(70000, 50, 300, 11), 3_000_000_000),
        ("rid_to_2p32", runs_column(70000, 50, 300, 12), (1 << 32) - 70000),
        ("zipf_runs", ia.generate(1 << 20, "zipf-runs", seed=3, cardinality=1 << 20, zipf_s=1.2,
                                  run_mean=4096), 0),
    ]
    return out


def run_starts(values, rid_base):
    """Rows that do not repeat their predecessor inside a 31-row chunk."""
    rows = rid_base + np.arange(values.size, dtype=np.uint64)
    start = np.ones(values.size, bool)
    start[1:] = (values[1:] != values[:-1]) | (rows[1:] % 31 == 0)
    return int(start.sum())


@pytest.mark.parametrize("name,values,rid_base", cases(), ids=[c[0] for c in cases()])
def test_run_mode_matches_oracle(ctx, name, values, rid_base):
    print(f"[runs] {name} n={values.size} rid_base={rid_base}", flush=True)
    got = ctx.build(values, rid_base=rid_base)
    expect_run = 3 * run_starts(values, rid_base) < 2 * values.size
    assert got.mode == (1 if expect_run else 0)
    assert_same(got, oracle.build(values, rid_base=rid_base), name)


def test_most_cases_use_run_mode():
    assert sum(3 * run_starts(v, r) < 2 * v.size for _, v, r in cases()) >= len(cases()) - 2


def test_uniform_stays_in_pair_mode(ctx):
    v = np.random.default_rng(1).integers(0, 65536, 100000).astype(np.uint32)
    got = ctx.build(v)
    assert got.mode == 0
    assert_same(got, oracle.build(v))


def test_batch_mixes_modes(ctx):
    rng = np.random.default_rng(12)
    cols = []
    for i in range(24):
        if i % 3 == 0:
            cols.append(rng.integers(0, 1 << 16, int(rng.integers(1, 40000))).astype(np.uint32))
        else:
            cols.append(runs_column(int(rng.integers(1, 60000)), float(rng.uniform(2, 300)),
                                    int(rng.integers(1, 5000)), 100 + i))
    buf, layout = pack(cols, gap=1)
    rid_bases = [(37 * i) % 97 for i in range(len(cols))]
    spec = [(o, n, r) for (o, n), r in zip(layout, rid_bases)]
    print(f"[runs] batch of {len(cols)}", flush=True)
    out = ctx.build_batch_device(dev(buf), spec)
    counts = ctx.read_counts_batch(out[4], len(cols))
    got = per_column(out, counts, layout)
    assert {c.mode for c in counts} == {0, 1}
    for i, c in enumerate(cols):
        assert_same(got[i], oracle.build(c, rid_base=rid_bases[i]), f"column {i}")


def key_width_columns():
    """Columns on both sides of the 16-bit key boundary (keys below 2^16 move
    between radix passes as 2-byte words), in pair and run mode."""
    rng = np.random.default_rng(21)
    top16 = rng.integers(0, 65536, 90000).astype(np.uint32)
    top16[rng.integers(0, top16.size)] = 65535            # max 65535: 2-byte keys
    over = top16.copy()
    over[rng.integers(0, over.size)] = 65536              # max 65536: 4-byte keys
    runs16 = runs_column(70000, 40, 65536, 22, 1 << 16)
    runs16[-1] = 65535
    runs17 = runs16.copy()
    runs17[100] = 1 << 16
    byte_keys = rng.integers(0, 256, 50000).astype(np.uint32)  # one active pass
    # 2-byte keys whose low byte is constant: the only active pass sorts on
    # bits 8..15 (the first pass packs key and row into one staged word)
    high_byte = (rng.integers(0, 256, 40001) << 8).astype(np.uint32)
    low_const = ((rng.integers(0, 256, 3 * 16384 + 5) << 8) | 0x37).astype(np.uint32)
    return [("max_65535", top16), ("max_65536", over), ("runs_max_65535", runs16),
            ("runs_max_65536", runs17), ("one_pass", byte_keys), ("high_byte_only", high_byte),
            ("low_byte_constant", low_const)]


@pytest.mark.parametrize("name,values", key_width_columns(), ids=[c[0] for c in key_width_columns()])
def test_key_width_boundary(ctx, name, values):
    print(f"[width] {name} n={values.size} max={int(values.max())}", flush=True)
    got = ctx.build(values, rid_base=3)
    assert got.bound == int(values.max())
    if name in ("high_byte_only", "low_byte_constant", "one_pass"):
        assert got.passes == 1  # the constant digit's pass is skipped
    assert_same(got, oracle.build(values, rid_base=3), name)


def test_key_widths_mixed_in_one_batch(ctx):
    cols = [v for _, v in key_width_columns()]
    rng = np.random.default_rng(23)
    cols += [rng.integers(0, 1 << 32, 20000, dtype=np.uint64).astype(np.uint32),
             rng.integers(0, 1000, 30001).astype(np.uint32)]
    buf, layout = pack(cols, gap=3, align=2)
    print(f"[width] batch of {len(cols)}", flush=True)
    out = ctx.build_batch_device(dev(buf), layout)
    counts = ctx.read_counts_batch(out[4], len(cols))
    got = per_column(out, counts, layout)
    for i, c in enumerate(cols):
        assert counts[i].fault == 0
        assert_same(got[i], oracle.build(c), f"column {i}")


def span_columns():
    """Columns of blocks of equal values (value = f(row // span)): blocks
    shorter than, equal to and longer than the histogram's 256-row spans and
    1024-row chunk windows, aligned to them and not, with keys that vary in
    their low digit only (the others the same in every row) or in all four.
    The histogram counts such chunks per span, per window or per row."""
    n = 3 * 4096 + 1234
    rows = np.arange(n, dtype=np.uint64)
    out = []
    for span in (64, 255, 256, 257, 1024, 1025, 3000):
        b = rows // np.uint64(span)
        out.append((f"span{span}_low", (np.uint64(0x01020300) + b % np.uint64(5)).astype(np.uint32)))
        out.append((f"span{span}_all", ((b * np.uint64(2654435761)) % np.uint64(1 << 32)).astype(np.uint32)))
    return out


@pytest.mark.parametrize("name,values", span_columns(), ids=[c[0] for c in span_columns()])
def test_span_patterns_match_oracle(ctx, name, values):
    print(f"[spans] {name} n={values.size}", flush=True)
    got = ctx.build(values, rid_base=7)
    assert_same(got, oracle.build(values, rid_base=7), name)
