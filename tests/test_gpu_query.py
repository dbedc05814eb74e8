"""GPU query side against the oracle and brute force:

  decode   each key's WAH words -> row bitmap, against the oracle's
           decoded_bitmap (src/cpu.cpp:47-74) and against the rows holding
           the key
  range    OR of the keys in [lo, hi], against the rows whose value is in
           [lo, hi]
  encode   row bitmap -> canonical WAH words, against the words written from
           the definition (per nonempty chunk: the nonzero gap fill, then the
           literal) and, for a key's rows, against the oracle's words
  combine  AND / OR / XOR / AND NOT of two keys' streams, against the
           canonical words of the bitwise result of their row sets
on random and zipf-runs columns (the fill-heavy shape of the 2^28 zipf-runs
configuration, scaled down) and on ragged sizes.
"""
import numpy as np
import pytest

import indexing_amd as ia
from oracle import oracle

pytestmark = pytest.mark.gpu


def u32(t):
    return t.detach().cpu().numpy().view(np.uint32)


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.uint32).view(np.int32)).to("cuda:0")


def rows_to_bitmap(rows, nrows):
    b = np.zeros(((nrows + 31) // 32) * 32, bool)
    b[np.asarray(rows, np.int64)] = True
    return np.packbits(b, bitorder="little").view(np.uint32)


def bitmap_rows(bm, nrows):
    bits = np.unpackbits(np.asarray(bm, np.uint32).view(np.uint8), bitorder="little")
    return np.flatnonzero(bits[:nrows])


def canonical_words(rows):
    """WAH words of a row set from the definition (src/cpu.cpp:223-257)."""
    rows = np.asarray(rows, np.int64)
    if rows.size == 0:
        return np.zeros(0, np.uint32)
    ch = rows // 31
    uc, idx = np.unique(ch, return_inverse=True)
    lit = np.zeros(uc.size, np.int64)
    np.bitwise_or.at(lit, idx, np.left_shift(1, rows % 31))
    fill = uc - np.concatenate([[-1], uc[:-1]]) - 1
    out = np.empty(2 * uc.size, np.int64)
    out[0::2], out[1::2] = fill, lit | 0x80000000
    keep = np.ones(2 * uc.size, bool)
    keep[0::2] = fill != 0
    return out[keep].astype(np.uint32)


def columns():
    rng = np.random.default_rng(11)
    return [
        ("uniform_small_card", rng.integers(0, 50, 200_003).astype(np.uint32)),
        ("uniform_u16", rng.integers(0, 65536, 300_007).astype(np.uint32)),
        ("zipf_runs", ia.generate(1 << 21, "zipf-runs", seed=1, cardinality=1 << 20, zipf_s=1.2,
                                  run_mean=4096.0)),
        ("ragged_62", rng.integers(0, 3, 62).astype(np.uint32)),
        ("one_row", np.array([7], np.uint32)),
    ]


COLS = columns()


@pytest.mark.parametrize("name,values", COLS, ids=[c[0] for c in COLS])
def test_decode_every_sampled_key(ctx, name, values):
    import torch
    idx = ctx.build(values)
    n = values.size
    rng = np.random.default_rng(3)
    pick = rng.choice(idx.keycnt, size=min(idx.keycnt, 40), replace=False)
    print(f"[query] decode {name} n={n} keys={pick.size}", flush=True)
    for i in pick:
        w = idx.key_words(i)
        bm, last = ctx.decode_device(dev(w), n)
        torch.cuda.synchronize()
        rows = np.flatnonzero(values == idx.keys[i])
        assert np.array_equal(bitmap_rows(u32(bm), n), rows), (name, int(idx.keys[i]))
        ref = oracle.decoded_bitmap(w)
        assert last == len(ref)
        got = np.zeros(last, np.uint8)
        got[rows] = 1
        assert "".join("1" if x else "0" for x in got) == ref


@pytest.mark.parametrize("name,values", COLS, ids=[c[0] for c in COLS])
def test_range_query(ctx, name, values):
    import torch
    n = values.size
    keys, offs, lens, words, counts = ctx.build_device(dev(values))
    torch.cuda.synchronize()
    c = ctx.read_counts(counts)
    K = int(c.keycnt)
    kd, od, ld = keys[:K], offs[:K], lens[:K]
    vmax = int(values.max())
    rng = np.random.default_rng(5)
    cases = [(0, vmax), (0, 0), (vmax, vmax), (vmax + 1, vmax + 5), (5, 3)]
    cases += [tuple(sorted(rng.integers(0, vmax + 2, 2))) for _ in range(6)]
    for lo, hi in cases:
        bm, matched = ctx.query_range_device(kd, od, ld, words, lo, hi, n)
        torch.cuda.synchronize()
        want = np.flatnonzero((values >= lo) & (values <= hi))
        assert np.array_equal(bitmap_rows(u32(bm), n), want), (name, lo, hi)
        uniq = np.unique(values)
        assert matched == int(((uniq >= lo) & (uniq <= hi)).sum())


@pytest.mark.parametrize("nrows", [1, 30, 31, 32, 62, 2047, 63488, 63489, 1_000_003])
def test_encode_matches_definition(ctx, nrows):
    import torch
    rng = np.random.default_rng(nrows)
    for dens in (0.0, 0.0005, 0.05, 0.7, 1.0):
        rows = np.flatnonzero(rng.random(nrows) < dens) if dens < 1.0 else np.arange(nrows)
        bm = rows_to_bitmap(rows, nrows)
        w = ctx.encode_device(dev(bm), nrows)
        torch.cuda.synchronize()
        assert np.array_equal(u32(w), canonical_words(rows)), (nrows, dens)
        back, _ = ctx.decode_device(w, nrows)
        torch.cuda.synchronize()
        assert np.array_equal(u32(back), bm)


def test_encode_of_a_key_equals_the_oracle_words(ctx):
    import torch
    values = ia.generate(500_000, "zipf", seed=2, cardinality=300, zipf_s=1.0)
    ref = oracle.build(values)
    for i in (0, 1, 17, ref["keycnt"] - 1):
        rows = np.flatnonzero(values == ref["keys"][i])
        w = ctx.encode_device(dev(rows_to_bitmap(rows, values.size)), values.size)
        torch.cuda.synchronize()
        o = ref["offsets"][i]
        assert np.array_equal(u32(w), ref["words"][o:o + ref["lengths"][i]])


@pytest.mark.parametrize("name,values", COLS[:3], ids=[c[0] for c in COLS[:3]])
def test_combine_pairs_of_keys(ctx, name, values):
    import torch
    idx = ctx.build(values)
    n = values.size
    rng = np.random.default_rng(9)
    print(f"[query] combine {name} n={n}", flush=True)
    for _ in range(6):
        i, j = rng.choice(idx.keycnt, 2, replace=False) if idx.keycnt > 1 else (0, 0)
        ra = np.flatnonzero(values == idx.keys[i])
        rb = np.flatnonzero(values == idx.keys[j])
        a, b = dev(idx.key_words(i)), dev(idx.key_words(j))
        # keys' row sets are disjoint: AND them with the union too
        union = np.union1d(ra, rb)
        for op, want in (("or", union), ("and", np.intersect1d(ra, rb)),
                         ("xor", np.setxor1d(ra, rb)), ("andnot", np.setdiff1d(ra, rb))):
            got = ctx.combine_device(op, a, b, n)
            torch.cuda.synchronize()
            assert np.array_equal(u32(got), canonical_words(want)), (name, op)
        uw = ctx.combine_device("or", a, b, n)
        got = ctx.combine_device("and", a, uw, n)
        torch.cuda.synchronize()
        assert np.array_equal(u32(got), canonical_words(ra))


def test_combine_capacity_is_checked(ctx):
    import torch
    values = np.random.default_rng(1).integers(0, 4, 10_000).astype(np.uint32)
    idx = ctx.build(values)
    a, b = dev(idx.key_words(0)), dev(idx.key_words(1))
    with pytest.raises(ia.WahError):
        ctx.combine_device("or", a, b, values.size, cap=3)
    torch.cuda.synchronize()
