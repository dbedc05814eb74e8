"""Pins the CPU oracle (oracle/wah_oracle.c) before it is trusted as the
checker: hand-derived known answers, the bitmap definition (decode and a
from-first-principles encoder), brute-force definitions of each stage on tiny
inputs, and the literal O(n^2) reference sort against the packed sort.
"""
import json
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle
from tests.wah_defs import check_invariants, decode_rows, encode_from_bitmaps

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "known_answers.json")))


def hexw(ws):
    return np.array([int(w, 16) for w in ws], dtype=np.uint32)


def fill100_values():
    v = np.ones(100, np.uint32)
    v[0], v[99], v[70] = 5, 5, 7
    return v


@pytest.mark.parametrize("name", ["sample32", "fill100"])
@pytest.mark.parametrize("sort", ["bubble", "packed"])
def test_known_answers(name, sort):
    g = GOLD[name]
    values = np.array(g["values"], np.uint32) if "values" in g else fill100_values()
    r = oracle.build(values, sort=sort)
    assert (r["k"], r["length"], r["keycnt"]) == (g["k"], g["length"], g["keycnt"])
    assert r["keys"].tolist() == g["keys"]
    assert r["offsets"].tolist() == g["offsets"]
    assert r["lengths"].tolist() == g["lengths"]
    assert np.array_equal(r["words"], hexw(g["words"]))


def test_cpu_cli_builtin_sample_stdout():
    out = subprocess.run([oracle.CPU_BIN], stdout=subprocess.PIPE, text=True, check=True).stdout
    assert out == GOLD["sample32"]["stdout"]


def test_cpu_cli_print_decodes_sample():
    out = subprocess.run([oracle.CPU_BIN, "-p"], stdout=subprocess.PIPE, text=True,
                         check=True).stdout.splitlines()[4:]
    values = GOLD["sample32"]["values"]
    lines = dict(line.split("\t") for line in out)
    assert len(out) == 43  # keys 0..42 including empty lines for absent values
    for key in range(43):
        rows = [i for i, v in enumerate(values) if v == key]
        want = "".join("1" if i in rows else "0" for i in range(max(rows) + 1)) if rows else ""
        assert lines[str(key)] == want


def random_cases():
    rng = np.random.default_rng(2024)
    cases = []
    for n in [1, 2, 5, 30, 31, 32, 62, 63, 64, 93, 200, 1000]:
        for card in [1, 2, 3, 10, 1 << 16]:
            cases.append(rng.integers(0, card, n).astype(np.uint32))
    cases.append(rng.integers(0, 1 << 32, 500, dtype=np.uint64).astype(np.uint32))
    cases.append(np.repeat(rng.integers(0, 20, 40), rng.integers(1, 80, 40)).astype(np.uint32))
    return cases


@pytest.mark.parametrize("case", range(len(random_cases())))
def test_oracle_equals_bitmap_definition(case):
    values = random_cases()[case]
    r = oracle.build(values)
    keys, offs, lens, words = encode_from_bitmaps(values)
    assert np.array_equal(r["keys"], keys)
    assert np.array_equal(r["offsets"], offs)
    assert np.array_equal(r["lengths"], lens)
    assert np.array_equal(r["words"], words)
    check_invariants(values, r["keys"], r["offsets"], r["lengths"], r["words"])
    # k = number of distinct (value, chunk) pairs
    assert r["k"] == len({(int(v), i // 31) for i, v in enumerate(values)})


@pytest.mark.parametrize("base", [1, 17, 31, 1000])
def test_rid_base_equals_definition(base):
    values = np.random.default_rng(base).integers(0, 6, 300).astype(np.uint32)
    r = oracle.build(values, rid_base=base)
    keys, offs, lens, words = encode_from_bitmaps(values, rid_base=base)
    assert np.array_equal(r["words"], words) and np.array_equal(r["offsets"], offs)


def test_bubble_equals_packed_sort():
    rng = np.random.default_rng(5)
    for n in [0, 1, 2, 7, 100, 777]:
        v = rng.integers(0, 9, n).astype(np.uint32)
        a = oracle.sort_rids_by_value(v, sort="bubble")
        b = oracle.sort_rids_by_value(v, sort="packed")
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
        # stable: rids ascending inside every run of equal values
        ks, rs = b
        assert np.all((np.diff(ks.astype(np.int64)) > 0) |
                      ((np.diff(ks.astype(np.int64)) == 0) & (np.diff(rs.astype(np.int64)) > 0)))


def test_threaded_build_equals_serial():
    """The CPU baseline's threaded build (slices sorted on threads, merged by
    rank ranges) gives the serial build's index: sizes below and above the
    threading threshold, thread counts that do not divide n, few and many
    keys, 32-bit keys and a rid_base."""
    rng = np.random.default_rng(8)
    cases = [(0, 4, 0), (7, 3, 0), (5000, 2, 0), (70001, 65536, 13), (300007, 300, 0),
             (1 << 18, 1 << 32, 31 * 1000 + 7)]
    for n, hi, base in cases:
        v = rng.integers(0, hi, n, dtype=np.uint64).astype(np.uint32)
        a = oracle.build(v, rid_base=base)
        for th in (2, 3, 7, 16):
            b = oracle.build(v, rid_base=base, threads=th)
            assert (b["k"], b["length"], b["keycnt"]) == (a["k"], a["length"], a["keycnt"])
            for key in ("keys", "offsets", "lengths", "words"):
                assert np.array_equal(a[key], b[key]), (n, th, key)


def test_empty_column():
    r = oracle.build(np.zeros(0, np.uint32))
    assert (r["k"], r["length"], r["keycnt"]) == (0, 0, 0)


# ---- stages against brute-force definitions --------------------------------

def test_chunk_literals_closed_form():
    rids = np.array([0, 1, 30, 31, 32, 61, 62, 2**32 - 1], np.uint32)
    ch, li = oracle.produce_chunk_id_literals(rids)
    assert ch.tolist() == [int(r) // 31 for r in rids]
    assert li.tolist() == [0x80000000 | (1 << (int(r) % 31)) for r in rids]


def test_merge_is_or_over_equal_runs():
    vals = np.array([1, 1, 1, 2, 2, 1, 1], np.uint32)
    chid = np.array([0, 0, 1, 1, 1, 1, 1], np.uint32)
    lits = np.array([1, 2, 4, 8, 16, 32, 64], np.uint32)
    v, c, l = oracle.merge_lit_by_val_chids(vals, chid, lits)
    assert v.tolist() == [1, 1, 2, 1] and c.tolist() == [0, 1, 1, 1]
    assert l.tolist() == [3, 4, 24, 96]


def test_fills_definition():
    vals = np.array([3, 3, 3, 4, 4], np.uint32)
    chid = np.array([2, 3, 7, 0, 5], np.uint32)
    f = oracle.produce_fills(vals, chid)
    assert f.tolist() == [2, 0, 3, 0, 4]


def test_fuse_drops_zero_words_only():
    f = np.array([0, 3, 0], np.uint32)
    l = np.array([0x80000001, 0xFFFFFFFF, 0x80000002], np.uint32)
    assert oracle.fuse_fill_literals(f, l).tolist() == [0x80000001, 3, 0xFFFFFFFF, 0x80000002]


def test_column_length_definition():
    vals = np.array([1, 1, 4, 9, 9, 9], np.uint32)
    fills = np.array([0, 5, 2, 0, 0, 1], np.uint32)
    k, o, l = oracle.compute_column_length(vals, fills)
    assert k.tolist() == [1, 4, 9] and l.tolist() == [3, 2, 4] and o.tolist() == [0, 3, 5]


def test_scans_and_compaction_brute_force():
    rng = np.random.default_rng(9)
    x = rng.integers(0, 1 << 32, 300, dtype=np.uint64).astype(np.uint32)
    heads = (rng.random(300) < 0.1).astype(np.uint32)
    ex = oracle.exclusive_scan(x)
    assert ex.tolist() == [int(sum(int(a) for a in x[:i])) % (1 << 32) for i in range(300)]
    seg = oracle.segmented_inclusive_scan(x, heads)
    acc = 0
    for i in range(300):
        acc = int(x[i]) if (i == 0 or heads[i]) else (acc + int(x[i])) % (1 << 32)
        assert seg[i] == acc
    valid = rng.integers(0, 2, 300).astype(np.uint32)
    assert oracle.stream_compaction(x, valid).tolist() == [int(a) for a, m in zip(x, valid) if m]


def test_decoder_known_words():
    # a literal, a zero fill of 2 chunks, a literal with the top payload bit
    words = np.array([0x80000005, 0x00000002, 0xC0000000], np.uint32)
    s = oracle.decoded_bitmap(words)
    assert s == "101" + "0" * 28 + "0" * 62 + "0" * 30 + "1"
    assert decode_rows(words) == [0, 2, 31 * 3 + 30]
    assert oracle.decoded_bitmap(np.array([0x80000000], np.uint32)) == ""


def test_max_size_closed_form_matches_oracle_on_small_cases():
    """The closed form tests/test_gpu_maxsize.py checks 2^30-value builds
    against, pinned here against the oracle on small columns (partial first
    and last chunks, rid_base not a multiple of 31)."""
    from tests.test_gpu_maxsize import expected_words
    keys = [40, 10, 70, 20, 60, 30, 50]
    for n, rb in ((10000, 0), (9999, 1000003), (217 * 3 + 5, 62), (31 * 7, 30)):
        vals = np.array([keys[i % 7] for i in range(n)], np.uint32)
        ref = oracle.build(vals, rid_base=rb)
        pos = 0
        for i, k in enumerate(sorted(keys)):
            w = expected_words(n, rb, keys.index(k))
            assert ref["keys"][i] == k and ref["offsets"][i] == pos and ref["lengths"][i] == w.size
            assert np.array_equal(ref["words"][pos:pos + w.size], w), (n, rb, k)
            pos += w.size
