"""Host-side parts of libwahgpu that need no GPU: the C ABI exports, the value
file reader (`ifstream >> uint32_t` semantics), the decoder, the seeded
generator and the exact stitching of partial indexes (the merger).
"""
import os
import re

import numpy as np
import pytest

import indexing_amd as ia
from indexing_amd import _native as N
from oracle import oracle
from tests.wah_defs import encode_from_bitmaps

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "wahgpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(wahgpu_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    L = N.lib()
    declared = header_functions()
    assert len(declared) >= 30
    for name in declared:
        assert hasattr(L, name), f"{name} missing from libwahgpu.so"
    assert set(declared) == set(N.SIGNATURES), "ctypes table out of sync with the header"


def test_no_device_here_is_an_error_not_a_fallback():
    # On a GPU-less host the context cannot be created; nothing computes.
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    with pytest.raises(ia.WahError):
        ia.Context(0)


def write(tmp_path, text, name="v.txt"):
    p = tmp_path / name
    p.write_bytes(text.encode() if isinstance(text, str) else text)
    return str(p)


@pytest.mark.parametrize("text,expect", [
    ("1\n2\n3\n", [1, 2, 3]),
    ("  7\t8\r\n9\v10\f11", [7, 8, 9, 10, 11]),
    ("4294967295 0", [4294967295, 0]),
    ("4294967296 5", []),              # overflow fails the extraction
    ("12a 13", [12]),                  # next extraction fails on 'a'
    ("1,2,3", [1]),                    # a non-space separator ends the input
    ("+5 -1 -0", [5, 4294967295, 0]),  # sign accepted, '-' wraps mod 2^32
    ("- 5", []),
    ("", []),
    ("\n\n", []),
    ("00042", [42]),
])
def test_text_reader_matches_ifstream_semantics(tmp_path, text, expect):
    assert ia.read_values_text(write(tmp_path, text)).tolist() == expect


def test_binary_reader_roundtrip(tmp_path):
    v = np.random.default_rng(1).integers(0, 1 << 32, 1001, dtype=np.uint64).astype(np.uint32)
    p = write(tmp_path, v.astype("<u4").tobytes(), "v.bin")
    assert np.array_equal(ia.read_values_binary(p), v)


def test_missing_file_is_an_error():
    with pytest.raises(ia.WahError):
        ia.read_values_text("/nonexistent/values.txt")


def test_decoder_agrees_with_oracle_decoder():
    rng = np.random.default_rng(3)
    for _ in range(50):
        values = rng.integers(0, 4, rng.integers(1, 400)).astype(np.uint32)
        r = oracle.build(values)
        for i in range(r["keycnt"]):
            w = r["words"][r["offsets"][i]:r["offsets"][i] + r["lengths"][i]]
            assert ia.decode_bitmap(w) == oracle.decoded_bitmap(w)


def mix64(z):
    M = (1 << 64) - 1
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
    return z ^ (z >> 31)


def test_generator_uniform_is_the_documented_splitmix_stream():
    seed, lo, hi = 42, 0, 65535
    got = ia.generate(16, "uniform", seed=seed, lo=lo, hi=hi)
    M = (1 << 64) - 1
    want = []
    for i in range(16):
        r = mix64((mix64(seed) + (i + 1) * 0x9E3779B97F4A7C15) & M)
        want.append(lo + (((r >> 32) * (hi - lo + 1)) >> 32))
    assert got.tolist() == want


def test_generator_ranges_and_reproducibility():
    a = ia.generate(100000, "uniform", seed=1, lo=10, hi=20)
    assert a.min() == 10 and a.max() == 20
    assert np.array_equal(a, ia.generate(100000, "uniform", seed=1, lo=10, hi=20))
    assert np.array_equal(a[500:900], ia.generate(400, "uniform", seed=1, lo=10, hi=20,
                                                  start=500))
    z = ia.generate(200000, "zipf-runs", seed=3, cardinality=1 << 20, zipf_s=1.2, run_mean=4096)
    assert np.array_equal(z[123456:130000], ia.generate(6544, "zipf-runs", seed=3,
                                                        cardinality=1 << 20, zipf_s=1.2,
                                                        run_mean=4096, start=123456))
    runs = np.count_nonzero(np.diff(z.astype(np.int64))) + 1
    assert 20 < runs < 200          # ~ n / 4096 runs
    f = ia.generate(100000, "zipf", seed=4, cardinality=1000, zipf_s=1.0)
    assert f.max() < 1000 and np.count_nonzero(f == 0) > np.count_nonzero(f == 10)
    full = ia.generate(100000, "uniform", seed=9, lo=0, hi=0xFFFFFFFF)
    assert full.max() > 0xF0000000


@pytest.mark.parametrize("cut", [[], [31], [30], [1, 2, 3], [100, 200, 248, 600], [500]])
def test_merger_stitches_partials_exactly(cut):
    """Partials built over consecutive row ranges (any split, aligned to 31 or
    not) stitch to the whole-column index."""
    values = np.random.default_rng(len(cut)).integers(0, 9, 1000).astype(np.uint32)
    bounds = [0] + cut + [values.size]
    m = ia.Merger()
    for a, b in zip(bounds[:-1], bounds[1:]):
        p = oracle.build(values[a:b], rid_base=a)
        m.add(p["keys"], p["offsets"], p["words"])
    got = m.result()
    keys, offs, lens, words = encode_from_bitmaps(values)
    assert np.array_equal(got.keys, keys) and np.array_equal(got.offsets, offs)
    assert np.array_equal(got.lengths, lens) and np.array_equal(got.words, words)


def test_merger_rejects_rows_out_of_order():
    values = np.zeros(100, np.uint32)
    late = oracle.build(values[50:], rid_base=50)
    early = oracle.build(values[:50], rid_base=0)
    m = ia.Merger()
    m.add(late["keys"], late["offsets"], late["words"])
    with pytest.raises(ia.WahError):
        m.add(early["keys"], early["offsets"], early["words"])
