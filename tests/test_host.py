"""Host-side parts of libwahgpu that need no GPU: the C ABI exports, the value
file reader (`ifstream >> uint32_t` semantics), the decoder, the seeded
generator and the exact stitching of partial indexes (the merger).
"""
import os
import re
import subprocess

import numpy as np
import pytest

import indexing_amd as ia
from indexing_amd import _native as N
from oracle import oracle
from tests.wah_defs import encode_from_bitmaps

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "indexing_amd", "build")


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


# ---- parallel text reader / writer ------------------------------------------

def py_ifstream_read(data):
    """Independent statement of the `while (in >> uint32_t)` loop over bytes."""
    out, i, e = [], 0, len(data)
    ws = b" \t\n\v\f\r"
    while True:
        while i < e and data[i] in ws:
            i += 1
        if i == e:
            return out
        neg = False
        if data[i] in b"+-":
            neg = data[i] == ord("-")
            i += 1
        j = i
        while j < e and 48 <= data[j] <= 57:
            j += 1
        if j == i:
            return out
        v = int(data[i:j])
        if v > 0xFFFFFFFF:
            return out
        out.append((-v) & 0xFFFFFFFF if neg else v)
        i = j


@pytest.mark.parametrize("threads", [1, 2, 3, 7, 16])
@pytest.mark.parametrize("text,expect", [
    ("1\n2\n3\n", [1, 2, 3]),
    ("12a 13", [12]),
    ("4294967296 5", []),
    ("+5 -1 -0", [5, 4294967295, 0]),
    ("1 2 3 4 5 6 7 8 9 x 10 11 12", [1, 2, 3, 4, 5, 6, 7, 8, 9]),
    ("  ", []),
])
def test_text_reader_threads_match_serial_semantics(tmp_path, threads, text, expect):
    p = write(tmp_path, text)
    assert ia.read_values_text(p, threads=threads).tolist() == expect


def test_text_reader_random_files_every_thread_count(tmp_path):
    rng = np.random.default_rng(17)
    seps = [b" ", b"\n", b"\t", b"\r\n", b"  \n ", b"\v", b"\f"]
    for trial in range(60):
        m = int(rng.integers(0, 400))
        toks = []
        for _ in range(m):
            r = rng.random()
            if r < 0.01:
                toks.append(b"x")                      # a failed extraction
            elif r < 0.02:
                toks.append(b"99999999999")            # overflow
            elif r < 0.04:
                toks.append(b"-" + str(int(rng.integers(0, 1000))).encode())
            elif r < 0.05:
                toks.append(b"7z")                     # a value, then a failure
            else:
                toks.append(str(int(rng.integers(0, 1 << 32))).encode())
        data = b"".join(t + seps[int(rng.integers(0, len(seps)))] for t in toks)
        if rng.random() < 0.3:
            data = data.rstrip()
        p = write(tmp_path, data, f"r{trial}.txt")
        want = py_ifstream_read(data)
        for threads in (1, 2, 5, 16):
            got = ia.read_values_text(p, threads=threads).tolist()
            assert got == want, (trial, threads)


def test_text_writer_roundtrip_and_matches_generate_cli(tmp_path):
    v = ia.generate(1 << 20, "uniform", seed=9, lo=0, hi=0xFFFFFFFF)
    p = str(tmp_path / "w.txt")
    ia.write_values_text(p, v, threads=8)
    assert np.array_equal(ia.read_values_text(p, threads=8), v)
    assert np.array_equal(ia.read_values_text(p, threads=1), v)
    q = str(tmp_path / "g.txt")
    subprocess.run([os.path.join(BUILD, "generate"), "-a", str(1 << 20), "-f", "0", "-t",
                    str(0xFFFFFFFF), "--seed", "9", "-o", q], check=True)
    assert open(p, "rb").read() == open(q, "rb").read()
    ia.write_values_text(p, v[:5], sep=",", threads=1)
    assert open(p, "rb").read() == (",".join(str(x) for x in v[:5]) + "\n").encode()


def test_tail_chunk_division_identities():
    """The tail's chunk arithmetic (wah_kernels.hip, k_tail).  With
    m = 0x84210843 = (2^36 + 29) / 31 and rid = 31q + k (k <= 30),
    rid * m / 2^36 = q + (k + 29 rid / 2^36) / 31, so mulhi(rid, m) >> 4 is
    rid // 31 whenever 29 rid < 2^36, i.e. for every rid up to 2369637128
    (all rids below 2^31); past that it fails first where k = 30, at
    2369637147.  Checked here at both ends of [0, 2^31), on a random sample
    and around the failure point; and (rid + q) & 31 == rid % 31 (rid + q =
    32q + k) on the same rids and the top of the 32-bit range."""
    magic = np.uint64(0x84210843)
    assert int(magic) * 31 == (1 << 36) + 29
    rng = np.random.default_rng(31)
    r = np.concatenate([np.arange(0, 1 << 22, dtype=np.uint64),
                        np.arange((1 << 31) - (1 << 22), 1 << 31, dtype=np.uint64),
                        rng.integers(0, 1 << 31, 1 << 22, dtype=np.uint64)])
    q = (r * magic) >> np.uint64(36)
    assert np.array_equal(q, r // np.uint64(31))
    assert np.array_equal((r + q) & np.uint64(31), r % np.uint64(31))
    r = np.arange(2_369_637_100, 2_369_637_200, dtype=np.uint64)
    bad = r[((r * magic) >> np.uint64(36)) != r // np.uint64(31)]
    assert bad[0] == 2_369_637_147 and 2_369_637_147 % 31 == 30
    assert 29 * 2_369_637_128 < (1 << 36) <= 29 * 2_369_637_129
    top = np.arange((1 << 32) - (1 << 22), 1 << 32, dtype=np.uint64)
    q = top // np.uint64(31)
    assert np.array_equal((top + q) & np.uint64(31), top % np.uint64(31))
