"""GPU parity: libwahgpu (through the C ABI) against the CPU oracle.

Bit-exact equality of every key, offset, length and word (the reference's
`cpu` program is the oracle for the whole path, src/cpu.cpp:313-328), plus the
reference's per-stage checks (src/phases.cpp WITH_CPU_TESTS, :775-807,
:825-854, :912-930, :935-948, :1097-1123) for the staged kernels.
"""
import numpy as np
import pytest

import indexing_amd as ia
from oracle import oracle
from tests.wah_defs import check_invariants

pytestmark = pytest.mark.gpu


def u32(t):
    return t.detach().cpu().numpy().view(np.uint32)


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.uint32).view(np.int32)).to("cuda:0")


def assert_same(got, ref, what=""):
    assert got.k == ref["k"], f"{what} k {got.k} != {ref['k']}"
    assert got.length == ref["length"], f"{what} L {got.length} != {ref['length']}"
    assert got.keycnt == ref["keycnt"], f"{what} K {got.keycnt} != {ref['keycnt']}"
    for name in ("keys", "offsets", "lengths", "words"):
        g, r = getattr(got, name), ref[name]
        if not np.array_equal(g, r):
            bad = np.flatnonzero(g != r)[:5] if g.size == r.size else "size"
            raise AssertionError(f"{what} {name} differ at {bad}")


def columns():
    """(name, values) cases: sizes straddling 31-row chunks and the 2048/4096
    tiles, empty and single columns, constant columns, full 32-bit keys."""
    rng = np.random.default_rng(7)
    out = [
        ("empty", np.zeros(0, np.uint32)),
        ("one", np.array([5], np.uint32)),
        ("sample32", np.array([10, 7, 22, 6, 7, 1, 9, 42, 2, 5, 13, 3, 2, 1, 0, 1, 18, 18, 3,
                               13, 5, 9, 0, 3, 2, 19, 5, 23, 22, 10, 6, 22], np.uint32)),
        ("const", np.full(10000, 77, np.uint32)),
        ("zeros", np.zeros(4097, np.uint32)),
        ("max32", np.full(333, 0xFFFFFFFF, np.uint32)),
        ("two_vals", rng.integers(0, 2, 65536).astype(np.uint32)),
        ("sorted", np.sort(rng.integers(0, 1000, 20000)).astype(np.uint32)),
        ("reverse", np.arange(50000, 0, -1).astype(np.uint32)),
        ("distinct32", rng.permutation(1 << 16).astype(np.uint32) * 65537),
    ]
    for n in (30, 31, 32, 61, 62, 63, 2047, 2048, 2049, 4095, 4096, 4097, 8191, 12289, 100003):
        out.append((f"u16_{n}", rng.integers(0, 65536, n).astype(np.uint32)))
    for n in (1000, 70000):
        out.append((f"u32_{n}", rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)))
        out.append((f"small_{n}", rng.integers(0, 7, n).astype(np.uint32)))
    out.append(("zipf_runs", ia.generate(300000, "zipf-runs", seed=3, cardinality=1 << 20,
                                         zipf_s=1.2, run_mean=300)))
    out.append(("zipf", ia.generate(200000, "zipf", seed=5, cardinality=5000, zipf_s=1.0)))
    return out


@pytest.mark.parametrize("name,values", columns(), ids=[c[0] for c in columns()])
def test_fused_build_matches_oracle(ctx, name, values):
    print(f"[fused] {name} n={values.size}", flush=True)
    got = ctx.build(values)
    ref = oracle.build(values)
    assert_same(got, ref, name)
    if values.size <= 5000:
        check_invariants(values, got.keys, got.offsets, got.lengths, got.words)


@pytest.mark.parametrize("name,values", columns()[:12], ids=[c[0] for c in columns()[:12]])
def test_staged_build_matches_oracle(ctx, name, values):
    print(f"[staged] {name} n={values.size}", flush=True)
    got = ctx.build(values, staged=True, timed=True)
    ref = oracle.build(values)
    assert_same(got, ref, name)
    assert set(got.times) >= {"sort", "merge", "fuse", "column_length", "total"}


# rows straddling 2^31, past 2369637147 and up to 2^32 - 1 take the tail's
# full division by 31 (below 2^31 it divides by a multiply-high)
@pytest.mark.parametrize("rid_base", [1, 30, 31, 1000, 123456789, (1 << 31) - 2500,
                                      3_000_000_000, (1 << 32) - 5000])
def test_rid_base(ctx, rid_base):
    values = np.random.default_rng(rid_base).integers(0, 50, 5000).astype(np.uint32)
    got = ctx.build(values, rid_base=rid_base)
    ref = oracle.build(values, rid_base=rid_base)
    assert_same(got, ref, f"rid_base={rid_base}")
    check_invariants(values, got.keys, got.offsets, got.lengths, got.words, rid_base=rid_base)


def test_build_device_matches_host_build(ctx):
    import torch
    values = ia.generate(300000, "uniform", seed=11, lo=0, hi=65535)
    ref = oracle.build(values)
    keys, offs, lens, words, counts = ctx.build_device(dev(values))
    torch.cuda.synchronize()
    c = ctx.read_counts(counts)
    assert (c.k, c.length, c.keycnt, c.bound, c.passes) == (
        ref["k"], ref["length"], ref["keycnt"], int(values.max()), 2)
    K, L = ref["keycnt"], ref["length"]
    assert np.array_equal(u32(keys)[:K], ref["keys"])
    assert np.array_equal(u32(offs)[:K], ref["offsets"])
    assert np.array_equal(u32(lens)[:K], ref["lengths"])
    assert np.array_equal(u32(words)[:L], ref["words"])


def test_builds_alternating_on_two_streams(ctx):
    """Builds queued on two torch streams in turn share the context's
    scratch: each switch orders the new stream after the old one.  Growing
    sizes make the arena regrow while the other stream's build may still
    run."""
    import torch
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    vals = [ia.generate(n, "uniform", seed=i, lo=0, hi=65535)
            for i, n in enumerate((100003, 300007, 1000033, 3000017, 200001, 2500009))]
    refs = [oracle.build(v) for v in vals]
    outs = []
    for i, v in enumerate(vals):
        with torch.cuda.stream(streams[i % 2]):
            d = dev(v)
            outs.append((d, ctx.build_device(d)))
    torch.cuda.synchronize()
    for i, ((d, (keys, offs, lens, words, counts)), ref) in enumerate(zip(outs, refs)):
        c = ctx.read_counts(counts)
        assert c.fault == 0
        K, L = ref["keycnt"], ref["length"]
        assert (c.k, c.length, c.keycnt) == (ref["k"], L, K), i
        assert np.array_equal(u32(keys)[:K], ref["keys"]), i
        assert np.array_equal(u32(offs)[:K], ref["offsets"]), i
        assert np.array_equal(u32(lens)[:K], ref["lengths"]), i
        assert np.array_equal(u32(words)[:L], ref["words"]), i


# ---- stages (the WITH_CPU_TESTS comparisons of src/phases.cpp) -------------

STAGE_CASES = [
    ("u16", lambda: np.random.default_rng(1).integers(0, 65536, 70001).astype(np.uint32)),
    ("small", lambda: np.random.default_rng(2).integers(0, 5, 9000).astype(np.uint32)),
    ("runs", lambda: ia.generate(50000, "zipf-runs", seed=9, cardinality=1000, run_mean=100)),
    ("u32", lambda: np.random.default_rng(3).integers(0, 1 << 32, 5000,
                                                      dtype=np.uint64).astype(np.uint32)),
]


@pytest.mark.parametrize("name,make", STAGE_CASES, ids=[c[0] for c in STAGE_CASES])
def test_each_stage_matches_oracle(ctx, name, make):
    import torch
    values = make()
    print(f"[stages] {name} n={values.size}", flush=True)
    # sort_rids_by_value (src/phases.cpp:775-807)
    keys, rids = ctx.sort_rids_by_value(dev(values))
    ok, orid = oracle.sort_rids_by_value(values)
    torch.cuda.synchronize()
    assert np.array_equal(u32(keys), ok) and np.array_equal(u32(rids), orid)
    # produce_chunk_id_literals (:825-854)
    chids, lits = ctx.produce_chunk_id_literals(rids)
    och, oli = oracle.produce_chunk_id_literals(orid)
    torch.cuda.synchronize()
    assert np.array_equal(u32(chids), och) and np.array_equal(u32(lits), oli)
    # merge_lit_by_val_chids (:912-930)
    mv, mc, ml = ctx.merge_lit_by_val_chids(keys, chids, lits)
    rv, rc, rl = oracle.merge_lit_by_val_chids(ok, och, oli)
    assert np.array_equal(u32(mv), rv) and np.array_equal(u32(mc), rc)
    assert np.array_equal(u32(ml), rl)
    # produce_fills (:935-948)
    fills = ctx.produce_fills(mv, mc)
    rf = oracle.produce_fills(rv, rc)
    torch.cuda.synchronize()
    assert np.array_equal(u32(fills), rf)
    # fuse_fill_literals + compute_column_length (:1097-1123)
    index = ctx.fuse_fill_literals(fills, ml)
    assert np.array_equal(u32(index), oracle.fuse_fill_literals(rf, rl))
    k2, o2, l2 = ctx.compute_column_length(mv, fills)
    rk, ro, rl2 = oracle.compute_column_length(rv, rf)
    assert np.array_equal(u32(k2), rk) and np.array_equal(u32(o2), ro)
    assert np.array_equal(u32(l2), rl2)


def test_merge_long_runs(ctx):
    """reduce-by-key with runs far longer than a tile (the primitive is general)."""
    n = 50000
    vals = np.repeat(np.arange(5, dtype=np.uint32), n // 5)
    chids = np.zeros(n, np.uint32)
    lits = np.random.default_rng(4).integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
    mv, mc, ml = ctx.merge_lit_by_val_chids(dev(vals), dev(chids), dev(lits))
    rv, rc, rl = oracle.merge_lit_by_val_chids(vals, chids, lits)
    assert np.array_equal(u32(mv), rv) and np.array_equal(u32(ml), rl)


# ---- primitives (src/sc.cpp, src/scan.cpp, src/segmented_scan.cpp checks) ---

@pytest.mark.parametrize("n", [0, 1, 100, 2048, 2049, 300001])
def test_primitives_match_oracle(ctx, n):
    import torch
    rng = np.random.default_rng(n)
    data = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
    valid = rng.integers(0, 2, n).astype(np.uint32)
    heads = (rng.random(n) < 0.01).astype(np.uint32)
    got = ctx.stream_compaction(dev(data), dev(valid))
    assert np.array_equal(u32(got), oracle.stream_compaction(data, valid))
    ex = ctx.exclusive_scan(dev(data))
    seg = ctx.segmented_inclusive_scan(dev(data), dev(heads))
    torch.cuda.synchronize()
    assert np.array_equal(u32(ex), oracle.exclusive_scan(data))
    assert np.array_equal(u32(seg), oracle.segmented_inclusive_scan(data, heads))


def test_both_rankings_match_oracle(ctx, monkeypatch):
    """The LDS fetch-add ranking (taken after the device check) and the
    ballot-match ranking give the same index; columns with 1-4 distinct
    digits put many lanes on one counter."""
    monkeypatch.setenv("WAHGPU_RANK", "match")
    m = ia.Context(0)
    try:
        assert m.rank_mode == 0
        print(f"[rank] default context rank mode {ctx.rank_mode}", flush=True)
        rng = np.random.default_rng(77)
        for card in (2, 3, 4, 256, 65536):
            v = (rng.integers(0, card, 300001) * 65793).astype(np.uint32)
            ref = oracle.build(v)
            assert_same(ctx.build(v), ref, f"default card={card}")
            assert_same(m.build(v), ref, f"match card={card}")
    finally:
        m.close()
