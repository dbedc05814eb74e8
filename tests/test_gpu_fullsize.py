"""GPU parity at the benchmark configurations' stated sizes and seeds.

The reference checks its whole path on the real input (src/phases.cpp
WITH_CPU_TESTS, :775-807, :1097-1123).  Here every configuration is built at
its full size and compared with the oracle:

  c1  2^20 uniform [0, 65535], seed 42: word for word against oracle.build,
      with the default radix ranking and with the ballot-match ranking
  c2  2^28 uniform [0, 65535], seed 1 ) digests of keys, offsets, lengths
  c3  2^28 zipf-runs, seed 1          ) and words against the oracle's
  c4  the 1024-column batch, seed 7   ) (tests/golden/fullsize.json, written
                                          by tests/golden/make_fullsize.py)

The generated inputs are digested too, so a mismatch cannot come from the
generator.  These sizes reach what the small cases do not: 30-bit look-back
fields, ~16K radix tiles and ~32K tail tiles per column, 2^28 row ids.
"""
import concurrent.futures as cf
import json
import os

import numpy as np
import pytest

import indexing_amd as ia
from indexing_amd import workloads as W
from indexing_amd.digest import array_digest, index_digest
from oracle import oracle

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "fullsize.json")))


def u32(t):
    return t.detach().cpu().numpy().view(np.uint32)


def make_single(name):
    p = dict(GOLD[name]["params"])
    n, dist = p.pop("n"), p.pop("dist")
    v = ia.generate(n, dist, **p)
    assert array_digest(v) == GOLD[name]["input"], f"{name}: generated input differs"
    return v


@pytest.fixture
def match_ctx(monkeypatch):
    monkeypatch.setenv("WAHGPU_RANK", "match")
    c = ia.Context(0)
    assert c.rank_mode == 0
    yield c
    c.close()


def device_digest(ctx, values):
    """Fused device build of one column -> (counts, digest of its index)."""
    import torch
    d = torch.from_numpy(values.view(np.int32)).to("cuda:0")
    keys, offs, lens, words, counts = ctx.build_device(d)
    torch.cuda.synchronize()
    c = ctx.read_counts(counts)
    assert c.fault == 0
    K, L = int(c.keycnt), int(c.length)
    dig = index_digest(u32(keys[:K]), u32(offs[:K]), u32(lens[:K]), u32(words[:L]))
    del d, keys, offs, lens, words
    torch.cuda.empty_cache()
    return c, dig


def check_single(ctx, name):
    g = GOLD[name]
    values = make_single(name)
    print(f"[fullsize] {name} n={values.size} rank_mode={ctx.rank_mode}", flush=True)
    c, dig = device_digest(ctx, values)
    assert (c.k, c.length, c.keycnt) == (g["k"], g["length"], g["keycnt"])
    assert dig == g["index"], f"{name}: index digest differs from the oracle's"


def test_c1_matches_oracle_both_rankings(ctx, match_ctx):
    values = make_single("c1")
    ref = oracle.build(values)
    assert ref["length"] == GOLD["c1"]["length"]
    for c in (ctx, match_ctx):
        print(f"[fullsize] c1 rank_mode={c.rank_mode}", flush=True)
        got = c.build(values)
        assert (got.k, got.length, got.keycnt) == (ref["k"], ref["length"], ref["keycnt"])
        for field in ("keys", "offsets", "lengths", "words"):
            assert np.array_equal(getattr(got, field), ref[field]), field
        again = c.build(values)  # a repeat gives the same bytes
        assert np.array_equal(again.words, got.words)


def test_c2_uniform_2p28_matches_oracle(ctx):
    check_single(ctx, "c2")


def test_c2_uniform_2p28_ballot_ranking_matches_oracle(match_ctx):
    check_single(match_ctx, "c2")


def test_c3_zipf_runs_2p28_matches_oracle(ctx):
    check_single(ctx, "c3")


def test_c4_1024_column_batch_matches_oracle(ctx):
    import torch
    g = GOLD["c4"]["per_column"]
    specs = W.column_specs(1024, 7)
    buf, layout = W.make_columns(specs)
    print(f"[fullsize] c4 1024 columns, {sum(n for _, n in layout)} values", flush=True)
    with cf.ThreadPoolExecutor(16) as ex:
        ins = list(ex.map(lambda ol: array_digest(buf[ol[0]:ol[0] + ol[1]]), layout))
    assert ins == [c["input"] for c in g], "generated c4 input differs"
    d = torch.from_numpy(buf.view(np.int32)).to("cuda:0")
    del buf
    keys, offs, lens, words, counts = ctx.build_batch_device(d, layout)
    torch.cuda.synchronize()
    del d
    cs = ctx.read_counts_batch(counts, len(layout))
    assert all(c.fault == 0 for c in cs)
    got_counts = [(int(c.k), int(c.length), int(c.keycnt)) for c in cs]
    assert got_counts == [(c["k"], c["length"], c["keycnt"]) for c in g]
    hk, ho, hl, hw = (u32(x) for x in (keys, offs, lens, words))

    def dig(i):
        o, _ = layout[i]
        K, L = got_counts[i][2], got_counts[i][1]
        return index_digest(hk[o:o + K], ho[o:o + K], hl[o:o + K], hw[2 * o:2 * o + L])

    with cf.ThreadPoolExecutor(16) as ex:
        digs = list(ex.map(dig, range(len(layout))))
    bad = [i for i, (a, c) in enumerate(zip(digs, g)) if a != c["index"]]
    assert not bad, f"c4 columns differ from the oracle: {bad[:10]}"
