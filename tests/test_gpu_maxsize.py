"""Columns of the largest size the ABI takes (WAHGPU_MAX_VALUES = 2^30 - 1
values), with the largest rid_base that keeps every row id in 32 bits,
against a closed form of the index instead of the oracle (which would take
minutes at this size).

values[i] = key[i % 7].  Row R = rid_base + i belongs to key v when
(R - rid_base) % 7 == v, so in a chunk c (rows 31c .. 31c + 30) key v's bits
are the i in [0, 31) with (31c + i - rid_base) % 7 == v: seven literal
patterns, periodic in c.  Every full chunk holds 4 or 5 rows of every key, so
a key's words are, by the build's definition (src/cpu.cpp:223-257), its
first chunk as a fill when nonzero and then one literal per chunk up to its
last.  This exercises the 30-bit digit counts of the look-back words, 2^30
rows through the sort and the tail, and chunk ids near 2^27.
"""
import numpy as np
import pytest

import indexing_amd as ia
from indexing_amd import _native as N
from tests.test_gpu_parity import u32

pytestmark = pytest.mark.gpu

N_MAX = N.MAX_VALUES


def expected_words(n, rid_base, v):
    """Words of key v (rows R = rid_base + i with i % 7 == v)."""
    r0, r1 = rid_base, rid_base + n  # rows [r0, r1)
    first_row = r0 + v
    last_row = r0 + v + ((n - 1 - v) // 7) * 7
    cf, cl = first_row // 31, last_row // 31
    c = np.arange(cf, cl + 1, dtype=np.int64)
    m = (31 * c - r0) % 7  # (row - r0) % 7 of the chunk's bit 0
    pat = np.zeros(7, np.uint32)
    for mm in range(7):
        bits = 0
        for i in range(31):
            if (mm + i) % 7 == v:
                bits |= 1 << i
        pat[mm] = bits
    lit = pat[m]
    # the first and last chunks may be cut by the column's ends
    lo_mask = np.uint32(((1 << 31) - 1) & ~((1 << (r0 - 31 * cf)) - 1)) if r0 > 31 * cf else np.uint32((1 << 31) - 1)
    lit[0] &= lo_mask
    hi_end = r1 - 31 * cl  # rows of the last chunk inside the column
    if hi_end < 31:
        lit[-1] &= np.uint32((1 << hi_end) - 1)
    assert lit[0] and lit[-1], "first and last chunks hold a row of the key"
    words = lit | np.uint32(0x80000000)
    if cf > 0:
        words = np.concatenate([np.array([cf], np.uint32), words])
    return words


def build_max(ctx, keys, rid_base):
    import torch
    n = N_MAX
    print(f"[maxsize] n={n} rid_base={rid_base} keys={keys[:2]}..", flush=True)
    kt = torch.tensor(np.array(keys, np.uint32).view(np.int32), device="cuda:0")
    d = kt[torch.arange(n, device="cuda:0", dtype=torch.int64) % 7]
    out = ctx.build_device(d, rid_base=rid_base)
    c = ctx.read_counts(out[4])
    del d
    return out, c


@pytest.mark.parametrize("rid_base,scale", [(0, 1), ((1 << 32) - (1 << 30), 1), (0, 65537)],
                         ids=["rows-from-0", "rows-to-2^32", "32-bit-keys"])
def test_max_size_column_closed_form(ctx, rid_base, scale):
    import torch
    keys = [x * scale for x in (40, 10, 70, 20, 60, 30, 50)]  # distinct, not in residue order
    out, c = build_max(ctx, keys, rid_base)
    try:
        assert c.fault == 0
        exp = {keys[v]: expected_words(N_MAX, rid_base, v) for v in range(7)}
        K = 7
        L = sum(w.size for w in exp.values())
        assert (c.keycnt, c.length) == (K, L), (c.keycnt, c.length, L)
        assert c.k == sum(w.size - (1 if w[0] < 0x80000000 else 0) for w in exp.values())
        gk = u32(out[0][:K])
        go = u32(out[1][:K])
        gl = u32(out[2][:K])
        assert list(gk) == sorted(keys)
        pos = 0
        for i, key in enumerate(sorted(keys)):
            w = exp[key]
            assert go[i] == pos and gl[i] == w.size, (i, go[i], gl[i], pos, w.size)
            got = out[3][pos:pos + w.size].cpu().numpy().view(np.uint32)
            assert np.array_equal(got, w), f"key {key} words differ"
            pos += w.size
    finally:
        del out
        torch.cuda.empty_cache()


def test_one_past_the_maximum_is_refused(ctx):
    import torch
    d = torch.zeros(1, dtype=torch.int32, device="cuda:0")
    small = (d.clone(), d.clone(), d.clone(), d.clone(),
             torch.zeros(16, dtype=torch.int32, device="cuda:0"))
    with pytest.raises(ia.WahError):  # refused before anything is read or written
        ctx.build_device(d.expand(N_MAX + 1), out=small)
    with pytest.raises(ia.WahError):  # rows past 2^32
        ctx.build_device(torch.zeros(16, dtype=torch.int32, device="cuda:0"),
                         rid_base=(1 << 32) - 8)
