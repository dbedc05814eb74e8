"""Columns of keys below 2^16 whose two 8-bit digits both vary (two radix
passes with 16-bit keys between them), with first-pass digit buckets of every
shape: ~n/256 positions each, two nonempty buckets and 254 empty ones, 65536
keys a few times each (tiles of the second pass spanning tens of buckets), a
skewed low byte, a nearly constant high digit; ragged lengths and rid_base,
one column alone and all of them in a batch; every index against the oracle
(src/cpu.cpp:313-328).
"""
import numpy as np
import pytest

import indexing_amd as ia
from oracle import oracle
from tests.test_gpu_batch import pack, per_column
from tests.test_gpu_parity import assert_same, dev

pytestmark = pytest.mark.gpu


def cases():
    rng = np.random.default_rng(16)
    out = []
    # uniform 16-bit keys: buckets of ~n/256 positions
    for n in (65536, 100003, 1 << 20, (1 << 20) + 12345):
        out.append((f"uniform-{n}", rng.integers(0, 1 << 16, n).astype(np.uint32)))
    # two low bytes only: 254 empty buckets, starts repeated
    hi = rng.integers(0, 256, 300001).astype(np.uint32)
    lo = np.where(rng.random(300001) < 0.5, 3, 250).astype(np.uint32)
    out.append(("two-low-bytes", hi << 8 | lo))
    # 65536 keys a few times each: tiles span tens of buckets (binary search)
    out.append(("tiny-buckets", rng.permutation(np.repeat(np.arange(1 << 16, dtype=np.uint32), 3))))
    # skewed low byte: a few huge buckets and many small ones
    z = ia.generate(500000, "zipf", seed=5, cardinality=65536, zipf_s=1.1)
    out.append(("zipf-16bit", z))
    # keys 256..511 only in the high byte's low bit: high digit nearly constant
    out.append(("high-digit-2", (rng.integers(256, 768, 200000)).astype(np.uint32)))
    return out


@pytest.mark.parametrize("name,v", cases(), ids=[c[0] for c in cases()])
def test_single_column(ctx, name, v):
    print(f"[two-pass] {name} n={v.size}", flush=True)
    for rid_base in (0, 77):
        d = dev(v)
        out = ctx.build_device(d, rid_base=rid_base)
        c = ctx.read_counts(out[4])
        assert c.fault == 0 and c.mode == 0
        got = per_column(out, [c], [(0, v.size)])[0]
        assert_same(got, oracle.build(v, rid_base=rid_base), f"{name} rid_base={rid_base}")


def test_batch_of_two_pass_columns(ctx):
    cols = [v for _, v in cases()]
    buf, layout = pack(cols, gap=5)
    out = ctx.build_batch_device(dev(buf), layout)
    counts = ctx.read_counts_batch(out[4], len(cols))
    got = per_column(out, counts, layout)
    for i, col in enumerate(cols):
        assert counts[i].fault == 0
        assert_same(got[i], oracle.build(col), f"column {i}")
