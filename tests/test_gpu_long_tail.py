"""Single columns around the length where the build switches to its long tail
tiles (16384 items per tile instead of 8192), against the oracle
(src/cpu.cpp:313-328): just below and at the switch, ragged last tiles,
32-bit keys with a nonzero rid base, few keys (long fills), and a run-mode
column, whose run tail takes the same long tiles.
"""
import numpy as np
import pytest

import indexing_amd as ia
from oracle import oracle
from tests.test_gpu_parity import assert_same

pytestmark = pytest.mark.gpu

SWITCH = 1 << 24  # the first length with long tail tiles


CASES = ["u16_below", "u16_at", "u32_ragged_rid_base", "three_keys", "zipf_runs"]


def make_case(name):
    """(values, rid_base) of a case."""
    rng = np.random.default_rng(24 + CASES.index(name))
    if name == "u16_below":
        return rng.integers(0, 65536, SWITCH - 1).astype(np.uint32), 0
    if name == "u16_at":
        return rng.integers(0, 65536, SWITCH).astype(np.uint32), 0
    if name == "u32_ragged_rid_base":
        return rng.integers(0, 1 << 32, SWITCH + 9000, dtype=np.uint64).astype(np.uint32), 12345
    if name == "three_keys":
        return rng.integers(0, 3, SWITCH + 16383).astype(np.uint32), 0
    return ia.generate(SWITCH + 777, "zipf-runs", seed=5, cardinality=1 << 20, zipf_s=1.2,
                       run_mean=512.0), 0


@pytest.mark.parametrize("name", CASES)
def test_long_tail_switch_matches_oracle(ctx, name):
    v, rid_base = make_case(name)
    print(f"[long-tail] {name} n={v.size} rid_base={rid_base}", flush=True)
    got = ctx.build(v, rid_base=rid_base)
    ref = oracle.build(v, rid_base=rid_base, threads=8)
    assert_same(got, ref, name)
    if name == "zipf_runs":
        assert got.mode == 1  # the run tail took the long tiles
