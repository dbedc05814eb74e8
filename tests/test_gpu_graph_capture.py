"""A caller captures the device-pointer builds into its own graph.

wahgpu_build_device returns nothing to the host, and
wahgpu_build_device_noalloc allocates nothing: both can be captured (torch's
CUDAGraph on a stream the context is bound to beforehand; the first, eager
build sizes the arena, so the captured call allocates nothing either).  The
graph is replayed over different inputs written into the captured input
buffer, with eager work and synchronizations between the replays, and every
replay is compared with the oracle (src/cpu.cpp:313-328).
"""
import numpy as np
import pytest

import indexing_amd as ia
from oracle import oracle
from tests.test_gpu_batch import per_column
from tests.test_gpu_parity import assert_same, dev

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("noalloc", [False, True], ids=["build_device", "noalloc"])
def test_build_captured_and_replayed(noalloc):
    import torch
    n = (1 << 20) + 77
    inputs = [ia.generate(n, "uniform", seed=90 + i, lo=0, hi=(1 << (8 + 3 * i)) - 1)
              for i in range(3)]
    c = ia.Context(0)
    s = torch.cuda.Stream()
    build = c.build_device_noalloc if noalloc else c.build_device
    try:
        d_in = dev(inputs[0])
        with torch.cuda.stream(s):
            out = build(d_in)     # binds the context to s and sizes its arena
            build(d_in, out=out)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        print(f"[capture] noalloc={noalloc}", flush=True)
        with torch.cuda.graph(g, stream=s):
            build(d_in, out=out)
        for rep, v in enumerate(inputs):
            d_in.copy_(dev(v))
            x = torch.ones(1 << 20, device="cuda:0").sum()  # eager work between replays
            g.replay()
            torch.cuda.synchronize()
            assert float(x) == float(1 << 20)
            cnt = c.read_counts(out[4])
            assert cnt.fault == 0
            got = per_column(out, [cnt], [(0, n)])[0]
            assert_same(got, oracle.build(v), f"replay {rep}")
    finally:
        c.close()
