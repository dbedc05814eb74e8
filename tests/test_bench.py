"""The benchmark's contract on the CPU: its byte model, its refusal of more
than one GPU and the HBM traffic figure it reads from the committed profile."""
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402

# 2^28 uniform [0, 65535], seed 1: the sizes the build reports for the column
N, k, L, K = 1 << 28, 268374080, 536621634, 65536


def test_byte_model_matches_canonical_c2_figure():
    """16nP + 8n + 4L + 12K at P = 2 is 40n + 4L + 12K, about 48 B/value
    (12.885e9 bytes for the column)."""
    bm = bench.byte_model(N, k, L, K, 2, 65535, 0)
    assert bm["b_alg"] == 40 * N + 4 * L + 12 * K
    assert bm["b_alg"] == pytest.approx(12.885e9, rel=1e-4)
    assert bm["alg_pass"] == [12.0 * N, 16.0 * N]   # read 4n, write 8n; then 8n + 8n
    # 2-byte keys: pass 1 writes 6n, pass 2 reads and writes 6n; the tail reads 6n
    assert bm["kb"] == 2.0
    assert bm["moved_pass"] == [10.0 * N, 12.0 * N]
    assert bm["b_moved"] == 4 * N + 22 * N + 6 * N + 4 * L + 12 * K


def test_byte_model_wide_keys_and_passes():
    """Keys of 2^16 and above move as 4 bytes: the moved bytes equal the
    canonical ones for pair mode."""
    for P in (1, 2, 3, 4):
        bm = bench.byte_model(N, k, L, K, P, 1 << 20, 0)
        assert bm["kb"] == 4.0
        assert bm["moved_pass"] == bm["alg_pass"]
        assert bm["b_moved"] == bm["b_alg"] == 16 * N * P + 8 * N + 4 * L + 12 * K
    empty = bench.byte_model(N, N, 2 * N, 1, 0, 7, 0)  # one value: no active pass
    assert empty["alg_pass"] == [] and empty["moved_pass"] == []


def test_byte_model_run_mode_counts_items():
    """Run mode sorts k items of 12 bytes, cut from a second read of the values."""
    kr, Lr, Kr = 8719340, 8781487, 10447
    bm = bench.byte_model(N, kr, Lr, Kr, 3, (1 << 20) - 1, 1)
    assert bm["moved_pass"] == [24.0 * kr, 24.0 * kr, 24.0 * kr]
    assert bm["b_moved"] == 8 * N + 12 * kr + 72 * kr + 12 * kr + 4 * Lr + 12 * Kr
    assert bm["b_alg"] == 16 * N * 3 + 8 * N + 4 * Lr + 12 * Kr


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_traffic_profile_parses():
    """The roofline's `traffic` comes from the committed counter summary, whose
    commands run bench.py."""
    t = bench.read_traffic()
    assert t is not None and 1e9 < t < 1e10
    text = open(bench.TRAFFIC_FILE).read()
    assert "--pmc FETCH_SIZE" in text and "--pmc WRITE_SIZE" in text
    assert "python3 bench.py" in text


class _FakeServer:
    """Stand-in for ia.Server with the same submit / wait / stats surface:
    a worker thread completes requests in order after a short delay, writing
    a result that depends only on the column."""

    def __init__(self):
        import queue
        import threading
        self.q = queue.Queue()
        self.n = 0
        self.th = threading.Thread(target=self._work, daemon=True)
        self.th.start()

    def _work(self):
        import time
        while True:
            p = self.q.get()
            time.sleep(0.0005)
            k, o, ln, w = p.bufs
            K = min(3, k.size)
            k[:K] = p.values[:K]
            o[:K] = 1
            ln[:K] = 2
            w[:2 * K] = p.values[0] if p.values.size else 0
            p.res.counts.keycnt, p.res.counts.length = K, 2 * K
            p.done.set()

    def submit(self, values, out):
        import threading
        import types
        p = types.SimpleNamespace(values=values, bufs=out, done=threading.Event(),
                                  res=types.SimpleNamespace(counts=types.SimpleNamespace(
                                      keycnt=0, length=0)))
        self.n += 1
        self.q.put(p)
        return p

    def wait(self, p):
        p.done.wait()

    def reset_stats(self):
        pass

    def stats(self):
        return {"p50_us": 0.0, "p99_us": 0.0, "values": 1, "busy_s": 1.0, "requests": self.n,
                "batches": self.n, "max_batch_cols": 1, "memory_waits": 0, "arena_peak": 0}


def test_serve_loop_retires_through_a_full_ring():
    """The open-loop submitter waits for result-ring space that only the
    retire thread frees: with a ring that holds a few requests, every request
    still completes and every result matches its column's recorded one."""
    import collections
    import concurrent.futures as cf
    import threading
    import zlib

    rng = np.random.default_rng(5)
    layout, cols = [], []
    for i in range(16):
        n = int(rng.integers(1, 400))
        cols.append(rng.integers(0, 1000, n).astype(np.uint32))
        layout.append((0, n))

    class Ring(bench._ResultRing):
        def __init__(self, words):
            self.buf = np.zeros(words, np.uint32)
            self.cap = words
            self.live = collections.deque()
            self.cv = threading.Condition()

    ring = Ring(5 * 400 + 4 + 100)   # room for about one large request
    req_col = [i % 16 for i in range(300)]
    alone = [None] * 16

    def checksum(k, o, ln, w):
        return zlib.crc32(k.tobytes() + o.tobytes() + ln.tobytes() + w.tobytes())

    with cf.ThreadPoolExecutor(4) as pool:
        srv = _FakeServer()
        r1 = bench.serve_run(srv, np.zeros(300), req_col, cols, layout, ring, pool, checksum,
                             alone, True)
        assert r1["mismatches"] == 0 and all(a is not None for a in alone)
        arrivals = np.cumsum(rng.exponential(0.0002, 300))
        r2 = bench.serve_run(_FakeServer(), arrivals, req_col, cols, layout, ring, pool,
                             checksum, alone, False)
        assert r2["mismatches"] == 0
    assert not ring.live
