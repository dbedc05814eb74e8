"""The benchmark's contract on the CPU: its byte model, its refusal of more
than one GPU and the HBM traffic figure it reads from the committed profile."""
import os
import subprocess
import sys

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
