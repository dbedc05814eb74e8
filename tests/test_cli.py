"""The command-line surface: `generate` (src/generate_test_data.cpp), and the
GPU programs `phases` (src/phases.cpp) and `gpu` (src/gpu.cpp) against the
oracle's `cpu` program (src/cpu.cpp) on the same value files."""
import os
import re
import subprocess

import numpy as np
import pytest

import indexing_amd as ia
from oracle import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "indexing_amd", "build")


def run(*args, **kw):
    return subprocess.run(list(args), stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                          text=True, timeout=kw.get("timeout", 300), check=kw.get("check", True))


# ---- generate ------------------------------------------------------------------

def test_generate_text_format_and_seed(tmp_path):
    out = run(os.path.join(BIN, "generate"), "-a", "1000", "--seed", "42").stdout
    assert out.endswith("\n") and out.count("\n") == 1000  # "\n" separators + final endl
    vals = np.array(out.split(), dtype=np.uint64)
    assert vals.size == 1000 and vals.max() <= 65535
    assert np.array_equal(vals.astype(np.uint32), ia.generate(1000, seed=42, lo=0, hi=65535))


def test_generate_flags_and_separator(tmp_path):
    out = run(os.path.join(BIN, "generate"), "-f", "5", "-t", "9", "-a", "50", "-s", " ",
              "--seed", "7").stdout
    assert "\n" not in out[:-1] and out.endswith("\n")
    vals = [int(x) for x in out.split(" ")]
    assert len(vals) == 50 and min(vals) >= 5 and max(vals) <= 9


def test_generate_zero_amount_prints_newline():
    assert run(os.path.join(BIN, "generate"), "-a", "0").stdout == "\n"


def test_generate_unseeded_differs():
    a = run(os.path.join(BIN, "generate"), "-a", "64").stdout
    b = run(os.path.join(BIN, "generate"), "-a", "64").stdout
    assert a != b  # seeded from random_device, as the reference


def test_generate_binary_and_file_output(tmp_path):
    p = tmp_path / "v.bin"
    run(os.path.join(BIN, "generate"), "-a", "300", "--seed", "3", "--binary", "-o", str(p))
    assert np.array_equal(np.fromfile(p, dtype="<u4"), ia.generate(300, seed=3))
    q = tmp_path / "z.txt"
    run(os.path.join(BIN, "generate"), "-a", "5000", "--seed", "3", "--dist", "zipf-runs",
        "--cardinality", "1000", "--run-mean", "50", "-o", str(q))
    assert np.array_equal(ia.read_values_text(q),
                          ia.generate(5000, "zipf-runs", seed=3, cardinality=1000,
                                      run_mean=50))


def test_generated_text_reads_back_like_ifstream(tmp_path):
    p = tmp_path / "v.txt"
    run(os.path.join(BIN, "generate"), "-a", "2000", "--seed", "11", "-o", str(p))
    assert np.array_equal(ia.read_values_text(p), ia.generate(2000, seed=11))


def test_phases_requires_a_file():
    out = run(os.path.join(BIN, "phases")).stdout
    assert out.startswith("A filename with test data is required")


def test_help_flags():
    for prog in ("phases", "gpu", "generate", "scan", "sc", "segmented_scan"):
        out = run(os.path.join(BIN, prog), "--help").stdout
        assert "--" in out


# ---- GPU programs against the oracle's cpu program -------------------------------

def value_file(tmp_path, values, name="v.txt"):
    p = tmp_path / name
    p.write_text("\n".join(str(int(v)) for v in values) + "\n")
    return str(p)


def cpu_print(path):
    out = run(oracle.CPU_BIN, "-f", path, "-p").stdout.splitlines()
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("n,card", [(1, 1), (32, 50), (1000, 7), (5000, 300), (20000, 64)])
def test_phases_print_matches_cpu(tmp_path, n, card):
    values = np.random.default_rng(n).integers(0, card, n).astype(np.uint32)
    path = value_file(tmp_path, values)
    out = run(os.path.join(BIN, "phases"), "-f", path, "-p").stdout.splitlines()
    timings = out[:9]
    assert all(re.fullmatch(r"\d+ us", line) for line in timings), timings
    ref = cpu_print(path)[4:]  # after the four header lines of cpu
    assert out[9:] == ref
    fused = run(os.path.join(BIN, "phases"), "-f", path, "-p", "--fused").stdout.splitlines()
    assert fused[9:] == ref


@pytest.mark.gpu
@pytest.mark.parametrize("batch,wg,jobs", [(1023, 1, 1), (100, 3, 2), (31, 2, 4), (4096, 8, 1),
                                           (17, 1, 0)])
def test_gpu_program_stitches_batches_exactly(tmp_path, batch, wg, jobs):
    values = np.random.default_rng(batch).integers(0, 40, 3000).astype(np.uint32)
    path = value_file(tmp_path, values)
    out = run(os.path.join(BIN, "gpu"), "-f", path, "-p", "--batch-size", str(batch),
              "-w", str(wg), "-j", str(jobs)).stdout.splitlines()
    assert out[0] == f"Reading data from '{path}' ... '3000' values."
    assert out[1] == f"Maximum value is '{values.max()}'."
    assert out[2].startswith("Using device named '")
    assert re.fullmatch(r"Time: '\d+' ms", out[-1])
    lines = out[3:-1]
    # the merger prints every key 0..bound with its bitmap (src/gpu.cpp:153-157)
    ref = {}
    for line in cpu_print(path)[4:]:
        k, bits = line.split("\t")
        ref[int(k)] = bits
    assert len(lines) == int(values.max()) + 1
    for line in lines:
        k, bits = line.split("\t")
        assert bits == ref.get(int(k), ""), f"key {k}"


@pytest.mark.gpu
def test_gpu_program_builtin_sample():
    out = run(os.path.join(BIN, "gpu"), "-p").stdout.splitlines()
    assert out[0] == "'32' values." and out[1] == "Maximum value is '42'."
    ref = run(oracle.CPU_BIN, "-p").stdout.splitlines()[4:]
    got = out[3:-1]
    refd = dict(line.split("\t") for line in ref)
    assert len(got) == 43
    for line in got:
        k, bits = line.split("\t")
        assert bits == refd.get(k, "")


# ---- primitive checks (src/scan.cpp, src/sc.cpp, src/segmented_scan.cpp) -------

@pytest.mark.gpu
@pytest.mark.parametrize("size", [1, 5000, 1048577 + 33])
def test_primitive_programs_self_check(size):
    out = run(os.path.join(BIN, "scan"), "--seed", "3", "--size", str(size)).stdout.splitlines()
    assert out == [f"Using {size} values.", "Success!"]
    out = run(os.path.join(BIN, "sc"), "--seed", "4", "--size", str(size), "-F", "0.3").stdout
    assert out.splitlines() == [f"Compacting {size} values.", "SUCCESS"]
    out = run(os.path.join(BIN, "segmented_scan"), "--seed", "5", "--size", str(size)).stdout
    assert out.splitlines() == [f"Using {size} values.", "Success"]


@pytest.mark.gpu
def test_primitive_programs_on_a_file(tmp_path):
    values = np.random.default_rng(9).integers(0, 1 << 32, 70001, dtype=np.uint64)
    path = value_file(tmp_path, values.astype(np.uint32))
    assert run(os.path.join(BIN, "scan"), "-f", path).stdout.splitlines() == ["Success!"]
    assert run(os.path.join(BIN, "sc"), "-f", path, "--seed", "1").stdout.splitlines() == ["SUCCESS"]
    out = run(os.path.join(BIN, "segmented_scan"), "-f", path, "--seed", "2", "-p").stdout
    lines = out.splitlines()
    assert lines[0] == "Success" and len(lines) == 1 + values.size
