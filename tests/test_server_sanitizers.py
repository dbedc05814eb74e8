"""The build server's host code (indexing_amd/csrc/wah_server.cpp) under
ThreadSanitizer and AddressSanitizer + UndefinedBehaviorSanitizer.

The server is compiled with g++ against a CPU test double of the HIP calls it
makes and of the library entry points it uses (tests/server_harness): streams
are worker threads, so copies, builds and event completions run
asynchronously as on a GPU.  tests/server_harness/stress.cpp submits ragged
and empty columns from several threads, completes them by wait, poll and a
second waiter on the same id, injects kernel-fault reports, starves the arena
so slots wait for memory, and destroys servers with requests queued; every
result is checked against its own column.  The reference builds its host
code under ASan by a configure switch (configure:175-176).
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HARNESS = os.path.join(ROOT, "tests", "server_harness")
SOURCES = [os.path.join(ROOT, "indexing_amd", "csrc", "wah_server.cpp"),
           os.path.join(HARNESS, "fake_device.cpp"),
           os.path.join(HARNESS, "stress.cpp")]


def _build(tmp_path, sanitize):
    exe = str(tmp_path / ("stress_" + sanitize.split(",")[0]))
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
           f"-fsanitize={sanitize}", "-fno-sanitize-recover=all", "-pthread",
           "-Wall", "-Wextra", "-Werror",
           "-I" + HARNESS, "-I" + os.path.join(ROOT, "include"),
           "-I" + os.path.join(ROOT, "indexing_amd", "csrc"), *SOURCES, "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    return exe


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not installed")
@pytest.mark.parametrize("sanitize", ["thread", "address,undefined"])
def test_server_loop_under_sanitizer(tmp_path, sanitize):
    exe = _build(tmp_path, sanitize)
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1 second_deadlock_stack=1",
               ASAN_OPTIONS="detect_leaks=1:halt_on_error=1")
    r = subprocess.run([exe, "100"], capture_output=True, text=True, timeout=600, env=env)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-6000:]
    assert "WARNING: ThreadSanitizer" not in out and "ERROR: AddressSanitizer" not in out, out[-6000:]
    assert "runtime error" not in out, out[-6000:]
    assert out.rstrip().endswith("ok") or "\nok\n" in out, out[-2000:]
    # the run exercised what it claims: batches of several columns, memory waits
    assert "memory waits: slots 4" in out
