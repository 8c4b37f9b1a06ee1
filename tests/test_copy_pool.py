"""The host staging path's copy pool (onefft_amd/csrc/copy_pool.hpp), on the CPU: built with g++
from tests/copy_pool_check.cpp (with ThreadSanitizer when the toolchain has it) and run; several
threads share pools of 0..8 helpers, and every copy must be exact with its guard bytes intact."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

SRC = os.path.join(ROOT, "tests", "copy_pool_check.cpp")


def build_and_run(tmp_path, extra, copies):
    exe = tmp_path / "copy_pool_check"
    cmd = ["g++", "-std=c++17", "-O2", "-pthread", *extra, SRC, "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        return None, r.stderr
    run = subprocess.run([str(exe), str(copies)], capture_output=True, text=True, timeout=300)
    return run, None


@pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
def test_copy_pool_exact_under_concurrency(tmp_path):
    run, err = build_and_run(tmp_path, [], 40)
    assert run is not None, err
    assert run.returncode == 0, run.stdout + run.stderr
    assert run.stdout.count("ok") == 4


@pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
def test_copy_pool_thread_sanitizer(tmp_path):
    run, err = build_and_run(tmp_path, ["-fsanitize=thread", "-g"], 8)
    if run is None:
        pytest.skip(f"ThreadSanitizer build unavailable: {err.splitlines()[-1] if err else ''}")
    if "FATAL: ThreadSanitizer" in run.stderr or "unexpected memory mapping" in run.stderr:
        pytest.skip("ThreadSanitizer cannot run in this environment")
    assert run.returncode == 0 and "WARNING: ThreadSanitizer" not in run.stderr, run.stdout + run.stderr[-4000:]
