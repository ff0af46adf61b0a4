"""The host-thread copy pool used by the staged host <-> device copies
(cudadecon_amd/csrc/copypool.hpp), built on the CPU with ThreadSanitizer."""

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not installed")
def test_copy_pool_copies_every_byte_without_races(tmp_path):
    exe = str(tmp_path / "copypool_check")
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=thread", "-pthread",
                           "-I", os.path.join(ROOT, "cudadecon_amd", "csrc"),
                           os.path.join(ROOT, "tests", "native", "copypool_check.cpp"), "-o", exe])
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1 exitcode=66")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "WARNING: ThreadSanitizer" not in r.stderr, r.stderr
    assert r.stdout.count(": ok") == 2, r.stdout
