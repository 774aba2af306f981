"""The host copy-thread pool (fastddm_amd/csrc/fddm_copy_pool.hpp) under
concurrent callers: built on its own with g++ and ThreadSanitizer, every
copy must be exact, the bytes around it untouched, and no data race
reported."""

import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "fastddm_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
@pytest.mark.parametrize("sanitize", [False, True])
def test_copy_pool_stress(tmp_path, sanitize):
    exe = str(tmp_path / "stress")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-pthread", "-I", CSRC,
           os.path.join(HERE, "copy_pool_stress.cpp"), "-o", exe]
    if sanitize:
        cmd[1:1] = ["-fsanitize=thread"]
    subprocess.run(cmd, check=True, capture_output=True, text=True, timeout=180)
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1")
    res = subprocess.run([exe, "60" if sanitize else "300"], capture_output=True, text=True, timeout=300, env=env)
    assert res.returncode == 0, res.stdout + res.stderr[-3000:]
    assert "bad=0" in res.stdout
    assert "ThreadSanitizer" not in res.stderr, res.stderr[-3000:]
