"""The host copy that the staging paths use (csrc/host_copy.hpp): several
threads calling it at once, at sizes around its split threshold, copy every
byte and nothing outside the destination; built once plainly and once under
ThreadSanitizer, which checks the pool's hand-off of parts between threads."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "cellularautomatacore.jl_amd", "csrc")
SRC = os.path.join(HERE, "native", "host_copy_check.cpp")


def build(tmp_path, extra):
    exe = str(tmp_path / "host_copy_check")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-pthread", "-I", CSRC, SRC, "-o", exe] + extra
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
@pytest.mark.parametrize("extra", [[], ["-fsanitize=thread"]], ids=["plain", "tsan"])
def test_concurrent_copies(tmp_path, extra):
    exe = build(tmp_path, extra)
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([exe, "4", "3" if extra else "6"], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr[-4000:]
