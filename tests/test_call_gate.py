"""The process-wide call gate (csrc/call_gate.h) that makes a second
thread's library call return VOSS_EBUSY instead of running beside another
thread's call: host-only checks compiled with g++ (ThreadSanitizer where
the toolchain has it), no GPU involved."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "voss-prime-gaps_v1_amd", "csrc")


@pytest.mark.parametrize("sanitize", [False, True])
def test_call_gate(tmp_path, sanitize):
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    exe = str(tmp_path / "call_gate_check")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-pthread", "-I", CSRC,
           os.path.join(HERE, "native", "call_gate_check.cpp"), "-o", exe]
    if sanitize:
        cmd.insert(1, "-fsanitize=thread")
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        if sanitize:
            pytest.skip("ThreadSanitizer unavailable: " + r.stderr[-200:])
        raise AssertionError(r.stderr)
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1")
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120, env=env)
    assert out.returncode == 0, out.stderr[-2000:]
    assert out.stdout.strip() == "ok"
