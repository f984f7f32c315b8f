"""The device loop's arena allocator (csrc/range_alloc.h): host-only check
compiled with g++ against a brute-force occupancy map, no GPU involved."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "voss-prime-gaps_v1_amd", "csrc")


def test_range_alloc(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    exe = str(tmp_path / "range_alloc_check")
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-I", CSRC,
                        os.path.join(HERE, "native", "range_alloc_check.cpp"), "-o", exe],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr[-2000:]
    assert out.stdout.strip() == "ok"
