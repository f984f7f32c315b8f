"""Certificates drop-in (src/generate_certificates.py) on a small results/
tree.  Parity unpinned: the reference ships no certificate fixture and its
script needs numba; the hashed record format is restated from
scripts/generate_certificates.py:105-108, which agrees with the format the
reference's README states (README.md:277-278), and the file layout from
:196-233.
What is checked here: rows in file order, verified flags taken from the
details file, top-20 selection by a stable merit sort, file names."""
import csv
import hashlib
import os
import shutil
import subprocess
import sys

import pytest

from conftest import PKG

ROWS = [  # maximal gaps (p_after, gap) with merit to 6 decimals, and two ties
    (1294268779, 288, 13.718236), (1453168433, 292, 13.855643), (2300942869, 320, 14.901488),
    (3842611109, 336, 15.285416), (4302407713, 354, 16.072053), (4302407713, 354, 16.072053),
    (1000000007, 6, 0.289530),
]


def make_tree(tmp_path, details):
    scripts = tmp_path / "scripts"
    results = tmp_path / "results"
    scripts.mkdir()
    results.mkdir()
    shutil.copy(os.path.join(PKG, "src", "generate_certificates.py"), scripts)
    with open(results / "large_gaps.csv", "w") as f:
        f.write("position,gap,merit\n")
        for p, g, m in ROWS:
            f.write(f"{p},{g},{m:.6f}\n")
    if details is not None:
        with open(results / "verification_details.csv", "w") as f:
            f.write("position,gap,merit,verified\n")
            for (p, g, m), ok in zip(ROWS, details):
                f.write(f"{p},{g},{m:.6f},{int(ok)}\n")
    return scripts / "generate_certificates.py"


def run(script, env=None):
    return subprocess.run([sys.executable, str(script)], capture_output=True, text=True,
                          env=env, check=True)


def test_certificates_from_details(tmp_path):
    flags = [True, True, False, True, True, True, False]
    script = make_tree(tmp_path, flags)
    run(script)
    with open(tmp_path / "results" / "certificates.csv") as f:
        rows = list(csv.DictReader(f))
    assert [r["p_after"] for r in rows] == [str(p) for p, _, _ in ROWS]
    for r, (p, g, m), ok in zip(rows, ROWS, flags):
        assert (int(r["p_before"]), int(r["gap"]), r["merit"]) == (p - g, g, f"{m:.6f}")
        assert r["verified"] == ("YES" if ok else "NO")
        text = f"VOSS|{p - g}|{p}|{g}|{m:.6f}"
        assert r["sha256"] == hashlib.sha256(text.encode()).hexdigest()
    names = sorted(os.listdir(tmp_path / "results" / "cert_top20"))
    assert len(names) == len(ROWS)
    assert names[0] == "cert_01_merit16.07_p4302407713.txt"
    assert names[-1] == "cert_07_merit0.29_p1000000007.txt"
    body = (tmp_path / "results" / "cert_top20" / names[0]).read_text(encoding="utf-8")
    assert "Prime Before (p)        : 4,302,407,359" in body
    assert "All 353 integers in (p, p+g) are proven composite" in body


def test_certificates_missing_input(tmp_path):
    (tmp_path / "scripts").mkdir()
    shutil.copy(os.path.join(PKG, "src", "generate_certificates.py"), tmp_path / "scripts")
    r = subprocess.run([sys.executable, str(tmp_path / "scripts" / "generate_certificates.py")],
                       capture_output=True, text=True)
    assert r.returncode == 1 and "not found" in r.stdout


@pytest.mark.gpu
def test_certificates_verify_on_gpu(tmp_path):
    """No details file: every gap goes through this build's verify_cpp (GPU);
    the last row (a 6-gap that skips primes) fails."""
    script = make_tree(tmp_path, None)
    env = dict(os.environ, VOSS_AMD_HOME=os.path.dirname(PKG))
    print("[gpu] certificates via verify_cpp", flush=True)
    run(script, env)
    with open(tmp_path / "results" / "certificates.csv") as f:
        flags = [r["verified"] for r in csv.DictReader(f)]
    assert flags == ["YES"] * 6 + ["NO"]
