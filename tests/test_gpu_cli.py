"""The reference-compatible programs on the GPU: voss_v7_live writes the
reference's files with the oracle's numbers, and verify_cpp produces the
same details CSV as the reference's own verifier (built from
src/verify_cpp.cpp into oracle/_ref)."""
import csv
import os
import subprocess

import numpy as np
import pytest

from conftest import PKG, ROOT

pytestmark = pytest.mark.gpu

SIEVE = os.path.join(PKG, "bin", "voss_v7_live")
VERIFY = os.path.join(PKG, "bin", "verify_cpp")
REF_VERIFY = os.path.join(ROOT, "oracle", "_ref", "verify_cpp")


def rows(path):
    with open(path) as f:
        return list(csv.reader(f))


@pytest.fixture(scope="module")
def run_1e9(tmp_path_factory):
    d = tmp_path_factory.mktemp("results")
    print("[gpu] voss_v7_live 1e9 300000000", flush=True)
    r = subprocess.run([SIEVE, "1000000000", "300000000"], cwd=d, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr
    return d, r.stdout


def test_stdout_protocol(run_1e9):
    _, out = run_1e9
    assert "Total primes: 50847534" in out
    assert out.count("-> PASS") == 3
    for f in ("gap_histogram.csv", "hl_trend.csv", "modq_counts.csv", "modq_gap_hist.csv",
              "chebyshev.csv", "gap_sequence.csv", "large_gaps.csv"):
        assert f"OK {f}" in out


def test_files_match_oracle(run_1e9, oracle):
    d, _ = run_1e9
    ref = oracle.run_plain(10**9, 300000000)
    g = rows(d / "gap_histogram.csv")
    assert g[0] == ["gap", "count"]
    assert {int(a): int(b) for a, b in g[1:]} == {i: int(c) for i, c in enumerate(ref["gaps"]) if c}
    m = rows(d / "modq_counts.csv")
    assert m[0] == ["q", "a", "count"] and len(m) == 1 + 6 + 30
    assert [int(r[2]) for r in m[1:7]] == [int(x) for x in ref["modq"][0:6]]
    assert [int(r[2]) for r in m[7:]] == [int(x) for x in ref["modq"][6:36]]
    mg = rows(d / "modq_gap_hist.csv")
    assert mg[0] == ["q", "residue", "gap", "count"]
    res6, res30 = [1, 5], [1, 7, 11, 13, 17, 19, 23, 29]
    exp = []
    for row in range(10):
        q, res = (6, res6[row]) if row < 2 else (30, res30[row - 2])
        for gap in range(1, 100):
            c = int(ref["modq_gap"][row][gap])
            if c:
                exp.append([str(q), str(res), str(gap), str(c)])
    assert mg[1:] == exp
    ch = rows(d / "chebyshev.csv")
    assert ch == [["N", "pi_4_1", "pi_4_3", "difference"],
                  ["1000000000", str(ref["c1"] + 1), str(ref["c3"] + 1),
                   str(ref["c3"] - ref["c1"])]]
    hl = rows(d / "hl_trend.csv")
    assert len(hl) == 6 and hl[1] == ["1000000000", "1.778300"]
    lg = rows(d / "large_gaps.csv")
    assert lg[0] == ["position", "gap", "merit"]
    assert lg[1:] == [[str(p), str(g), f"{m:.6f}"] for p, g, m in ref["large"]]


@pytest.mark.skipif(not os.path.exists(REF_VERIFY), reason="oracle/_ref/verify_cpp not built")
def test_verify_cli_matches_reference_verifier(run_1e9, tmp_path):
    d, _ = run_1e9
    inp = tmp_path / "in.csv"
    with open(d / "large_gaps.csv") as f:
        lines = f.read().splitlines()
    # add rows that must fail: shifted ends and a gap skipping a prime
    extra = []
    for line in lines[1:30]:
        p, g, m = line.split(",")
        extra.append(f"{int(p) + 2},{int(g) + 2},{m}")
        extra.append(f"{p},{int(g) - 2},{m}")
    inp.write_text("\n".join(lines + extra) + "\n")
    outs = {}
    for name, exe in (("gpu", VERIFY), ("ref", REF_VERIFY)):
        print(f"[gpu] verify_cpp ({name})", flush=True)
        r = subprocess.run([exe, str(inp), str(tmp_path / f"{name}.csv"),
                            str(tmp_path / f"{name}_cache.csv")], capture_output=True,
                           text=True, timeout=180, env=dict(os.environ, OMP_NUM_THREADS="8"))
        assert r.returncode == 1, r.stdout[-1500:] + r.stderr  # the added rows fail
        outs[name] = (open(tmp_path / f"{name}.csv").read(),
                      open(tmp_path / f"{name}_cache.csv").read())
    assert outs["gpu"][0] == outs["ref"][0]
    assert sorted(outs["gpu"][1].splitlines()) == sorted(outs["ref"][1].splitlines())


def test_gap_sequence_file(run_1e9, oracle):
    d, _ = run_1e9
    ref = oracle.run_plain(10**9, 300000000, sample=True)
    rws = rows(d / "gap_sequence.csv")
    assert rws[0] == ["index", "gap", "p"]
    idx = np.nonzero(ref["gap_seq"] > 0)[0]
    assert len(rws) - 1 == len(idx)
    for k in (0, len(idx) // 2, len(idx) - 1):
        i = idx[k]
        assert rws[1 + k] == [str(i), str(ref["gap_seq"][i]), str(ref["pos_seq"][i])]


def test_stop_file_and_resume(run_1e9, tmp_path):
    """stop.txt ends the run cleanly after the next checkpoint (segment 0);
    the rerun resumes from checkpoint.bin and produces the same outputs as
    the uninterrupted run (src/voss_v7_live.cu:599-621, :407-421)."""
    full_dir, _ = run_1e9
    d = tmp_path / "resume"
    d.mkdir()
    (d / "stop.txt").write_text("")
    print("[gpu] voss_v7_live with stop.txt", flush=True)
    r = subprocess.run([SIEVE, "1000000000", "300000000"], cwd=d, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0 and "STOP-FILE detected at seg 0" in r.stdout
    assert (d / "checkpoint.bin").exists() and not (d / "stop.txt").exists()
    assert not (d / "gap_histogram.csv").exists()
    print("[gpu] voss_v7_live resume", flush=True)
    r = subprocess.run([SIEVE, "1000000000", "300000000"], cwd=d, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0 and "RESUMING FROM SEGMENT 1 / 4" in r.stdout
    assert "Total primes: 50847534" in r.stdout and not (d / "checkpoint.bin").exists()
    for f in ("gap_histogram.csv", "modq_counts.csv", "modq_gap_hist.csv", "chebyshev.csv",
              "large_gaps.csv"):
        assert (d / f).read_text() == (full_dir / f).read_text(), f


def test_resume_from_reference_layout_checkpoint(tmp_path):
    """A checkpoint in the reference's 800,064-byte layout (no extension):
    pi and the gap histogram are restored exactly; like the reference, the
    mod-q state of the earlier segments is not."""
    d = tmp_path / "ref_ckpt"
    d.mkdir()
    (d / "stop.txt").write_text("")
    subprocess.run([SIEVE, "100000000", "30000000"], cwd=d, capture_output=True, text=True,
                   timeout=120, check=True)
    raw = (d / "checkpoint.bin").read_bytes()
    (d / "checkpoint.bin").write_bytes(raw[:800064])
    r = subprocess.run([SIEVE, "100000000", "30000000"], cwd=d, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0 and "RESUMING FROM SEGMENT 1 / 4" in r.stdout
    assert "Total primes: 5761455" in r.stdout
    assert r.stdout.count("-> PASS") >= 2   # checks 1 and 2; check 3 needs the full mod-4 counts
