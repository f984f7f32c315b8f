"""The process boundary the reference pipeline depends on (run_voss.py,
src/voss_master.py, bin/verify_cpp), exercised on the CPU with stand-in
programs where a GPU would be needed."""
import importlib.util
import os
import re
import shutil
import stat
import subprocess
import sys

import pytest

from conftest import GOLDEN, PKG

DRIVER = os.path.join(PKG, "src", "voss_master.py")


def update_master(path, n_value, seg_num):
    """run_voss.py:111-123 restated: first-match regex rewrite."""
    src = open(path).read()
    src = re.sub(r"(N_VALUE\s*=\s*)\d+", r"\g<1>" + str(n_value), src, count=1)
    src = re.sub(r"(SEG_NUM\s*=\s*)\d+", r"\g<1>" + str(seg_num), src, count=1)
    open(path, "w").write(src)


def load(path):
    spec = importlib.util.spec_from_file_location("vm", path)
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def make_tree(tmp_path, stub_body):
    root = tmp_path / "ref"
    (root / "src").mkdir(parents=True)
    (root / "bin").mkdir()
    (root / "data").mkdir()
    shutil.copy(DRIVER, root / "src" / "voss_master.py")
    shutil.copy(os.path.join(GOLDEN, "reference", "verification_db.json"),
                root / "data" / "verification_db.json")
    stub = root / "bin" / "voss_v7_live"
    stub.write_text("#!/bin/sh\n" + stub_body)
    stub.chmod(stub.stat().st_mode | stat.S_IXUSR)
    return root


def test_regex_patch_hits_the_assignments(tmp_path):
    p = tmp_path / "voss_master.py"
    shutil.copy(DRIVER, p)
    update_master(p, 1234567890, 300000000)
    m = load(str(p))
    assert m.N_VALUE == 1234567890 and m.SEG_NUM == 300000000


def test_driver_runs_binary_in_results_and_checks_oeis(tmp_path):
    root = make_tree(tmp_path, 'echo "args $1 $2 cwd $(pwd)"; echo "Total primes: 4"\n')
    update_master(root / "src" / "voss_master.py", 10, 30)
    r = subprocess.run([sys.executable, str(root / "src" / "voss_master.py")],
                       capture_output=True, text=True, env=dict(os.environ, VOSS_AMD_HOME=""))
    assert r.returncode == 0, r.stdout + r.stderr
    assert "args 10 30" in r.stdout
    assert f"cwd {root / 'results'}" in r.stdout
    assert "PASS" in r.stdout


def test_driver_reports_fail_and_propagates_exit_code(tmp_path):
    root = make_tree(tmp_path, 'echo "Total primes: 5"; exit 3\n')
    update_master(root / "src" / "voss_master.py", 10, 30)
    r = subprocess.run([sys.executable, str(root / "src" / "voss_master.py")],
                       capture_output=True, text=True, env=dict(os.environ, VOSS_AMD_HOME=""))
    assert "FAIL" in r.stdout
    assert r.returncode == 3


def test_driver_missing_binary(tmp_path):
    root = make_tree(tmp_path, "")
    os.remove(root / "bin" / "voss_v7_live")
    r = subprocess.run([sys.executable, str(root / "src" / "voss_master.py")],
                       capture_output=True, text=True, env=dict(os.environ, VOSS_AMD_HOME=""))
    assert r.returncode == 1 and "is not built" in r.stdout


VERIFY = os.path.join(PKG, "bin", "verify_cpp")


@pytest.mark.skipif(not os.path.exists(VERIFY), reason="CLI not built")
def test_verify_cli_input_errors(tmp_path):
    r = subprocess.run([VERIFY, str(tmp_path / "nope.csv"), str(tmp_path / "o.csv"),
                        str(tmp_path / "c.csv")], capture_output=True, text=True)
    assert r.returncode == 1 and "cannot open" in r.stderr
    empty = tmp_path / "e.csv"
    empty.write_text("position,gap,merit\n")
    r = subprocess.run([VERIFY, str(empty), str(tmp_path / "o.csv"), str(tmp_path / "c.csv")],
                       capture_output=True, text=True)
    assert r.returncode == 1 and "0 gap rows" in r.stdout
    bad = tmp_path / "b.csv"
    bad.write_text("position,gap,merit\nx12,4,0.5\n")
    r = subprocess.run([VERIFY, str(bad), str(tmp_path / "o.csv"), str(tmp_path / "c.csv")],
                       capture_output=True, text=True)
    assert r.returncode == 1 and "not a gap row" in r.stderr


@pytest.mark.skipif(not os.path.exists(VERIFY), reason="CLI not built")
def test_verify_cli_all_cached(tmp_path):
    """Rows the cache holds are verified without a device: the details file
    lists every input row in order with verified = 1, merits as %.6f, the
    ignored lines (empty, no comma) are skipped, the cache gains nothing,
    and the exit code is 0."""
    inp = tmp_path / "in.csv"
    inp.write_text("position,gap,merit\n1327,34,4.7234567\n\nno comma here\n9551,36\n")
    cache = tmp_path / "cache.csv"
    cache.write_text("p_after,gap\n9551,36\n1327,34\n")
    det = tmp_path / "det.csv"
    r = subprocess.run([VERIFY, str(inp), str(det), str(cache)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert det.read_text() == ("position,gap,merit,verified\n1327,34,4.723457,1\n"
                               "9551,36,0.000000,1\n")
    assert cache.read_text() == "p_after,gap\n9551,36\n1327,34\n"


@pytest.mark.skipif(not os.path.exists(os.path.join(PKG, "bin", "voss_v7_live")),
                    reason="CLI not built")
def test_sieve_cli_usage():
    r = subprocess.run([os.path.join(PKG, "bin", "voss_v7_live")], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr
    r = subprocess.run([os.path.join(PKG, "bin", "voss_v7_live"), "1000", "31"],
                       capture_output=True, text=True)
    assert r.returncode == 2


def test_bench_mr_work_count_matches_oracle(oracle):
    """bench.py's restated Miller-Rabin work count equals the oracle's."""
    import random
    import bench
    rng = random.Random(0x564F5353 + 31)
    ns = [rng.randrange(2**40, 2**64) | 1 for _ in range(300)]
    ns += [2**61 - 1, 3825123056546413051, 15, 2, 37, 1, 0, 341550071728321]
    for n in ns:
        assert bench.mr_mulmods(n) == oracle.mr_mulmods(n), n
