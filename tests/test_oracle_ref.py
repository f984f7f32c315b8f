"""Cross-checks the oracle's Miller-Rabin and verify_gap restatement against
the reference's own CPU verifier, compiled from src/verify_cpp.cpp by
oracle/Makefile into oracle/_ref/verify_cpp (skipped where it is absent).

The reference binary is driven through its own CLI: rows "n,0,0" make
verify_gap(n, 0) = is_prime_u64(n) (src/verify_cpp.cpp:59-63), so the
`verified` column of its details CSV is the reference's MR verdict.
"""
import csv
import os
import random
import subprocess

import numpy as np
import pytest

from conftest import ROOT

REF_BIN = os.path.join(ROOT, "oracle", "_ref", "verify_cpp")

pytestmark = pytest.mark.skipif(not os.path.exists(REF_BIN),
                                reason="oracle/_ref/verify_cpp not built")


def run_ref(tmp_path, rows):
    inp = tmp_path / "in.csv"
    out = tmp_path / "out.csv"
    cache = tmp_path / "cache.csv"
    with open(inp, "w") as f:
        f.write("position,gap,merit\n")
        for p, g in rows:
            f.write(f"{p},{g},0.000000\n")
    r = subprocess.run([REF_BIN, str(inp), str(out), str(cache)], capture_output=True,
                       text=True, env=dict(os.environ, OMP_NUM_THREADS="4"), timeout=300)
    assert r.returncode in (0, 1), r.stdout + r.stderr
    with open(out) as f:
        return [int(row["verified"]) for row in csv.DictReader(f)]


def test_mr_and_verify_gap_match_reference_binary(oracle, tmp_path):
    """One invocation (the reference's monitor thread sleeps 30 s per run)."""
    rng = random.Random(0x564F5353 + 7)
    ns = [rng.randrange(2**30, 2**48) | 1 for _ in range(4000)]
    ns += [2047, 1373653, 25326001, 3215031751, 2152302898747, 3474749660383,
           561, 41041, 2, 3, 37, 41, 2**31 - 1, 2**47 - 115]
    r = oracle.run_plain(10**9, 30000000000)
    gaps = [(p, g) for p, g, _ in r["large"]]
    # wrong gaps: shifted ends, a gap that skips a prime, odd gaps
    gaps += [(p, g - 2) for p, g in gaps[:50]] + [(p + 2, g + 2) for p, g in gaps[:50]]
    gaps += [(11, 4), (13, 6), (7, 0), (15, 4), (9, 1)]
    # malformed rows: a negative gap (p_before = p_after + |gap|, "verified"
    # when both are prime: gap < 2), and gaps beyond p_after whose p_before
    # wraps to a prime just below 2^64 (2^64 - 59, 2^64 - 83)
    gaps += [(7, -4), (11, -2), (7, -6), (9, -2), (7, 66), (13, 72), (2, 61), (3, 62),
             (31, 114), (5, 88)]
    ref = run_ref(tmp_path, [(n, 0) for n in ns] + gaps)
    ref_mr, ref_gap = ref[:len(ns)], ref[len(ns):]
    mine_mr = oracle.is_prime_batch(np.array(ns, dtype=np.uint64))
    assert ref_mr == [int(x) for x in mine_mr]
    mine_gap = oracle.verify_gaps(np.array([p for p, _ in gaps], dtype=np.uint64),
                                  np.array([g for _, g in gaps], dtype=np.int32))
    assert ref_gap == [int(x) for x in mine_gap]
    assert sum(ref_gap[:len(r["large"])]) == len(r["large"])
