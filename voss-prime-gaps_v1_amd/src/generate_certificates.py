#!/usr/bin/env python3
"""VOSS certificates, MI355X drop-in for scripts/generate_certificates.py.

Reads results/large_gaps.csv and writes, like the reference
(scripts/generate_certificates.py:136-240):
  results/certificates.csv  -- p_before,p_after,gap,merit,sha256,verified for
                               every gap, in file order
  results/cert_top20/*.txt  -- the 20 largest merits (stable sort)

The certificate hash is SHA-256 over "VOSS|p_before|p_after|gap|merit" with
the merit to 6 decimals (:105-108).  Verification flags come from
results/verification_details.csv when it exists (:163-195); gaps missing
from it are verified by the GPU verifier (bin/verify_cpp of this build,
GPU Miller-Rabin plus an interval sieve of the interior) instead of the
reference's numba functions (:22-92), which need no numba here.  Paths are
relative to this file's parent directory, as in the reference.
"""
import csv
import hashlib
import math
import os
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, '..'))
INPUT = os.path.join(ROOT, 'results', 'large_gaps.csv')
CERT_CSV = os.path.join(ROOT, 'results', 'certificates.csv')
TOP20_DIR = os.path.join(ROOT, 'results', 'cert_top20')
DETAILS = os.path.join(ROOT, 'results', 'verification_details.csv')
TOP_N = 20


def sha256_cert(p_before, p_after, gap, merit):
    """Deterministic SHA-256 over the certificate content."""
    content = f"VOSS|{p_before}|{p_after}|{gap}|{merit:.6f}"
    return hashlib.sha256(content.encode()).hexdigest()


def find_verifier():
    """bin/verify_cpp of this build: $VOSS_AMD_HOME, or this tree's bin/."""
    cands = []
    home = os.environ.get('VOSS_AMD_HOME')
    if home:
        cands += [os.path.join(home, 'voss-prime-gaps_v1_amd', 'bin', 'verify_cpp'),
                  os.path.join(home, 'bin', 'verify_cpp')]
    cands += [os.path.join(ROOT, 'bin', 'verify_cpp')]
    for c in cands:
        if os.path.isfile(c) and os.access(c, os.X_OK):
            return c
    return None


def read_details(path):
    out = {}
    with open(path) as f:
        for r in csv.DictReader(f):
            out[(int(r['position']), int(r['gap']))] = (r['verified'] == '1')
    return out


def verify_missing(rows):
    """Verification flags of rows through the GPU verifier program."""
    exe = find_verifier()
    if exe is None:
        print("ERROR: gaps without verification details and no verify_cpp "
              "(build voss-prime-gaps_v1_amd and set VOSS_AMD_HOME)")
        sys.exit(1)
    with tempfile.TemporaryDirectory() as tmp:
        inp = os.path.join(tmp, 'large_gaps.csv')
        det = os.path.join(tmp, 'details.csv')
        with open(inp, 'w') as f:
            f.write('position,gap,merit\n')
            for p_after, gap, merit in rows:
                f.write(f'{p_after},{gap},{merit:.6f}\n')
        subprocess.run([exe, inp, det, os.path.join(tmp, 'cache.csv')],
                       stdout=subprocess.DEVNULL)
        return read_details(det)


def write_text_cert(filepath, p_before, p_after, gap, merit, sha):
    with open(filepath, 'w', encoding='utf-8') as f:
        f.write("=" * 66 + "\n")
        f.write("           VOSS NUMERICAL CERTIFICATE OF PRIME GAP\n")
        f.write("=" * 66 + "\n")
        f.write(f"Prime Before (p)        : {p_before:,}\n")
        f.write(f"Prime After  (p+g)      : {p_after:,}\n")
        f.write(f"Gap Size (g)            : {gap:,}\n")
        f.write(f"Merit (g / ln p)        : {merit:.6f}\n")
        f.write(f"ln(p)                   : {math.log(p_before):.6f}\n")
        f.write(f"Max Factor Limit (sqrtp+g) : {math.isqrt(p_after):,}\n")
        f.write("-" * 66 + "\n")
        f.write("VERIFICATION METHOD:\n")
        f.write("  1. p_before and p_after are proven prime via\n")
        f.write("     Deterministic Miller-Rabin (bases 2,3,5,7,11,13,17,19,23,29,31,37).\n")
        f.write(f"  2. All {gap-1} integers in (p, p+g) are proven composite\n")
        f.write("     via Interval Sieve with base primes up to √(p+g).\n")
        f.write("-" * 66 + "\n")
        f.write(f"SHA-256: {sha}\n")
        f.write(f"Generated: {time.strftime('%Y-%m-%d %H:%M:%S')}\n")
        f.write("=" * 66 + "\n")


def main():
    if not os.path.exists(INPUT):
        print(f"ERROR: {INPUT} not found")
        sys.exit(1)
    print(f"Reading: {INPUT}")
    rows = []
    with open(INPUT) as f:
        for r in csv.DictReader(f):
            rows.append((int(r['position']), int(r['gap']), float(r['merit'])))
    print(f"Gaps: {len(rows):,}")
    if not rows:
        print("Nothing to do")
        return

    t0 = time.time()
    verified = {}
    if os.path.exists(DETAILS):
        print(f"Reading verification details: {DETAILS}")
        verified = read_details(DETAILS)
        print(f"  Loaded {len(verified):,} verification results")
    missing = [r for r in rows if (r[0], r[1]) not in verified]
    if missing:
        print(f"Verifying {len(missing):,} gaps on the GPU")
        verified.update(verify_missing(missing))

    os.makedirs(os.path.dirname(CERT_CSV), exist_ok=True)
    print(f"\nGenerating: {CERT_CSV}")
    n_ok = 0
    with open(CERT_CSV, 'w', newline='') as f:
        w = csv.writer(f)
        w.writerow(['p_before', 'p_after', 'gap', 'merit', 'sha256', 'verified'])
        for p_after, gap, merit in rows:
            p_before = p_after - gap
            ok = verified.get((p_after, gap), False)
            w.writerow([p_before, p_after, gap, f"{merit:.6f}",
                        sha256_cert(p_before, p_after, gap, merit), "YES" if ok else "NO"])
            n_ok += ok
    dt = time.time() - t0
    print(f"  Wrote {len(rows):,} certificates in {dt:.2f}s")
    print(f"  Verified: {n_ok:,}/{len(rows):,}")

    os.makedirs(TOP20_DIR, exist_ok=True)
    top = sorted(rows, key=lambda x: -x[2])[:TOP_N]
    print(f"\nGenerating top {TOP_N} text certificates in: {TOP20_DIR}")
    for i, (p_after, gap, merit) in enumerate(top, 1):
        p_before = p_after - gap
        sha = sha256_cert(p_before, p_after, gap, merit)
        fname = f"cert_{i:02d}_merit{merit:.2f}_p{p_after}.txt"
        write_text_cert(os.path.join(TOP20_DIR, fname), p_before, p_after, gap, merit, sha)
        print(f"  [{i:2d}] {fname}")
    print(f"\n  CSV  : {CERT_CSV} ({len(rows):,} certificates)")
    print(f"  TXT  : {TOP20_DIR}/ ({len(top)} top certificates by merit)")


if __name__ == '__main__':
    main()
