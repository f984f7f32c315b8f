"""Certificates drop-in (src/generate_certificates.py) on a small results/
tree.  The reference ships no certificate fixture and its script needs numba
(absent here), so its digests are pinned by what fixes them: the record
format the reference states twice (README.md:277-278 and
scripts/generate_certificates.py:105-108, SHA-256 over
VOSS|p_before|p_after|gap|merit:.6f) and SHA-256 itself, checked here with
an implementation written from FIPS 180-4 (constants derived from the
primes, not typed) that must reproduce the standard's published examples
before it checks any certificate.  The file layout follows :196-233.
What is checked here: rows in file order, digests, verified flags taken
from the details file, top-20 selection by a stable merit sort, file
names."""
import csv
import math
import os
import shutil
import subprocess
import sys

import pytest

from conftest import PKG

def _icbrt(n):
    """floor(n ** (1/3)) for a non-negative integer n."""
    x = 1 << ((n.bit_length() + 2) // 3)
    while True:
        y = (2 * x + n // (x * x)) // 3
        if y >= x:
            break
        x = y
    while x ** 3 > n:
        x -= 1
    while (x + 1) ** 3 <= n:
        x += 1
    return x


def _first_primes(k):
    out, n = [], 2
    while len(out) < k:
        if all(n % p for p in out if p * p <= n):
            out.append(n)
        n += 1
    return out


# FIPS 180-4 4.2.2 / 5.3.3: the first 32 bits of the fractional parts of the
# cube roots of the first 64 primes, and of the square roots of the first 8
_P64 = _first_primes(64)
_K = [_icbrt(p << 96) & 0xFFFFFFFF for p in _P64]
_H0 = [math.isqrt(p << 64) & 0xFFFFFFFF for p in _P64[:8]]


def spec_sha256(msg: bytes) -> str:
    """SHA-256 (FIPS 180-4 section 6.2) in plain integer arithmetic."""
    M = 0xFFFFFFFF

    def rotr(x, n):
        return ((x >> n) | (x << (32 - n))) & M

    # 5.1.1: a 1 bit, zeros to 448 mod 512, the 64-bit message length
    data = msg + b"\x80" + b"\x00" * ((55 - len(msg)) % 64) + (8 * len(msg)).to_bytes(8, "big")
    h = list(_H0)
    for blk in range(0, len(data), 64):
        w = [int.from_bytes(data[blk + 4 * t: blk + 4 * t + 4], "big") for t in range(16)]
        for t in range(16, 64):
            s0 = rotr(w[t - 15], 7) ^ rotr(w[t - 15], 18) ^ (w[t - 15] >> 3)
            s1 = rotr(w[t - 2], 17) ^ rotr(w[t - 2], 19) ^ (w[t - 2] >> 10)
            w.append((w[t - 16] + s0 + w[t - 7] + s1) & M)
        a, b, c, d, e, f, g, hh = h
        for t in range(64):
            t1 = (hh + (rotr(e, 6) ^ rotr(e, 11) ^ rotr(e, 25)) + ((e & f) ^ (~e & g)) +
                  _K[t] + w[t]) & M
            t2 = ((rotr(a, 2) ^ rotr(a, 13) ^ rotr(a, 22)) + ((a & b) ^ (a & c) ^ (b & c))) & M
            a, b, c, d, e, f, g, hh = (t1 + t2) & M, a, b, c, (d + t1) & M, e, f, g
        h = [(x + y) & M for x, y in zip(h, (a, b, c, d, e, f, g, hh))]
    return "".join(f"{x:08x}" for x in h)


def test_spec_sha256_known_answers():
    """The FIPS 180-2 examples (one block, two blocks, one million 'a'),
    and the empty message."""
    assert spec_sha256(b"abc") == (
        "ba7816bf8f01cfea414140de5dae2223b00361a396177a9cb410ff61f20015ad")
    assert spec_sha256(b"") == (
        "e3b0c44298fc1c149afbf4c8996fb92427ae41e4649b934ca495991b7852b855")
    assert spec_sha256(b"abcdbcdecdefdefgefghfghighijhijkijkljklmklmnlmnomnopnopq") == (
        "248d6a61d20638b8e5c026930c3e6039a33ce45964ff2167f6ecedd419db06c1")
    assert spec_sha256(b"a" * 1000000) == (
        "cdc76e5c9914fb9281a1c7e284d73e67f1809a48a497200e046d39ccc7112cd0")


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
        # the record as README.md:277-278 states it
        text = f"VOSS|{p - g}|{p}|{g}|{m:.6f}"
        assert r["sha256"] == spec_sha256(text.encode())
    names = sorted(os.listdir(tmp_path / "results" / "cert_top20"))
    assert len(names) == len(ROWS)
    assert names[0] == "cert_01_merit16.07_p4302407713.txt"
    assert names[-1] == "cert_07_merit0.29_p1000000007.txt"
    body = (tmp_path / "results" / "cert_top20" / names[0]).read_text(encoding="utf-8")
    assert "Prime Before (p)        : 4,302,407,359" in body
    assert "All 353 integers in (p, p+g) are proven composite" in body
    assert "SHA-256: " + spec_sha256(b"VOSS|4302407359|4302407713|354|16.072053") in body


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
