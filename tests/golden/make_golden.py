"""Writes the oracle fixtures under tests/golden/ (run once, output committed).

    python tests/golden/make_golden.py [plain_1e11] [interval_cfg3]

The oracle (oracle/libvoss_oracle.so, pinned by tests/test_oracle_pins.py
against the reference's own fixtures) produces the outputs at sizes the
tests cannot recompute quickly; the GPU tests compare against these files.
"""
import hashlib
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import oracle as O  # noqa: E402


def large_digest(large):
    """sha256 of the canonical large-gap CSV rows (position,gap,%.6f merit)."""
    h = hashlib.sha256()
    for pos, gap, merit in large:
        h.update(f"{pos},{gap},{merit:.6f}\n".encode())
    return h.hexdigest()


def pack(r, extra):
    d = dict(extra)
    d.update({
        "total_primes": r["total_primes"], "first_prime": r["first_prime"],
        "last_prime": r["last_prime"], "c1": r["c1"], "c3": r["c3"],
        "modq": [int(x) for x in r["modq"]],
        "gaps": {str(g): int(c) for g, c in enumerate(r["gaps"]) if c},
        "modq_gap": {f"{row},{g}": int(c) for row in range(10) for g, c in
                     enumerate(r["modq_gap"][row]) if c},
        "large_count": r["large_count"],
        "large_sha256": large_digest(r["large"]),
        "large_top": [[p, g, m] for p, g, m in r["large"][:200]],
    })
    return d


def plain_1e11():
    t = time.time()
    r = O.run_plain(10**11, 30000000000, large_cap=1 << 20)
    d = pack(r, {"mode": "plain", "N": 10**11, "SEG_NUM": 30000000000,
                 "oracle_seconds": round(time.time() - t, 1)})
    with open(os.path.join(HERE, "oracle_plain_1e11.json"), "w") as f:
        json.dump(d, f, indent=0)


def interval_cfg3():
    lo, hi = 10**13, 10**13 + 10**12
    t = time.time()
    r = O.run_interval(lo, hi, large_cap=1 << 22)
    d = pack(r, {"mode": "interval", "lo": lo, "hi": hi,
                 "oracle_seconds": round(time.time() - t, 1)})
    with open(os.path.join(HERE, "oracle_interval_1e13_1e12.json"), "w") as f:
        json.dump(d, f, indent=0)


if __name__ == "__main__":
    which = sys.argv[1:] or ["plain_1e11"]
    for w in which:
        globals()[w]()
        print("wrote", w)
