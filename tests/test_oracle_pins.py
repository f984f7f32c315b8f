"""Pins the CPU oracle against the reference's own fixtures, published
values and closed forms (no reference code is executed here; the reference
CPU verifier built from its source under oracle/_ref is used in
test_oracle_ref.py).

Sources (reference tree): data/verification_db.json (OEIS A006880),
results/hist_snapshots/*.csv, results/*.csv, README.md:122-127,305-310,
results/submission_candidates.csv, results/anomalous_gaps.json.
"""
import csv
import json
import math
import os
import random

import numpy as np
import pytest

from conftest import GOLDEN

REF = os.path.join(GOLDEN, "reference")


def read_hist(path):
    with open(path) as f:
        rows = list(csv.reader(f))[1:]
    return {int(a): int(b) for a, b in rows}


def nonzero(arr):
    return {g: int(c) for g, c in enumerate(arr) if c}


def check_invariants(r, N):
    """The reference's three runtime checks (src/voss_v7_live.cu:693-705)."""
    g = r["gaps"]
    tgc = int(g.sum())
    wsum = int((np.arange(len(g), dtype=np.int64) * g).sum())
    assert tgc == r["total_primes"] - 1
    assert wsum == r["last_prime"] - 2
    assert (r["c1"] + 1) + (r["c3"] + 1) + 1 == r["total_primes"]


@pytest.mark.parametrize("k", range(1, 11))
def test_pi_oeis(oracle, k):
    db = json.load(open(os.path.join(REF, "verification_db.json")))["pi_of_10_pow_n"]
    r = oracle.run_plain(10**k, 30000000000)
    assert r["total_primes"] == db[str(k)]
    check_invariants(r, 10**k)


def test_hist_1000_fixture(oracle):
    r = oracle.run_plain(1000, 30000000000)
    assert nonzero(r["gaps"]) == read_hist(os.path.join(REF, "hist_1000.csv"))


@pytest.mark.slow
def test_hist_1e11_fixture_and_chebyshev(oracle):
    r = oracle.run_plain(10**11, 30000000000)
    assert nonzero(r["gaps"]) == read_hist(os.path.join(REF, "hist_100000000000.csv"))
    vals = json.load(open(os.path.join(REF, "readme_values.json")))
    assert r["total_primes"] == vals["pi"]["100000000000"]
    assert [r["c1"] + 1, r["c3"] + 1] == vals["chebyshev"]["100000000000"]
    check_invariants(r, 10**11)


@pytest.mark.parametrize("N", [10**9, 10**10])
def test_chebyshev_readme(oracle, N):
    vals = json.load(open(os.path.join(REF, "readme_values.json")))
    r = oracle.run_plain(N, 30000000000)
    assert [r["c1"] + 1, r["c3"] + 1] == vals["chebyshev"][str(N)]


def test_modq_structure(oracle):
    """q=6 and q=30 counts: zero off the coprime residues, sums = pi - 3,
    and each residue class matches a direct count (p mod q) of the primes."""
    N = 10**7
    r = oracle.run_plain(N, 30000000000)
    modq = r["modq"]
    assert int(modq[0:6].sum()) == r["total_primes"] - 3
    assert int(modq[6:36].sum()) == r["total_primes"] - 3
    for a in range(6):
        if math.gcd(a, 6) != 1:
            assert modq[a] == 0
    for a in range(30):
        if math.gcd(a, 30) != 1:
            assert modq[6 + a] == 0
    # direct count with an independent sieve (numpy Eratosthenes)
    s = np.ones(N + 1, dtype=bool)
    s[:2] = False
    for i in range(2, int(N**0.5) + 1):
        if s[i]:
            s[i * i::i] = False
    ps = np.nonzero(s)[0]
    ps = ps[ps >= 7]
    for a in (1, 5):
        assert modq[a] == int((ps % 6 == a).sum())
    for a in (1, 7, 11, 13, 17, 19, 23, 29):
        assert modq[6 + a] == int((ps % 30 == a).sum())
    assert r["c1"] == int((ps % 4 == 1).sum())


def brute_reference_loop(N, SEG):
    """Independent restatement of the reference's per-segment host loop
    (src/voss_v7_live.cu:434-635) on a plain list of primes: in-segment pairs
    feed the gap, per-residue and large-gap outputs; the boundary pair feeds
    the main histogram only."""
    s = np.ones(N + 1, dtype=bool)
    s[:2] = False
    for i in range(2, int(N**0.5) + 1):
        if s[i]:
            s[i * i::i] = False
    gaps = {}
    mq = {}
    large = []
    last = 5
    num_seg = (N + SEG - 1) // SEG
    w30 = [1, 7, 11, 13, 17, 19, 23, 29]
    q6 = [0, 0, 1, 0, 1, 0, 1, 1]
    for seg in range(num_seg):
        lo, hi = seg * SEG + 1, min((seg + 1) * SEG, N)
        v = [int(x) for x in np.nonzero(s[lo:hi + 1])[0] + lo if x >= 7]
        if v:
            d0 = v[0] - last
            gaps[d0] = gaps.get(d0, 0) + 1
        for a, b in zip(v, v[1:]):
            g = b - a
            gaps[g] = gaps.get(g, 0) + 1
            if g < 100:
                w = w30.index(a % 30)
                for row in (q6[w], 2 + w):
                    mq[(row, g)] = mq.get((row, g), 0) + 1
            merit = g / math.log(a) if g > 100 else 0.0
            if g >= 500 or merit >= 10.0:
                large.append((b, g))
        if v:
            last = v[-1]
    gaps[1] = gaps.get(1, 0) + 1
    gaps[2] = gaps.get(2, 0) + 1
    return gaps, mq, sorted(large)


@pytest.mark.parametrize("N,SEG", [(300000, 30000), (1000003, 30030), (2 * 10**6, 90)])
def test_segment_straddle_rule(oracle, N, SEG):
    gaps, mq, large = brute_reference_loop(N, SEG)
    r = oracle.run_plain(N, SEG)
    assert nonzero(r["gaps"]) == gaps
    got = {(row, g): int(c) for row in range(10) for g, c in enumerate(r["modq_gap"][row]) if c}
    assert got == mq
    assert sorted((p, g) for p, g, _ in r["large"]) == large


def test_straddle_shortfall_structure_1e14_fixture():
    """In the reference's 1e14 outputs, summing modq_gap_hist over the q=6
    residues falls short of gap_histogram only at gaps a SEG_NUM boundary
    (a multiple of 30) can split: never at 4, 6, 10 or 16."""
    main = read_hist(os.path.join(REF, "gap_histogram_1e14.csv"))
    with open(os.path.join(REF, "modq_gap_hist_1e14.csv")) as f:
        rows = list(csv.reader(f))[1:]
    q6 = {}
    for q, res, g, c in rows:
        if q == "6":
            q6[int(g)] = q6.get(int(g), 0) + int(c)
    short = {g: main.get(g, 0) - q6.get(g, 0) for g in range(1, 100)}
    for g in (4, 6, 10, 16):
        assert short[g] == 0
    assert all(v >= 0 for v in short.values())
    assert short[1] == 1      # (2,3)
    assert short[2] >= 2      # (3,5), (5,7) and boundary pairs
    # 3,333 boundaries at SEG_NUM = 3e10 split 2,865 pairs below 100
    assert sum(short.values()) == 2868


def test_straddle_shortfall_structure_oracle(oracle):
    """The oracle shows the same structure with many boundaries."""
    r = oracle.run_plain(10**8, 3000000)
    main = r["gaps"]
    q6 = r["modq_gap"][0] + r["modq_gap"][1]
    for g in (4, 6, 10, 16):
        assert main[g] == q6[g]
    boundaries = 10**8 // 3000000
    short = int(main[2:100].sum() - q6[2:100].sum())
    assert 2 <= short <= boundaries + 2 + 1


def merit_kats():
    kats = []
    with open(os.path.join(REF, "submission_candidates.csv")) as f:
        for row in list(csv.DictReader(f)):
            kats.append((int(row["position"]), int(row["gap"]), row["merit"]))
    js = json.load(open(os.path.join(REF, "anomalous_gaps.json")))
    for e in js["record_gaps_summary"]:
        kats.append((int(e["p"]), int(e["gap"]), f"{e['merit']:.6f}"))
    return kats


@pytest.mark.parametrize("pos,gap,merit", merit_kats())
def test_merit_kats(oracle, pos, gap, merit):
    """merit = float(g / log(p_before)) with position = p_after."""
    assert f"{oracle.merit(pos - gap, gap):.6f}" == merit


def test_maximal_gaps_1e11(oracle):
    """Known maximal gaps below 1e11 (public table A002386/A005250) appear in
    the large-gap list with position = p_after."""
    r = oracle.run_plain(10**11 // 10, 30000000000)  # 1e10 keeps the test fast
    got = {(p - g, g) for p, g, _ in r["large"]}
    for p_before, g in [(1294268491, 288), (1453168141, 292), (2300942549, 320),
                        (3842610773, 336), (4302407359, 354)]:
        assert (p_before, g) in got
        assert oracle.verify_gap(p_before + g, g)


# ------------------------------------------------------------ Miller-Rabin

PSEUDOPRIMES = [2047, 1373653, 25326001, 3215031751, 2152302898747, 3474749660383,
                341550071728321, 3825123056546413051]


@pytest.mark.parametrize("n", PSEUDOPRIMES + [561, 41041, 1105, 1729, 2465])
def test_mr_kat_composite(oracle, n):
    assert oracle.is_prime(n) is False


@pytest.mark.parametrize("n", [2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41,
                               2**61 - 1, 2**64 - 59, 2**31 - 1, 1000000007])
def test_mr_kat_prime(oracle, n):
    assert oracle.is_prime(n) is True


def test_mr_base_multiples(oracle):
    for b in [2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37]:
        for k in (2, 3, 1000003):
            assert oracle.is_prime(b * k) is False
    assert oracle.is_prime(0) is False and oracle.is_prime(1) is False


def test_mr_bruteforce_small(oracle):
    N = 2 * 10**6
    s = np.ones(N, dtype=bool)
    s[:2] = False
    for i in range(2, int(N**0.5) + 1):
        if s[i]:
            s[i * i::i] = False
    v = oracle.is_prime_batch(np.arange(N, dtype=np.uint64))
    assert np.array_equal(v.astype(bool), s)


def test_mr_vs_sympy_random(oracle):
    sympy = pytest.importorskip("sympy")
    rng = random.Random(0x564F5353)
    ns = [rng.randrange(2**40, 2**64) | 1 for _ in range(3000)]
    ns += [int(sympy.nextprime(rng.randrange(2**40, 2**64))) for _ in range(300)]
    v = oracle.is_prime_batch(np.array(ns, dtype=np.uint64))
    assert [bool(x) for x in v] == [bool(sympy.isprime(n)) for n in ns]


def test_mr_mulmod_count(oracle):
    """Work unit: a composite rejected at base 2 costs bitlen(d)+popcount(d)
    plus its squarings; trial-division hits cost nothing."""
    assert oracle.mr_mulmods(3 * 5) == 0
    n = 2**61 - 1  # prime: d = (n-1)/2, r = 1 -> 12 * (bitlen + popcount)
    d = (n - 1) // 2
    assert oracle.mr_mulmods(n) == 12 * (d.bit_length() + bin(d).count("1"))


def witness_squarings(n, a):
    """Squarings is_prime_u64's witness loop runs for base a
    (src/verify_cpp.cpp:47-54): none when a^d is 1 or -1, else one per step
    until -1 appears or r - 1 steps are spent; and whether a passes."""
    d, r = n - 1, 0
    while d % 2 == 0:
        d //= 2
        r += 1
    x = pow(a, d, n)
    if x in (1, n - 1):
        return 0, True
    for j in range(1, r):
        x = x * x % n
        if x == n - 1:
            return j, True
    return r - 1, False


@pytest.mark.parametrize("n,stop_base", [
    (41, None),                    # prime, r = 3: witness loops of 0..2 squarings
    (998244353, None),             # prime 119 * 2^23 + 1, r = 23
    (2**64 - 59, None),            # prime, r = 1
    (1373653, 5),                  # psi_2 = 829 * 1657: passes 2 and 3, fails at 5
    (25326001, 7),                 # psi_3: passes 2, 3, 5, fails at 7
    (3825123056546413051, 37),     # psi_9..psi_11: passes 2..31, fails at 37
    (341, 2),                      # 11 * 31: trial division (11 is a base)
    (4033, None),                  # 37 * 109, a base-2 strong pseudoprime: trial (37)
])
def test_mr_mulmod_count_witness_loop(oracle, n, stop_base):
    """The count with the witness-loop squarings, on primes with r > 1 and
    on strong pseudoprimes that pass some bases before one rejects them:
    per base tried, bitlen(d) + popcount(d) products of mod_pow plus the
    squarings that base's loop runs (each base's loop counted from the
    base's own power, so a dropped or extra squaring in any base fails)."""
    bases = (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37)
    if any(n % b == 0 for b in bases):
        assert oracle.mr_mulmods(n) == 0
        return
    d = n - 1
    while d % 2 == 0:
        d //= 2
    per = d.bit_length() + bin(d).count("1")
    want = 0
    for a in bases:
        sq, ok = witness_squarings(n, a)
        want += per + sq
        if not ok:
            assert a == stop_base, (n, a)
            break
    else:
        assert stop_base is None
    assert oracle.mr_mulmods(n) == want
    # at least one tested number exercises loops of 2+ squarings
    if n == 998244353:
        assert max(witness_squarings(n, a)[0] for a in bases) >= 2


def test_verify_gap_semantics(oracle):
    assert oracle.verify_gap(11, 4)          # 7 -> 11
    assert not oracle.verify_gap(13, 6)      # 7 -> 13 skips 11
    assert not oracle.verify_gap(15, 4)      # 15 composite
    assert oracle.verify_gap(25056082543, 456)   # maximal gap below 1e11
    assert not oracle.verify_gap(25056082543, 454)
    assert oracle.verify_gap(7, 0)           # gap < 2: only the ends


def test_mr_base2_pseudoprimes_above_2_31(oracle):
    """Composites built from their factors that pass the base-2 strong test:
    the oracle's later witnesses reject every one (and base 2 alone would
    not: each is checked to pass it)."""
    from pseudoprimes import chernick_carmichaels, semiprime_pseudoprimes
    semi = semiprime_pseudoprimes(oracle.is_prime_batch, 0x564F5353 + 40, starts=2)
    carm = chernick_carmichaels(oracle.is_prime_batch)
    assert len(semi) > 500 and len(carm) > 1000 and sum(s for _, s in carm) > 100
    ns = np.array(semi + [n for n, _ in carm], dtype=np.uint64)
    assert not oracle.is_prime_batch(ns).any()


@pytest.mark.parametrize("lo,hi", [(10**12, 10**12 + 2 * 10**6), (10**15 + 7, 10**15 + 10**6),
                                   (49, 100000), (3 * 10**9, 3 * 10**9 + 10**7)])
def test_outputs_from_a_prime_list(oracle, lo, hi):
    """tests/from_primes.py (the interval outputs derived from the list of the
    interval's primes) equals the oracle's sieve: what lets a primality test
    stand in for the sieve where the oracle is too slow."""
    from from_primes import interval_results, interval_results_np
    ns = np.arange(lo | 1, hi, 2, dtype=np.uint64)
    ps = ns[oracle.is_prime_batch(ns).astype(bool)]
    r = oracle.run_interval(lo, hi)
    for d in (interval_results(ps), interval_results_np(ps)):
        for k in ("total_primes", "first_prime", "last_prime", "c1", "c3", "large_count"):
            assert d[k] == r[k], k
        assert np.array_equal(d["modq"], r["modq"])
        assert np.array_equal(d["gaps"], r["gaps"])
        assert np.array_equal(d["modq_gap"], r["modq_gap"])
        assert d["large"] == [(p, g, float(np.float32(m))) for p, g, m in r["large"]]
    if lo == 3 * 10**9:
        assert r["large_count"] > 0  # the merit rule is exercised
