"""GPU Miller-Rabin and gap verification parity against the oracle
(is_prime_u64 / verify_gap restatements, pinned by the reference binary)."""
import random

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KAT = [0, 1, 2, 3, 4, 5, 9, 25, 37, 41, 561, 1105, 1729, 2047, 41041, 1373653, 25326001,
       3215031751, 2152302898747, 3474749660383, 341550071728321, 3825123056546413051,
       2**61 - 1, 2**64 - 59, 2**64 - 1, 2**63 + 1, 2**62 + 135, 2**31 - 1, 1000000007]


def test_mr_kats(engine, oracle):
    v = engine.mr(np.array(KAT, dtype=np.uint64))
    assert [int(x) for x in v] == [int(oracle.is_prime(n)) for n in KAT]


def test_mr_small_exhaustive(engine, oracle):
    ns = np.arange(0, 3 * 10**6, dtype=np.uint64)
    assert np.array_equal(engine.mr(ns), oracle.is_prime_batch(ns))


def test_mr_random_64bit(engine, oracle):
    rng = np.random.default_rng(0x564F5353 + 2)
    ns = rng.integers(0, 2**63, size=10**6, dtype=np.uint64) * np.uint64(2) + np.uint64(1)
    assert np.array_equal(engine.mr(ns), oracle.is_prime_batch(ns))


def test_mr_device_path(engine, oracle):
    rng = np.random.default_rng(7)
    ns = rng.integers(2**40, 2**63, size=200000, dtype=np.uint64) | np.uint64(1)
    v, times = engine.mr_device(ns, repeats=2)
    assert np.array_equal(v, oracle.is_prime_batch(ns))
    assert len(times) == 2 and all(t > 0 for t in times)


def test_verify_gaps(engine, oracle):
    r = oracle.run_plain(10**9, 30000000000)
    rows = [(p, g) for p, g, _ in r["large"]]
    rows += [(p, g - 2) for p, g in rows[:40]] + [(p + 2, g + 2) for p, g in rows[:40]]
    rows += [(11, 4), (13, 6), (7, 0), (15, 4), (9, 1), (25056082543, 456), (25056082543, 454)]
    rows += [(3, 1), (5, 3), (7, 5), (7, 2), (5, 2), (29, 27), (37, 2)]
    # malformed rows, decided row by row as the reference does (the oracle's
    # verdicts are pinned to the reference binary in test_oracle_ref.py): a
    # negative gap, and gaps beyond p_after whose p_before wraps to a prime
    # just below 2^64
    rows += [(7, -4), (11, -2), (7, -6), (9, -2), (7, 66), (13, 72), (2, 61), (3, 62),
             (31, 114), (5, 88)]
    # claimed gaps beyond the per-thread interior check (65,536), both ends
    # prime: they go through the interval-sieve path and fail (primes inside)
    for pa in (1000000007, 10**12 + 39):
        g = 65538
        while not oracle.is_prime(pa - g):
            g += 2
        rows.append((pa, g))
    pa = np.array([p for p, _ in rows], dtype=np.uint64)
    g = np.array([x for _, x in rows], dtype=np.int32)
    assert np.array_equal(engine.verify_gaps(pa, g), oracle.verify_gaps(pa, g))
    # the last primes below 2^64, found by testing every odd number downwards:
    # consecutive pairs verify, a pair that skips one does not (the interior
    # ends next to 2^64; the oracle's sieve up to 2^32 is too slow here)
    top = [n for n in range(2**64 - 1, 2**64 - 400, -2) if oracle.is_prime(n)][:3]
    pa = np.array([top[0], top[1], top[0]], dtype=np.uint64)
    g = np.array([top[0] - top[1], top[1] - top[2], top[0] - top[2]], dtype=np.int32)
    assert engine.verify_gaps(pa, g).tolist() == [1, 1, 0]
    # claimed gaps of 2^16 .. 2^31 - 1 with both ends prime, next to 2^63
    # and 2^64 (beyond the interval sieve's 2^62 bound): a prime lies inside
    # every one of them (no prime gap below 2^64 exceeds 1550), so none
    # verifies, and one malformed row does not fail the others
    rows = [(top[0], top[0] - top[1])]
    for pa0 in (top[0], [n for n in range(2**63 + 1, 2**63 + 4000, 2) if oracle.is_prime(n)][0]):
        for g0 in (65538, 10**6, 2**31 - 4002):  # even: pa0 - g1 stays odd
            g1 = next(x for x in range(g0, g0 + 4000, 2) if oracle.is_prime(pa0 - x))
            rows.append((pa0, g1))
    pa = np.array([p for p, _ in rows], dtype=np.uint64)
    g = np.array([x for _, x in rows], dtype=np.int32)
    assert engine.verify_gaps(pa, g).tolist() == [1] + [0] * (len(rows) - 1)


def test_mr_near_powers_of_two(engine, oracle):
    """Odd n around every 2^k: crosses the ranges in which the kernels
    compute 2^64 mod n differently (below 2^13, up to 2^62, 2^62..2^63,
    2^63 and above) and the 32-bit halves of the Montgomery square."""
    ns = sorted({(1 << k) + d for k in range(2, 65) for d in range(-301, 302, 2)
                 if 0 < (1 << k) + d < 2**64})
    ns = np.array(ns, dtype=np.uint64)
    assert np.array_equal(engine.mr(ns), oracle.is_prime_batch(ns))


def test_mr_every_odd_below_2_31_against_a_sieve(engine):
    """Every n < 2^31 against a sieve of Eratosthenes: the range in which
    the kernels' witnesses are reduced mod n (the largest is 1795265022)
    and a witness that n divides is skipped."""
    lim = 1 << 31
    odd_prime = np.ones(lim // 2, dtype=np.bool_)  # index i <-> 2i + 1
    odd_prime[0] = False
    for i in range(1, (int(lim ** 0.5) + 1) // 2 + 1):
        if odd_prime[i]:
            p = 2 * i + 1
            odd_prime[p * p // 2::p] = False
    assert int(odd_prime.sum()) + 1 == 105097565  # pi(2^31)
    assert [int(x) for x in engine.mr(np.arange(0, 4, dtype=np.uint64))] == [0, 0, 1, 1]
    chunk = 1 << 26
    for start in range(0, lim // 2, chunk):
        ns = np.arange(start, start + chunk, dtype=np.uint64) * np.uint64(2) + np.uint64(1)
        v = engine.mr(ns)
        assert np.array_equal(v.astype(np.bool_), odd_prime[start:start + chunk]), start


def test_mr_base2_pseudoprimes_above_2_31(engine, oracle):
    """Base-2 strong pseudoprimes above 2^31 (semiprimes p(2p - 1) and
    Chernick Carmichael numbers, built from their factors): the kernels pass
    them through the base-2 stage and the further witnesses must reject
    every one, as the oracle's do."""
    from pseudoprimes import chernick_carmichaels, semiprime_pseudoprimes
    semi = semiprime_pseudoprimes(oracle.is_prime_batch, 0x564F5353 + 41)
    carm = chernick_carmichaels(oracle.is_prime_batch)
    ns = np.array(semi + [n for n, _ in carm], dtype=np.uint64)
    print(f"[gpu] {len(ns)} base-2 pseudoprimes", flush=True)
    assert len(ns) > 3000
    # interleaved with primes so the waves also hold lanes that run every witness
    rng = np.random.default_rng(0x564F5353 + 42)
    odd = rng.integers(2**31, 2**63, size=40 * len(ns), dtype=np.uint64) | np.uint64(1)
    primes = odd[oracle.is_prime_batch(odd).astype(bool)]
    assert len(primes) >= len(ns)
    mix = np.empty(2 * len(ns), dtype=np.uint64)
    mix[0::2] = ns
    mix[1::2] = primes[:len(ns)]
    v = engine.mr(mix)
    assert not v[0::2].any()
    assert np.array_equal(v, oracle.is_prime_batch(mix))
