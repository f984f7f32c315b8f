"""Base-2 strong pseudoprimes above 2^31, built from their factors: the
composites on which the base-2 Miller-Rabin round passes and only further
witnesses decide.

* semiprimes p(2p - 1) with p and 2p - 1 prime: n - 1 = (2p + 1)(p - 1),
  and when 2p - 1 = +-1 (mod 8) the order of 2 divides p - 1 modulo both
  factors, so n is a base-2 Fermat pseudoprime; those that pass the strong
  test are kept;
* Chernick's Carmichael numbers (6k + 1)(12k + 1)(18k + 1) with all three
  factors prime: Fermat pseudoprimes to every coprime base; all are
  returned, with the strong-test outcome for base 2.
"""
import numpy as np


def strong_probable_prime_base2(n):
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    x = pow(2, d, n)
    if x in (1, n - 1):
        return True
    for _ in range(s - 1):
        x = x * x % n
        if x == n - 1:
            return True
    return False


def semiprime_pseudoprimes(is_prime_batch, seed, starts=8, span=1 << 20):
    """Base-2 strong pseudoprimes p(2p - 1) in (2^31, 2^64), p drawn from
    `starts` seeded windows of `span` odd numbers below 2^31."""
    rng = np.random.default_rng(seed)
    out = []
    for start in rng.integers(1 << 16, (1 << 31) - span, size=starts):
        cand = np.arange(int(start) | 1, int(start) + span, 2, dtype=np.uint64)
        p = cand[is_prime_batch(cand).astype(bool)]
        p = p[is_prime_batch(np.uint64(2) * p - np.uint64(1)).astype(bool)]
        for x in p:
            n = int(x) * (2 * int(x) - 1)
            if 2**31 < n < 2**64 and strong_probable_prime_base2(n):
                out.append(n)
    return out


def chernick_carmichaels(is_prime_batch, kmax=150000):
    """[(n, strong probable prime to base 2)] for n = (6k+1)(12k+1)(18k+1)
    in (2^31, 2^64) with all three factors prime."""
    k = np.arange(1, kmax, dtype=np.uint64)
    a, b, c = np.uint64(6) * k + np.uint64(1), np.uint64(12) * k + np.uint64(1), \
        np.uint64(18) * k + np.uint64(1)
    m = is_prime_batch(a).astype(bool) & is_prime_batch(b).astype(bool) & \
        is_prime_batch(c).astype(bool)
    out = []
    for x in k[m]:
        n = (6 * int(x) + 1) * (12 * int(x) + 1) * (18 * int(x) + 1)
        if 2**31 < n < 2**64:
            out.append((n, strong_probable_prime_base2(n)))
    return out
