"""Seeded synthetic workloads for the batched, verification and serving forms.

SplitMix64 (Steele, Lea, Flood 2014) with seed 0x564F5353 ("VOSS"); stream k
uses seed + k.  Uniform reals are (x >> 11) * 2^-53.

  stream 1  batch intervals: start = 1e12 + floor(u * (1e16 - 1e12)),
            length = floor(10 ** (4 + 5 u'))
  stream 2  Miller-Rabin candidates: even indices odd-uniform in [2^40, 2^64),
            odd indices drawn from a list of sieve gap endpoints
  stream 3  serving requests: Poisson arrivals, half interval requests
            (start uniform in [1e10, 1e16), length log-uniform in [1e5, 1e10]),
            half MR batches (size log-uniform in [1e3, 1e6])
"""
import numpy as np

SEED0 = 0x564F5353
_M64 = (1 << 64) - 1


def splitmix64(seed):
    """Infinite generator of SplitMix64 outputs."""
    x = seed & _M64
    while True:
        x = (x + 0x9E3779B97F4A7C15) & _M64
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        yield z ^ (z >> 31)


def splitmix64_array(seed, n):
    """n SplitMix64 outputs as uint64 (vectorised; same values as the generator)."""
    i = np.arange(1, n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (np.uint64(seed & _M64) + i * np.uint64(0x9E3779B97F4A7C15))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def unit(x):
    return float(x >> 11) * (2.0 ** -53)


def batch_intervals(n=4096, seed=SEED0 + 1, lo_min=10**12, lo_max=10**16,
                    len_exp=(4.0, 9.0)):
    """The batched workload: n intervals [start, start+len)."""
    g = splitmix64(seed)
    out = []
    for _ in range(n):
        start = lo_min + int(unit(next(g)) * (lo_max - lo_min))
        length = int(10 ** (len_exp[0] + (len_exp[1] - len_exp[0]) * unit(next(g))))
        out.append((start, start + length))
    return out


def mr_candidates(n, endpoints, seed=SEED0 + 2):
    """n MR candidates: even indices odd-uniform in [2^40, 2^64), odd indices
    a seeded choice of the given gap endpoints."""
    endpoints = np.asarray(endpoints, dtype=np.uint64)
    raw = splitmix64_array(seed, 2 * n)
    out = np.empty(n, dtype=np.uint64)
    ev = raw[0:n:2] if n > 0 else raw[:0]
    # redraw-free mapping into [2^40, 2^64): values below 2^40 get bit 40 set
    ev = np.where(ev < np.uint64(1 << 40), ev | np.uint64(1 << 40), ev) | np.uint64(1)
    out[0::2] = ev[: (n + 1) // 2]
    if n > 1:
        if len(endpoints) == 0:
            raise ValueError("no gap endpoints to draw from")
        pick = raw[n:n + n // 2] % np.uint64(len(endpoints))
        out[1::2] = endpoints[pick.astype(np.int64)]
    return out


def serve_requests(n, rate_per_s, seed=SEED0 + 3):
    """n serving requests: (arrival_s, kind, payload).  kind "interval":
    payload (lo, hi); kind "mr": payload = candidate count."""
    g = splitmix64(seed)
    t = 0.0
    out = []
    for _ in range(n):
        t += -np.log(max(unit(next(g)), 1e-300)) / rate_per_s
        if unit(next(g)) < 0.5:
            lo = 10**10 + int(unit(next(g)) * (10**16 - 10**10))
            ln = int(10 ** (5 + 5 * unit(next(g))))
            out.append((t, "interval", (lo, lo + ln)))
        else:
            cnt = int(10 ** (3 + 3 * unit(next(g))))
            out.append((t, "mr", cnt))
    return out
