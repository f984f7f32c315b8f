"""The interval outputs (the layout of the oracle's run_interval and of
Engine.run_interval) derived from the sorted list of the primes in
[lo, hi): every consecutive pair inside the interval counts, as in the
reference's in-segment rules (src/voss_v7_live.cu:255-312): gap histogram
for g < 100000; for p_before >= 7 the gap-by-residue rows (q=6 row of
p_before, then its q=30 row) for g < 100 and the large gaps (g >= 500, or
g > 100 with merit g / log(p_before) >= 10, the merit a float of the double
quotient, listed as (p_after, g, merit) by merit descending, then p_after);
mod-4 and mod-6/mod-30 counts over the primes.  For a list of primes that
comes from somewhere other than a sieve (a primality test), this states what
the sieve must report."""
import math

import numpy as np

W30 = (1, 7, 11, 13, 17, 19, 23, 29)
Q6_ROW = (0, 0, 1, 0, 1, 0, 1, 1)  # residue mod 6 of W30[w]: 1 -> row 0, 5 -> row 1


def interval_results(primes):
    ps = [int(x) for x in primes]
    gaps = np.zeros(100000, dtype=np.int64)
    modq_gap = np.zeros((10, 100), dtype=np.int64)
    modq = np.zeros(40, dtype=np.uint64)
    c1 = c3 = 0
    for v in ps:
        c1 += (v & 3) == 1
        c3 += (v & 3) == 3
        modq[v % 6] += 1
        modq[6 + v % 30] += 1
    large = []
    for p, q in zip(ps, ps[1:]):
        g = q - p
        if g < 100000:
            gaps[g] += 1
        if p < 7:
            continue
        if g < 100:
            w = W30.index(p % 30)
            modq_gap[Q6_ROW[w], g] += 1
            modq_gap[2 + w, g] += 1
        merit = g / math.log(float(p)) if g > 100 else 0.0
        if (g >= 500 or merit >= 10.0) and g < 100000:
            large.append((q, g, float(np.float32(merit))))
    large.sort(key=lambda r: (-r[2], r[0]))
    return {"total_primes": len(ps), "first_prime": ps[0] if ps else 0,
            "last_prime": ps[-1] if ps else 0, "c1": int(c1), "c3": int(c3), "modq": modq,
            "gaps": gaps, "modq_gap": modq_gap, "large_count": len(large), "large": large}


def interval_results_np(primes):
    """interval_results for long prime lists, vectorised with numpy: the
    same rules and output layout (merits through math.log, the C library's
    log, for the few pairs that can be large)."""
    ps = np.asarray(primes, dtype=np.uint64)
    n = len(ps)
    gaps = np.zeros(100000, dtype=np.int64)
    modq_gap = np.zeros((10, 100), dtype=np.int64)
    modq = np.zeros(40, dtype=np.uint64)
    r6 = (ps % np.uint64(6)).astype(np.int64)
    r30 = (ps % np.uint64(30)).astype(np.int64)
    modq[:6] = np.bincount(r6, minlength=6)[:6].astype(np.uint64)
    modq[6:36] = np.bincount(r30, minlength=30)[:30].astype(np.uint64)
    r4 = (ps & np.uint64(3)).astype(np.int64)
    c1, c3 = int((r4 == 1).sum()), int((r4 == 3).sum())
    large = []
    if n > 1:
        g = np.diff(ps).astype(np.int64)
        gaps += np.bincount(g[g < 100000], minlength=100000)[:100000]
        pb = ps[:-1]
        inner = pb >= np.uint64(7)
        small = inner & (g < 100)
        row30 = np.searchsorted(np.array(W30), r30[:-1][small])
        gs = g[small]
        np.add.at(modq_gap, (np.array(Q6_ROW)[row30], gs), 1)
        np.add.at(modq_gap, (2 + row30, gs), 1)
        for k in np.nonzero(inner & (g > 100))[0]:
            p, gg = int(pb[k]), int(g[k])
            merit = gg / math.log(float(p))
            if (gg >= 500 or merit >= 10.0) and gg < 100000:
                large.append((p + gg, gg, float(np.float32(merit))))
    large.sort(key=lambda r: (-r[2], r[0]))
    return {"total_primes": n, "first_prime": int(ps[0]) if n else 0,
            "last_prime": int(ps[-1]) if n else 0, "c1": c1, "c3": c3, "modq": modq,
            "gaps": gaps, "modq_gap": modq_gap, "large_count": len(large), "large": large}
