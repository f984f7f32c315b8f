"""Pins the oracle's gap_sequence sampler (oracle/voss_oracle.c, the block
after the +1 at g=1,2) against an independent restatement of the
reference's sampling written from its source, not from the oracle:

  host logic   src/voss_v7_live.cu:526-547  per_seg = GAP_SEQ_SIZE / NUM_SEG;
               per segment with n_pos > 1: start = n_pos / 2,
               sample_n = min(n_pos - start - 1, per_seg), written when
               sample_n > 1 and seg * per_seg + sample_n <= GAP_SEQ_SIZE
  kernel       src/voss_v7_live.cu:235-253  row i of the segment's slot
               range holds (v[start+i+1] - v[start+i], v[start+i]) over the
               segment's sorted primes >= 7
  writer       src/voss_v7_live.cu:785-803  rows with gap > 0, by index

Here the segment's primes come from a numpy sieve of [1, N], so an
off-by-one in start, in sample_n or in the per-segment offset of either
side fails.  The 1e14 sample itself is pinned on the GPU against the
reference's results/autocorr.json (tests/test_gpu_reference_scale.py)."""
import numpy as np
import pytest

GAP_SEQ_SIZE = 5000000


def sieve(n):
    s = np.ones(n + 1, dtype=bool)
    s[:2] = False
    for i in range(2, int(n**0.5) + 1):
        if s[i]:
            s[i * i::i] = False
    return s


def brute_sample(N, SEG):
    """gap_seq, pos_seq (GAP_SEQ_SIZE each, zero where nothing is written)."""
    s = sieve(N)
    num_seg = (N + SEG - 1) // SEG
    per_seg = GAP_SEQ_SIZE // num_seg
    gap_seq = np.zeros(GAP_SEQ_SIZE, dtype=np.int32)
    pos_seq = np.zeros(GAP_SEQ_SIZE, dtype=np.uint64)
    if per_seg <= 0:
        return gap_seq, pos_seq
    for seg in range(num_seg):
        lo, hi = seg * SEG + 1, min((seg + 1) * SEG, N)
        v = np.nonzero(s[lo:hi + 1])[0] + lo
        v = v[v >= 7]  # the wheel bitmap holds the primes >= 7 only
        n_pos = len(v)
        if n_pos <= 1:
            continue
        start = n_pos // 2
        sample_n = min(n_pos - start - 1, per_seg)
        if sample_n <= 1:
            continue
        off = seg * per_seg
        if off + sample_n > GAP_SEQ_SIZE:
            continue
        w = v[start:start + sample_n + 1].astype(np.int64)
        gap_seq[off:off + sample_n] = np.diff(w)
        pos_seq[off:off + sample_n] = w[:-1]
    return gap_seq, pos_seq


@pytest.mark.parametrize("N,SEG", [
    (10**6, 30000000000),   # one logical segment
    (10**6, 30000),         # 34 segments, every sample shorter than per_seg
    (2 * 10**6 + 17, 90000),  # ragged last segment
    (100000, 210),          # 477 segments of a few primes: sample_n <= 1 skips
    (3 * 10**6, 2010),      # per_seg = 3350 > what a segment holds
    (4 * 10**5, 60),        # segments of 0..3 primes
])
def test_oracle_sampler_matches_brute_force(oracle, N, SEG):
    gs, ps = brute_sample(N, SEG)
    r = oracle.run_plain(N, SEG, sample=True)
    assert np.array_equal(r["gap_seq"], gs)
    assert np.array_equal(r["pos_seq"], ps)
    assert r["sample_rows"] == int((gs > 0).sum())


def test_sampler_full_segments_row_count(oracle):
    """results/autocorr.json's sample_size at 1e14 is NUM_SEG * per_seg =
    3,334 * 1,499 = 4,997,666: every segment writes exactly per_seg rows
    once it holds more than 2 * per_seg + 1 primes.  [1, 3e8] with
    SEG_NUM = 3e7 is such a case (10 segments of ~1.6e6 primes, per_seg =
    5e5): both restatements fill all 5e6 rows, and agree row for row."""
    N, SEG = 3 * 10**8, 3 * 10**7
    gs, ps = brute_sample(N, SEG)
    assert int((gs > 0).sum()) == 10 * (GAP_SEQ_SIZE // 10) == GAP_SEQ_SIZE
    r = oracle.run_plain(N, SEG, sample=True)
    assert r["sample_rows"] == GAP_SEQ_SIZE
    assert np.array_equal(r["gap_seq"], gs)
    assert np.array_equal(r["pos_seq"], ps)
    assert (10**14 + 30000000000 - 1) // 30000000000 == 3334
    assert 3334 * (GAP_SEQ_SIZE // 3334) == 4997666
