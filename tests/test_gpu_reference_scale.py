"""Plain runs at the reference's own published scales, through the C ABI,
against the reference's result files (copied to tests/golden/reference):
results/hist_snapshots/hist_1e12, hist_1e13 (gap histograms), and the full
1e14 output set of results/ (gap_histogram.csv, modq_counts.csv,
modq_gap_hist.csv, chebyshev.csv, hl_trend.csv), with pi and the Chebyshev
counts of README.md:122-127, :305-310.  The 1e14 run uses SEG_NUM = 3e10,
whose 3,333 segment boundaries the reference's modq_gap_hist reflects (the
straddle rule), so every per-residue gap count must match too.  Bit-exact.
"""
import csv
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

REF = os.path.join(GOLDEN, "reference")
SEG_NUM = 30000000000


def read_rows(name):
    with open(os.path.join(REF, name)) as f:
        return list(csv.reader(f))


def read_hist(name):
    return {int(g): int(c) for g, c in read_rows(name)[1:]}


def readme():
    with open(os.path.join(REF, "readme_values.json")) as f:
        return json.load(f)


def check_pi_chebyshev(r, N):
    ref = readme()
    assert r["total_primes"] == ref["pi"][str(N)]
    # the reference's pi(x;4,1), pi(x;4,3) include 5 and 3 (src/voss_v7_live.cu:682)
    assert [r["c1"] + 1, r["c3"] + 1] == ref["chebyshev"][str(N)]


@pytest.mark.parametrize("N,fname", [(10**12, "hist_1000000000000.csv"),
                                     (10**13, "hist_10000000000000.csv")])
def test_plain_matches_reference_histogram(engine, N, fname):
    print(f"[gpu] plain [1, {N:.0e}] vs the reference's {fname}", flush=True)
    r = engine.run_plain(N, SEG_NUM, large_cap=16)
    check_pi_chebyshev(r, N)
    assert {g: int(c) for g, c in enumerate(r["gaps"]) if c} == read_hist(fname)


def test_plain_1e14_matches_reference_outputs(engine):
    N = 10**14
    print("[gpu] plain [1, 1e14] SEG_NUM=3e10 vs the reference's 1e14 results (~35 s)",
          flush=True)
    r = engine.run_plain(N, SEG_NUM, large_cap=16)
    check_pi_chebyshev(r, N)
    gaps = {g: int(c) for g, c in enumerate(r["gaps"]) if c}
    assert gaps == read_hist("gap_histogram_1e14.csv")
    assert gaps == read_hist("hist_100000000000000.csv")
    # modq_counts.csv: q=6 then q=30 residue rows (slot layout of modq_d)
    m = read_rows("modq_counts_1e14.csv")
    assert m[0] == ["q", "a", "count"]
    want = [[6, a, int(r["modq"][a])] for a in range(6)] + \
           [[30, a, int(r["modq"][6 + a])] for a in range(30)]
    assert [[int(x) for x in row] for row in m[1:]] == want
    # modq_gap_hist.csv: in-segment gaps < 100 by residue of p_before
    res = [(6, 1), (6, 5)] + [(30, x) for x in (1, 7, 11, 13, 17, 19, 23, 29)]
    got = [[q, a, g, int(r["modq_gap"][row][g])] for row, (q, a) in enumerate(res)
           for g in range(1, 100) if r["modq_gap"][row][g]]
    mg = read_rows("modq_gap_hist_1e14.csv")
    assert mg[0] == ["q", "residue", "gap", "count"]
    assert [[int(x) for x in row] for row in mg[1:]] == got
    ch = read_rows("chebyshev_1e14.csv")
    assert ch[1] == [str(N), str(r["c1"] + 1), str(r["c3"] + 1), str(r["c3"] - r["c1"])]
    hl = read_rows("hl_trend_1e14.csv")
    assert hl[-1] == [str(N), f"{gaps[6] / gaps[2]:.6f}"]


def test_gap_sequence_1e14_matches_reference_autocorr(engine):
    """The gap_sequence.csv sample of [1, 1e14], SEG_NUM = 3e10 (the middle
    of each of the 3,334 logical segments, src/voss_v7_live.cu:526-547,
    :235-253), against the reference's own summary of its 1e14 sample,
    results/autocorr.json, which scripts/autocorrelation.py computed from
    that file's rows (gap > 0, in index order): sample_size and
    mean_gap_raw exactly (a float64 sum of integers below 2^53 is exact),
    mean_normalized, std_normalized and the lag autocorrelations of
    gap / ln(p) within 1e-9 (numpy's log may differ in the last ulp between
    machines).  Runs right after the 1e14 test on the same context, whose
    per-window prime counts the sampler reuses (otherwise it re-runs the
    1e14 sieve first)."""
    N = 10**14
    with open(os.path.join(REF, "autocorr_1e14.json")) as f:
        ref = json.load(f)
    assert ref["N"] == N
    print("[gpu] gap_sequence sample of [1, 1e14] vs the reference's autocorr.json", flush=True)
    gs, ps, rows = engine.sample_plain(N, SEG_NUM)
    keep = gs > 0
    assert rows == int(keep.sum()) == ref["sample_size"] == 4997666
    gaps = gs[keep].astype(np.float64)
    pos = ps[keep].astype(np.float64)
    assert float(gaps.mean()) == ref["mean_gap_raw"]
    # every row lies in the middle of its logical segment, p_before < p_after
    idx = np.nonzero(keep)[0]
    seg = (ps[keep] - 1) // SEG_NUM
    assert np.array_equal(seg, idx // (5000000 // 3334))
    norm = gaps / np.log(pos)
    assert abs(float(norm.mean()) - ref["mean_normalized"]) < 1e-9
    assert abs(float(norm.std()) - ref["std_normalized"]) < 1e-9
    c = norm - norm.mean()
    var = c.var()
    for lag in (1, 2, 5, 10):
        ac = float((c[:-lag] * c[lag:]).mean() / var)
        assert abs(ac - ref[f"ac_lag{lag}"]) < 1e-9, (lag, ac, ref[f"ac_lag{lag}"])


def test_large_gap_top_k_equals_full_sort(engine):
    """A caller that keeps K of many large gaps gets the selection path
    (per-thread best K, then the best K of those); it must equal the first
    K of the fully sorted list, every field."""
    N = 10**12
    print("[gpu] plain [1, 1e12]: top-32 selection vs full sort", flush=True)
    full = engine.run_plain(N, SEG_NUM, large_cap=1 << 22)
    assert 4 * 65536 < full["large_count"] <= 1 << 22
    assert len(full["large"]) == full["large_count"]
    top = engine.run_plain(N, SEG_NUM, large_cap=32)
    assert top["large_count"] == full["large_count"]
    assert top["large"].tobytes() == full["large"][:32].tobytes()
