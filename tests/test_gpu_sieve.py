"""GPU sieve parity: the gfx950 library (through the C ABI) against the CPU
oracle on the same ranges, and against the committed 1e11 fixtures.
Integer outputs are compared bit-exactly; merits are the same float32."""
import json
import math
import os
import random

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

FIELDS = ("total_primes", "first_prime", "last_prime", "c1", "c3", "large_count")


def assert_same(got, ref, large=True):
    for k in FIELDS:
        assert got[k] == ref[k], (k, got[k], ref[k])
    assert np.array_equal(got["modq"], ref["modq"])
    assert np.array_equal(got["gaps"], ref["gaps"])
    assert np.array_equal(got["modq_gap"], ref["modq_gap"])
    if large:
        got_l = got["large"].tolist() if hasattr(got["large"], "tolist") else got["large"]
        assert got_l == [(p, g, float(np.float32(m))) for p, g, m in ref["large"]]


@pytest.mark.parametrize("N", [5, 6, 7, 10, 30, 31, 100, 1000, 12345, 10**6, 10**7 + 19,
                               10**8, 10**9])
def test_plain_matches_oracle(engine, oracle, N):
    print(f"[gpu] plain N={N}", flush=True)
    assert_same(engine.run_plain(N, 30000000000), oracle.run_plain(N, 30000000000))


@pytest.mark.parametrize("N,SEG", [(10**6, 30030), (10**8, 3000000), (2 * 10**7, 4000020),
                                   (10**5, 30), (10**9, 30000000), (123456789, 3932160)])
def test_plain_segment_rule_matches_oracle(engine, oracle, N, SEG):
    print(f"[gpu] plain N={N} SEG_NUM={SEG}", flush=True)
    assert_same(engine.run_plain(N, SEG), oracle.run_plain(N, SEG))


def test_plain_1e11_fixture(engine):
    """N = 1e11, SEG_NUM = 3e10: the reference's histogram snapshot and the
    oracle fixture for every other output."""
    print("[gpu] plain N=1e11", flush=True)
    fx = json.load(open(os.path.join(GOLDEN, "oracle_plain_1e11.json")))
    r = engine.run_plain(10**11, 30000000000, large_cap=1 << 20)
    for k in FIELDS:
        assert r[k] == fx[k], k
    assert [int(x) for x in r["modq"]] == fx["modq"]
    assert {str(g): int(c) for g, c in enumerate(r["gaps"]) if c} == fx["gaps"]
    assert {f"{row},{g}": int(c) for row in range(10) for g, c in enumerate(r["modq_gap"][row])
            if c} == fx["modq_gap"]
    import hashlib
    h = hashlib.sha256()
    for p, g, m in r["large"].tolist():
        h.update(f"{p},{g},{m:.6f}\n".encode())
    assert h.hexdigest() == fx["large_sha256"]
    # and the reference's own snapshot of the histogram at 1e11
    import csv
    with open(os.path.join(GOLDEN, "reference", "hist_100000000000.csv")) as f:
        ref = {int(a): int(b) for a, b in list(csv.reader(f))[1:]}
    assert {g: int(c) for g, c in enumerate(r["gaps"]) if c} == ref


@pytest.mark.parametrize("window", [65536, 32768])
def test_window_sizes_agree(engine, oracle, window):
    print(f"[gpu] window={window}", flush=True)
    ref = oracle.run_plain(10**9 + 7, 300000000)
    old = engine.tuning()
    try:
        engine.set_tuning(window_bytes=window, wg_per_cu={65536: 2, 32768: 3}[window])
        assert_same(engine.run_plain(10**9 + 7, 300000000), ref)
    finally:
        engine.set_tuning(**old)


def test_tier_split_agrees(engine, oracle):
    ref = oracle.run_plain(10**8, 30000000000)
    old = engine.tuning()
    try:
        for lim in (80, 4096, 200000):
            engine.set_tuning(tier_a_limit=lim)
            assert_same(engine.run_plain(10**8, 30000000000), ref)
    finally:
        engine.set_tuning(**old)


INTERVALS = [(0, 100), (1, 2), (2, 3), (2, 4), (3, 8), (5, 6), (7, 7), (7, 8), (0, 10**6),
             (10**9, 10**9 + 10**7), (10**12, 10**12 + 10**6), (2**40, 2**40 + 5 * 10**6),
             (10**15, 10**15 + 10**5), (10**16 - 10**5, 10**16 + 10**5),
             (25056082000, 25056083000)]


@pytest.mark.parametrize("lo,hi", INTERVALS)
def test_interval_matches_oracle(engine, oracle, lo, hi):
    print(f"[gpu] interval [{lo},{hi})", flush=True)
    assert_same(engine.run_interval(lo, hi), oracle.run_interval(lo, hi))


def seeded_intervals(n, seed, lo_min=10**9, lo_max=10**13, lmin=3, lmax=7):
    rng = random.Random(seed)
    out = []
    for _ in range(n):
        lo = rng.randrange(lo_min, lo_max)
        ln = int(10 ** (lmin + (lmax - lmin) * rng.random()))
        out.append((lo, lo + ln))
    return out


def test_large_batch_host_paths(engine, oracle):
    """2,500 intervals: enough for the multi-threaded host paths (planning
    by job ranges, per-job accumulation, per-interval results).  Every
    record equals the per-interval form's (planned and finished one
    interval at a time) byte for byte; a sample also equals the oracle."""
    import ctypes
    import vossgpu
    iv = seeded_intervals(2500, 0x564F5353 + 21, lo_min=10**9, lo_max=10**14, lmin=3, lmax=6.3)
    iv += [(0, 1000), (10**6, 10**6), (7, 8), (2, 3)]
    lo = [a for a, _ in iv]
    hi = [b for _, b in iv]
    print(f"[gpu] batch of {len(iv)} (multi-threaded host paths)", flush=True)
    rb, lb, _ = engine.run_batch(lo, hi, large_cap=32, raw=True)
    rp, lp, _ = engine.run_batch_plain(lo, hi, large_cap=32, raw=True)
    off = vossgpu.IntervalResult.large.offset
    for k in range(len(iv)):
        assert bytes(rb[k])[:off] == bytes(rp[k])[:off], (k, iv[k])
        n = min(rb[k].large_count, 32)
        assert vossgpu.large_array(rb[k].large, n).tobytes() == \
            vossgpu.large_array(rp[k].large, n).tobytes(), (k, iv[k])
    for k in list(range(0, len(iv), 97)) + list(range(len(iv) - 4, len(iv))):
        assert_same(vossgpu.interval_to_dict(rb[k]), oracle.run_interval(*iv[k]), large=False)


def test_batch_matches_single_and_oracle(engine, oracle):
    iv = seeded_intervals(48, 0x564F5353 + 1)
    iv += [(0, 1000), (10**6, 10**6), (7, 8)]
    lo = [a for a, _ in iv]
    hi = [b for _, b in iv]
    print(f"[gpu] batch of {len(iv)}", flush=True)
    batch, _ = engine.run_batch(lo, hi, large_cap=4096)
    single, _ = engine.run_batch_plain(lo, hi, large_cap=4096)
    for (a, b), rb, rs in zip(iv, batch, single):
        ref = oracle.run_interval(a, b)
        assert_same(rb, ref)
        assert_same(rs, ref)


def test_invariants_interval_large(engine):
    """Size-independent checks at a config-3-like offset: sum N(g) = pi - 1,
    sum g N(g) = p_last - p_first, Chebyshev split adds up."""
    lo, hi = 10**13, 10**13 + 10**10
    print(f"[gpu] interval [{lo},{hi}) invariants", flush=True)
    r = engine.run_interval(lo, hi)
    g = r["gaps"]
    assert int(g.sum()) == r["total_primes"] - 1
    assert int((np.arange(len(g), dtype=np.int64) * g).sum()) == r["last_prime"] - r["first_prime"]
    assert r["c1"] + r["c3"] == r["total_primes"]
    # prime number theorem sanity: within 1% of li-density estimate
    est = (hi - lo) / math.log((lo + hi) / 2)
    assert abs(r["total_primes"] / est - 1) < 0.01


def test_bad_arguments(engine):
    import vossgpu
    with pytest.raises(vossgpu.VossError):
        engine.run_plain(4, 30)
    with pytest.raises(vossgpu.VossError):
        engine.run_plain(1000, 31)
    with pytest.raises(vossgpu.VossError):
        engine.run_interval(10, 5)


def test_interval_1e13_fixture(engine):
    """[1e13, 1e13 + 1e12): base primes to 3.3e6, 254k windows; every output
    against the oracle fixture (sha256 over the full large-gap list)."""
    import hashlib
    print("[gpu] interval [1e13, 1.1e13)", flush=True)
    fx = json.load(open(os.path.join(GOLDEN, "oracle_interval_1e13_1e12.json")))
    r = engine.run_interval(fx["lo"], fx["hi"], large_cap=1 << 22)
    for k in FIELDS:
        assert r[k] == fx[k], k
    assert [int(x) for x in r["modq"]] == fx["modq"]
    assert {str(g): int(c) for g, c in enumerate(r["gaps"]) if c} == fx["gaps"]
    assert {f"{row},{g}": int(c) for row in range(10) for g, c in enumerate(r["modq_gap"][row])
            if c} == fx["modq_gap"]
    h = hashlib.sha256()
    for p, g, m in r["large"].tolist():
        h.update(f"{p},{g},{m:.6f}\n".encode())
    assert h.hexdigest() == fx["large_sha256"]


@pytest.fixture
def buckets_forced(engine):
    """Every prime >= the window length through HBM bucket lists, for every
    job regardless of its length."""
    old = engine.tuning()
    engine.set_tuning(bucket_limit=131072, bucket_min_windows=1)
    yield engine
    engine.set_tuning(**{k: v for k, v in old.items()})


@pytest.mark.parametrize("lo,hi", [(10**12, 10**12 + 10**9), (10**15, 10**15 + 10**8),
                                   (10**16 - 3 * 10**7, 10**16 + 10**7), (10**11, 10**11 + 10**5)])
def test_buckets_interval_matches_oracle(buckets_forced, oracle, lo, hi):
    print(f"[gpu] buckets interval [{lo},{hi})", flush=True)
    assert_same(buckets_forced.run_interval(lo, hi), oracle.run_interval(lo, hi))


def test_buckets_plain_1e11_fixture(buckets_forced):
    """25,435 windows in 4 bucket tiles, primes 131,101..316,223 bucketed."""
    print("[gpu] buckets plain 1e11", flush=True)
    fx = json.load(open(os.path.join(GOLDEN, "oracle_plain_1e11.json")))
    r = buckets_forced.run_plain(10**11, 30000000000)
    for k in FIELDS:
        assert r[k] == fx[k], k
    assert {str(g): int(c) for g, c in enumerate(r["gaps"]) if c} == fx["gaps"]
    assert {f"{row},{g}": int(c) for row in range(10) for g, c in enumerate(r["modq_gap"][row])
            if c} == fx["modq_gap"]


def test_buckets_batch_matches_oracle(buckets_forced, oracle):
    iv = seeded_intervals(24, 0x564F5353 + 9, lo_min=10**13, lo_max=10**15, lmin=5, lmax=8)
    lo = [a for a, _ in iv]
    hi = [b for _, b in iv]
    print(f"[gpu] buckets batch of {len(iv)}", flush=True)
    batch, _ = buckets_forced.run_batch(lo, hi, large_cap=4096)
    for (a, b), rb in zip(iv, batch):
        assert_same(rb, oracle.run_interval(a, b))


@pytest.mark.parametrize("N,SEG", [(10**9, 300000000), (10**8, 30000000000), (123456789, 3932160)])
def test_gap_sequence_sample_matches_oracle(engine, oracle, N, SEG):
    """gap_sequence.csv rows (src/voss_v7_live.cu:526-547, :785-803)."""
    print(f"[gpu] sample N={N} SEG={SEG}", flush=True)
    gs, ps, rows = engine.sample_plain(N, SEG)
    ref = oracle.run_plain(N, SEG, sample=True)
    assert rows == ref["sample_rows"]
    assert np.array_equal(gs, ref["gap_seq"])
    assert np.array_equal(ps, ref["pos_seq"])


def test_plain_range_pieces_compose(engine):
    """Pieces of logical segments joined by a main-histogram-only pair equal
    the whole run (what the CLI does at checkpoints)."""
    N, SEG = 10**9, 30000000
    whole = engine.run_plain(N, SEG)
    nseg = (N + SEG - 1) // SEG
    cuts = [0, 1, 5, 10, 17, nseg]
    gaps = np.zeros_like(whole["gaps"])
    mq = np.zeros_like(whole["modq_gap"])
    modq = np.zeros_like(whole["modq"])
    total, last, c1, c3, nlarge = 3, 5, 0, 0, 0
    for a, b in zip(cuts, cuts[1:]):
        r = engine.run_plain_range(N, SEG, a, b)
        if r["total_primes"]:
            gaps[r["first_prime"] - last] += 1
            last = r["last_prime"]
            total += r["total_primes"]
        gaps += r["gaps"]
        mq += r["modq_gap"]
        modq += r["modq"]
        c1 += r["c1"]
        c3 += r["c3"]
        nlarge += r["large_count"]
    gaps[1] += 1
    gaps[2] += 1
    assert np.array_equal(gaps, whole["gaps"])
    assert np.array_equal(mq, whole["modq_gap"])
    assert np.array_equal(modq, whole["modq"])
    assert (total, last, c1, c3, nlarge) == (whole["total_primes"], whole["last_prime"],
                                             whole["c1"], whole["c3"], whole["large_count"])


def test_large_gap_truncation_is_canonical(engine, oracle):
    """A caller capacity smaller than the number of large gaps keeps the
    first K in canonical order (merit desc, position asc) in every form."""
    lo, hi = 10**12, 10**12 + 2 * 10**9
    print("[gpu] truncation K=8", flush=True)
    ref = oracle.run_interval(lo, hi)
    assert ref["large_count"] > 100
    top = [(p, g, float(np.float32(m))) for p, g, m in ref["large"][:8]]
    single = engine.run_interval(lo, hi, large_cap=8)
    assert single["large"].tolist() == top and single["large_count"] == ref["large_count"]
    lo2, hi2 = [lo, 10**13], [hi, 10**13 + 10**9]
    (b0, b1), _ = engine.run_batch(lo2, hi2, large_cap=8)
    (p0, p1), _ = engine.run_batch_plain(lo2, hi2, large_cap=8)
    assert b0["large"].tolist() == top == p0["large"].tolist()
    assert b1["large"].tolist() == p1["large"].tolist()


def test_autotune_installs_a_valid_tuning_and_keeps_results(engine, oracle):
    """The autotuner only changes timing: the batch after tuning is the
    oracle's, and the installed tuning is one of the searched shapes."""
    iv = seeded_intervals(32, 0x564F5353 + 11, lo_min=10**12, lo_max=10**15, lmin=4, lmax=8)
    lo = [a for a, _ in iv]
    hi = [b for _, b in iv]
    old = engine.tuning()
    try:
        print("[gpu] autotune on 32 intervals (sample 16)", flush=True)
        t, sec = engine.autotune(lo, hi, max_intervals=16, repeats=1)
        assert sec > 0
        assert (t["window_bytes"], t["wg_per_cu"]) in {(131072, 1), (65536, 2), (65536, 1),
                                                       (32768, 3)}
        assert t["bucket_limit"] >= t["window_bytes"]
        assert engine.tuning() == t
        batch, _ = engine.run_batch(lo, hi, large_cap=4096)
        for (a, b), rb in zip(iv, batch):
            assert_same(rb, oracle.run_interval(a, b))
    finally:
        engine.set_tuning(**old)


def _mixed_batch(seed):
    """Short and long intervals, the long ones through bucket lists."""
    iv = seeded_intervals(20, seed, lo_min=10**12, lo_max=10**15, lmin=4, lmax=7)
    iv += seeded_intervals(6, seed + 1, lo_min=10**14, lo_max=10**16, lmin=7.2, lmax=8.3)
    iv += [(0, 1000), (10**6, 10**6), (7, 8)]
    return [a for a, _ in iv], [b for _, b in iv]


def test_graph_mode_matches_eager_and_oracle(engine, oracle):
    lo, hi = _mixed_batch(0x564F5353 + 21)
    print(f"[gpu] graph vs eager batch of {len(lo)}", flush=True)
    try:
        engine.set_graph_mode(True)
        g, gst = engine.run_batch(lo, hi, large_cap=4096)
        engine.set_graph_mode(False)
        e, est = engine.run_batch(lo, hi, large_cap=4096)
    finally:
        engine.set_graph_mode(True)
    assert gst["sieve_ms"] == -1 and est["sieve_ms"] > 0
    assert gst["bucket_entries"] > 0
    for a, b, rg, re_ in zip(lo, hi, g, e):
        assert_same(rg, re_)
    for k in range(0, len(lo), 4):
        assert_same(g[k], oracle.run_interval(lo[k], hi[k]))


def test_prepared_batch_replays(engine):
    """Three replays of one captured graph, with other eager work and
    synchronisations between them; each equals the eager batch."""
    lo, hi = _mixed_batch(0x564F5353 + 23)
    engine.set_graph_mode(False)
    try:
        ref, _ = engine.run_batch(lo, hi, large_cap=512)
    finally:
        engine.set_graph_mode(True)
    with engine.prepare_batch(lo, hi) as pb:
        for rep in range(3):
            print(f"[gpu] prepared batch replay {rep}", flush=True)
            got, st = pb.run(large_cap=512)
            assert st["device_ms"] > 0
            for rg, rr in zip(got, ref):
                assert_same(rg, rr)
            # eager work on the same context between replays
            engine.run_interval(10**13 + rep * 10**9, 10**13 + rep * 10**9 + 10**8)
            engine.run_batch(lo[:3], hi[:3])


def test_prepared_batch_rejects_changed_tuning(engine):
    old = engine.tuning()
    with engine.prepare_batch([10**12], [10**12 + 10**6]) as pb:
        try:
            engine.set_tuning(tier_a_limit=old["tier_a_limit"] * 2)
            with pytest.raises(Exception, match="tuning changed"):
                pb.run()
        finally:
            engine.set_tuning(**old)
        got, _ = pb.run()
        assert got[0]["total_primes"] > 0


@pytest.mark.parametrize("graph", [True, False])
def test_bucket_overflow_reruns_exactly(engine, oracle, graph):
    """Rooms at 50% of the estimate: every bucket overflows, the fill kernel
    flags it, and the run repeats with exact counts.  Compared with the
    default run (pinned to the oracle by the graph-mode test) and, for two
    long intervals, with the oracle itself."""
    lo, hi = _mixed_batch(0x564F5353 + 25)
    ref, _ = engine.run_batch(lo, hi, large_cap=4096)
    old = engine.tuning()
    try:
        engine.set_graph_mode(graph)
        engine.set_tuning(bucket_room_pct=50)
        print(f"[gpu] bucket overflow rerun graph={graph}", flush=True)
        got, st = engine.run_batch(lo, hi, large_cap=4096)
    finally:
        engine.set_tuning(**old)
        engine.set_graph_mode(True)
    assert st["sieve_ms"] > 0  # the exact re-run is timed per kernel
    for rg, rr in zip(got, ref):
        assert_same(rg, rr)
    longest = sorted(range(len(lo)), key=lambda k: hi[k] - lo[k])[-2:]
    for k in longest:
        assert_same(got[k], oracle.run_interval(lo[k], hi[k]))


def test_prepared_batch_with_mr(engine):
    """A prepared batch with two Miller-Rabin batches on device buffers:
    every replay gives the interval results of the eager batch and the
    verdicts of voss_mr_batch; a batch with MR work only also runs."""
    import ctypes
    import vossgpu
    L = vossgpu.lib()
    lo, hi = _mixed_batch(0x564F5353 + 27)
    ref, _ = engine.run_batch(lo, hi, large_cap=256)
    rng = np.random.default_rng(0x564F5353 + 28)
    c1 = (rng.integers(1 << 40, 1 << 63, size=20000, dtype=np.uint64) * 2 + 1).astype(np.uint64)
    c2 = np.array([2, 3, 4, 37, 41, 3825123056546413051, 2**61 - 1, 2**64 - 59,
                   (2**32 - 5) * (2**32 - 17)], dtype=np.uint64)
    want = [engine.mr(c1), engine.mr(c2)]
    bufs = []
    try:
        for c in (c1, c2):
            dn, dv = ctypes.c_void_p(), ctypes.c_void_p()
            vossgpu._check(L.voss_device_alloc(engine._h, c.nbytes, ctypes.byref(dn)))
            vossgpu._check(L.voss_device_alloc(engine._h, max(c.size, 8), ctypes.byref(dv)))
            vossgpu._check(L.voss_memcpy_h2d(engine._h, dn, c.ctypes.data, c.nbytes))
            bufs.append((dn, dv, c.size))
        with engine.prepare_batch(lo, hi) as pb:
            for dn, dv, n in bufs:
                pb.add_mr(dn, n, dv)
            for rep in range(3):
                print(f"[gpu] prepared batch + MR replay {rep}", flush=True)
                for dn, dv, n in bufs:  # stale verdicts must be overwritten
                    z = np.full(n, 7, dtype=np.uint8)
                    vossgpu._check(L.voss_memcpy_h2d(engine._h, dv, z.ctypes.data, n))
                got, _ = pb.run(large_cap=256)
                for rg, rr in zip(got, ref):
                    assert_same(rg, rr)
                for (dn, dv, n), w in zip(bufs, want):
                    v = np.zeros(n, dtype=np.uint8)
                    vossgpu._check(L.voss_memcpy_d2h(engine._h, v.ctypes.data, dv, n))
                    assert np.array_equal(v, w)
                engine.mr(c2[:3])  # eager work between replays
        with engine.prepare_batch([], []) as pb:
            dn, dv, n = bufs[1]
            pb.add_mr(dn, n, dv)
            pb.run()
            v = np.zeros(n, dtype=np.uint8)
            vossgpu._check(L.voss_memcpy_d2h(engine._h, v.ctypes.data, dv, n))
            assert np.array_equal(v, want[1])
    finally:
        for dn, dv, _ in bufs:
            L.voss_device_free(engine._h, dn)
            L.voss_device_free(engine._h, dv)


def test_server_owned_context_refuses_direct_calls(engine):
    """While a server runs, its context answers direct calls with VOSS_EBUSY
    (no device work), requests still complete, and after the stop the
    context is usable again with the same results."""
    import vossgpu
    lo, hi = 10**9, 10**9 + 10**6
    with engine.serve(max_active=4, windows_per_step=64, large_capacity=64) as srv:
        print("[gpu] direct call on a served context", flush=True)
        with pytest.raises(vossgpu.VossError) as e:
            engine.run_plain(10**6, 30000000000)
        assert e.value.code == -5
        with pytest.raises(vossgpu.VossError) as e:
            engine.mr(np.array([7, 9], dtype=np.uint64))
        assert e.value.code == -5
        with pytest.raises(vossgpu.VossError) as e:
            engine.serve(max_active=2)  # one server per context
        assert e.value.code == -5
        served, _ = srv.wait(srv.submit_interval(lo, hi))
    alone = engine.run_batch([lo], [hi], large_cap=64)[0][0]
    for k in ("total_primes", "first_prime", "last_prime", "c1", "c3", "large_count"):
        assert served[k] == alone[k]
    assert np.array_equal(served["gaps"], alone["gaps"])


def test_second_thread_is_refused_while_a_call_runs(engine):
    """A call on another context from a second thread while one thread's
    call runs returns VOSS_EBUSY at once, before it touches the device; the
    running call is unaffected, and afterwards the second context works.
    The two threads never run device work at the same time."""
    import threading
    import time
    import vossgpu
    other = vossgpu.Engine(0)
    try:
        want = other.run_plain(10**7, 30000000000)
        res = {}

        def long_call():
            res["r"] = engine.run_plain(10**13, 30000000000, large_cap=16)  # ~2.2 s

        print("[gpu] second thread refused while a 1e13 run holds the library", flush=True)
        t = threading.Thread(target=long_call)
        t.start()
        time.sleep(0.5)
        code = None
        try:
            other.run_plain(10**7, 30000000000)
        except vossgpu.VossError as e:
            code = e.code
        t.join()
        assert code == -5
        assert res["r"]["total_primes"] == 346065536839
        got = other.run_plain(10**7, 30000000000)
        assert got["total_primes"] == want["total_primes"] == 664579
        assert np.array_equal(got["gaps"], want["gaps"])
    finally:
        other.close()


@pytest.mark.skip(reason="the run that stalled (two contexts driven from two host threads with "
                         "device work overlapping): such calls are now refused with VOSS_EBUSY "
                         "by the library's call gate, so this body would see refusals; kept as "
                         "the record of the stalled case")
def test_concurrent_contexts_share_the_host_pool(engine):
    """Two contexts driven from two host threads at once: their planning,
    accumulation and result steps meet in the one host worker pool (the
    caller that finds it busy runs its chunks itself).  Each thread's
    records equal the same calls made alone, byte for byte."""
    import ctypes
    import threading
    import vossgpu
    iv = seeded_intervals(2500, 0x564F5353 + 22, lo_min=10**9, lo_max=10**14, lmin=3, lmax=6.3)
    lo = [a for a, _ in iv]
    hi = [b for _, b in iv]
    off = vossgpu.IntervalResult.large.offset

    def records(rs):
        return [bytes(r)[:off] + vossgpu.large_array(r.large, min(r.large_count, 32)).tobytes()
                for r in rs]

    def plain_key(d):
        return (d["total_primes"], d["gaps"].tobytes(), d["large_count"])

    print("[gpu] concurrent contexts: reference calls", flush=True)
    want_batch = records(engine.run_batch(lo, hi, large_cap=32, raw=True)[0])
    want_plain = plain_key(engine.run_plain(10**10, 30000000000, large_cap=64))
    other = vossgpu.Engine(0)
    got, errors = {}, []

    def batches(eng, tag):
        try:
            for rep in range(3):
                got[(tag, rep)] = records(eng.run_batch(lo, hi, large_cap=32, raw=True)[0])
        except Exception as e:  # reported below, not swallowed
            errors.append(repr(e))

    def plains(eng, tag):
        try:
            for rep in range(3):
                got[(tag, rep)] = plain_key(eng.run_plain(10**10, 30000000000, large_cap=64))
        except Exception as e:
            errors.append(repr(e))

    print("[gpu] concurrent contexts: two threads", flush=True)
    try:
        ts = [threading.Thread(target=batches, args=(engine, "b0")),
              threading.Thread(target=batches, args=(other, "b1"))]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        ts = [threading.Thread(target=batches, args=(engine, "b2")),
              threading.Thread(target=plains, args=(other, "p"))]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
    finally:
        other.close()
    assert not errors, errors
    for key, v in got.items():
        assert v == (want_plain if key[0] == "p" else want_batch), key


def test_interval_at_the_2_62_bound():
    """[2^62 - 3e5, 2^62): the largest hi the library accepts, base primes up
    to 2^31.  The oracle's sieve takes minutes there, so the expected outputs
    come from the interval's primes as the GPU Miller-Rabin finds them (pinned
    on its own against a sieve and the oracle) through tests/from_primes.py
    (pinned against the oracle's sieve).  A context of its own, so the
    session's does not keep 1.7 GB of base-prime records."""
    import vossgpu
    from from_primes import interval_results
    lo, hi = 2**62 - 300000, 2**62
    other = vossgpu.Engine(0)
    try:
        ns = np.uint64(lo + 1) + np.arange(0, (hi - lo) // 2, dtype=np.uint64) * np.uint64(2)
        print("[gpu] primes of the interval below 2^62 by Miller-Rabin", flush=True)
        want = interval_results(ns[other.mr(ns).astype(bool)])
        print("[gpu] interval below 2^62 (base primes to 2^31)", flush=True)
        got = other.run_interval(lo, hi, large_cap=4096)
    finally:
        other.close()
    assert want["total_primes"] > 6000
    assert_same(got, want)


def test_reciprocal_error_bound_of_the_tile_remainder():
    """The bucket fill's double-precision remainder (start_mod) takes the
    quotient within 1 of X/p < 2^47 only if 1/p (recip_newton: v_rcp_f64 and
    one Newton step) is accurate to 2^-48 relative: checked on the device for
    every odd p in [3, 2^30) (libvoss_probe.so)."""
    import ctypes
    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                        "voss-prime-gaps_v1_amd", "lib", "libvoss_probe.so")
    L = ctypes.CDLL(path)
    L.voss_probe_recip.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(ctypes.c_double)]
    err = ctypes.c_double()
    print("[gpu] recip_newton over the odd p below 2^30", flush=True)
    assert L.voss_probe_recip(3, 1 << 30, ctypes.byref(err)) == 0
    print("max relative error 2^%.2f" % math.log2(err.value), flush=True)
    assert 0.0 < err.value < 2.0 ** -48, err.value


def test_interval_below_2_60_double_precision_remainders():
    """[2^60 - 8e6, 2^60): three LDS windows whose bucket primes (2^20 up to
    just below 2^30) take the bucket fill's double-precision remainder at its
    largest quotient (X/p up to 2^40 per tile); expected outputs from the
    interval's primes as the GPU Miller-Rabin finds them
    (tests/from_primes.py)."""
    import vossgpu
    from from_primes import interval_results
    lo, hi = 2**60 - 8 * 10**6, 2**60
    other = vossgpu.Engine(0)
    try:
        ns = np.uint64(lo + 1) + np.arange(0, (hi - lo) // 2, dtype=np.uint64) * np.uint64(2)
        print("[gpu] primes of the interval below 2^60 by Miller-Rabin", flush=True)
        want = interval_results(ns[other.mr(ns).astype(bool)])
        print("[gpu] interval below 2^60 (bucket primes below 2^30)", flush=True)
        got = other.run_interval(lo, hi, large_cap=4096)
    finally:
        other.close()
    assert want["total_primes"] > 150000
    assert_same(got, want)


def test_batch_of_high_intervals_matches_primality():
    """A batch of 12 intervals between 1e17 and 2^62 (six of 1e5 numbers, six
    of 8e6: three LDS windows each, so their primes from 2^20 to 2^31 go
    through HBM bucket lists), each against the outputs derived from its
    primes as the GPU Miller-Rabin finds them (tests/from_primes.py)."""
    import random as _random
    import vossgpu
    from from_primes import interval_results
    rng = _random.Random(0x564F5353 + 50)
    iv = []
    for k in range(12):
        ln = 10**5 if k % 2 == 0 else 8 * 10**6
        lo = rng.randrange(10**17, 2**62 - ln)
        iv.append((lo, lo + ln))
    other = vossgpu.Engine(0)
    try:
        print("[gpu] batch of 12 intervals in [1e17, 2^62)", flush=True)
        got, _ = other.run_batch([a for a, _ in iv], [b for _, b in iv], large_cap=4096)
        for (a, b), g in zip(iv, got):
            ns = np.uint64(a | 1) + np.arange(0, (b - (a | 1) + 1) // 2, dtype=np.uint64) * np.uint64(2)
            assert int(ns[-1]) < b
            want = interval_results(ns[other.mr(ns).astype(bool)])
            assert want["total_primes"] > 2000
            assert_same(g, want)
    finally:
        other.close()
