"""The batched and serving workloads as configured (workloads.py), through
the C ABI, against checks independent of the HIP sieve:

* the 4096 seeded intervals (starts uniform in [1e12, 1e16), lengths
  log-uniform in [1e4, 1e9]) in one batched launch: the reference's own
  runtime invariants (src/voss_v7_live.cu:693-705, restated per interval)
  on every interval; a seeded sample of short intervals against the
  oracle's interval sieve; a seeded sample of longer ones (up to 3e7
  numbers) against the outputs derived from their primes as the GPU
  Miller-Rabin finds them (tests/from_primes.py, pinned against the
  oracle's sieve);
* the 1e7 Miller-Rabin candidates (half odd uniform in [2^40, 2^64), half
  gap endpoints of that batch) against the oracle's is_prime_u64
  restatement (pinned to the reference's verify_cpp built from source);
* the serving stream (Poisson arrivals, interval requests with start in
  [1e10, 1e16) and length log-uniform in [1e5, 1e10], MR batches of 1e3..1e6
  candidates) through voss_serve_*: every interval result equal to the same
  request run alone, every MR verdict equal to the oracle's.
Integer outputs are compared exactly (tolerance 0)."""
import random
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("total_primes", "first_prime", "last_prime", "c1", "c3", "large_count")


def assert_same(got, ref):
    for k in FIELDS:
        assert got[k] == ref[k], (k, got[k], ref[k])
    assert np.array_equal(got["modq"], ref["modq"])
    assert np.array_equal(got["gaps"], ref["gaps"])
    assert np.array_equal(got["modq_gap"], ref["modq_gap"])
    got_l = got["large"].tolist() if hasattr(got["large"], "tolist") else got["large"]
    assert got_l == [(p, g, float(np.float32(m))) for p, g, m in ref["large"]]


def check_invariants(r, lo, hi):
    """Per interval: sum N(g) = pi - 1, sum g N(g) = last - first,
    pi(x;4,1) + pi(x;4,3) = pi (no 2 in [1e12, ...)), the residue counts
    sum to pi, and the ends lie in [lo, hi)."""
    n = r["total_primes"]
    g = r["gaps"]
    assert int(g.sum()) == max(n - 1, 0)
    assert int((np.arange(len(g), dtype=np.int64) * g).sum()) == \
        (r["last_prime"] - r["first_prime"] if n else 0)
    assert r["c1"] + r["c3"] == n
    assert int(r["modq"][:6].sum()) == n and int(r["modq"][6:36].sum()) == n
    if n:
        assert lo <= r["first_prime"] <= r["last_prime"] < hi
    assert r["large_count"] >= len(r["large"])


@pytest.fixture(scope="module")
def batch(engine):
    import workloads as W
    iv = W.batch_intervals(4096)
    lo = [a for a, _ in iv]
    hi = [b for _, b in iv]
    print("[gpu] the configured batch: 4096 seeded intervals in [1e12, 1e16)", flush=True)
    res, _ = engine.run_batch(lo, hi, large_cap=4096)
    return iv, res


def test_batch_4096_invariants(batch):
    iv, res = batch
    assert len(res) == 4096
    for (a, b), r in zip(iv, res):
        check_invariants(r, a, b)
    assert sum(r["total_primes"] for r in res) > 10**9


def test_batch_4096_sample_vs_oracle(batch, oracle):
    """The shortest intervals (about 1e4..1e5 numbers) of a seeded sample
    against the oracle's interval sieve."""
    iv, res = batch
    rng = random.Random(0x564F5353 + 60)
    short = [k for k, (a, b) in enumerate(iv) if b - a <= 10**5]
    pick = sorted(rng.sample(short, 12))
    assert max(iv[k][0] for k in pick) > 10**15  # base primes beyond 3e7
    for k in pick:
        a, b = iv[k]
        print(f"[gpu] batch interval {k}: [{a}, {b}) vs the oracle", flush=True)
        assert_same(res[k], oracle.run_interval(a, b))


def test_batch_4096_sample_vs_primality(batch, engine):
    """Longer intervals of a seeded sample (1e5 .. 3e7 numbers, bucketed
    large primes among them) against the outputs derived from their primes
    as the GPU Miller-Rabin finds them."""
    from from_primes import interval_results_np
    iv, res = batch
    rng = random.Random(0x564F5353 + 61)
    mid = [k for k, (a, b) in enumerate(iv) if 10**5 < b - a <= 3 * 10**7]
    pick = sorted(rng.sample(mid, 10))
    assert max(iv[k][1] - iv[k][0] for k in pick) > 10**7
    for k in pick:
        a, b = iv[k]
        print(f"[gpu] batch interval {k}: [{a}, {b}) vs its primes", flush=True)
        ns = np.arange(a | 1, b, 2, dtype=np.uint64)
        assert_same(res[k], interval_results_np(ns[engine.mr(ns).astype(bool)]))


def test_mr_1e7_configured_candidates(batch, engine, oracle):
    """The configured Miller-Rabin set: 1e7 candidates, odd indices drawn
    from the large-gap endpoints of the 4096-interval batch."""
    import workloads as W
    _, res = batch
    ends = np.array([x for r in res for p, g, _ in r["large"].tolist() for x in (p, p - g)],
                    dtype=np.uint64)
    assert len(ends) > 1000
    cand = W.mr_candidates(10**7, ends)
    print("[gpu] Miller-Rabin over the configured 1e7 candidates", flush=True)
    v_dev, _ = engine.mr_device(cand)
    want = oracle.is_prime_batch(cand)
    assert np.array_equal(v_dev, want)
    assert int(want[1::2].sum()) == len(cand) // 2  # every endpoint is prime
    assert np.array_equal(engine.mr(cand[:10**6]), want[:10**6])


def test_serving_stream_configured(engine, oracle):
    """60 requests of the configured stream, submitted at their (scaled)
    Poisson arrival times from two client threads: each interval result
    equals the request run alone, each MR verdict the oracle's."""
    import time
    import workloads as W
    reqs = W.serve_requests(60, 200.0)
    rng = np.random.default_rng(W.SEED0 + 4)
    payload = {}
    for k, (_, kind, p) in enumerate(reqs):
        if kind == "mr":
            payload[k] = rng.integers(2**40, 2**63, size=p, dtype=np.uint64) | np.uint64(1)
        else:
            payload[k] = p
    assert max(p[1] - p[0] for _, kind, p in reqs if kind == "interval") > 10**9
    ids = {}
    lock = threading.Lock()
    print("[gpu] serving: the configured stream (60 requests)", flush=True)
    with engine.serve(max_active=64, windows_per_step=4096, large_capacity=64) as srv:
        t0 = time.perf_counter()

        def client(ks):
            for k in ks:
                wait = reqs[k][0] - (time.perf_counter() - t0)
                if wait > 0:
                    time.sleep(wait)
                if reqs[k][1] == "interval":
                    rid = srv.submit_interval(*payload[k])
                else:
                    rid = srv.submit_mr(payload[k])
                with lock:
                    ids[k] = rid

        ts = [threading.Thread(target=client, args=(list(range(t, len(reqs), 2)),))
              for t in range(2)]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        got = {k: srv.wait(rid) for k, rid in ids.items()}
    print("[gpu] serving: each request alone", flush=True)
    for k, (_, kind, _) in enumerate(reqs):
        res, lat = got[k]
        assert lat >= 0
        if kind == "interval":
            a, b = payload[k]
            alone = engine.run_batch([a], [b], large_cap=64)[0][0]
            for f in FIELDS:
                assert res[f] == alone[f], (k, f)
            assert np.array_equal(res["gaps"], alone["gaps"])
            assert np.array_equal(res["modq_gap"], alone["modq_gap"])
            assert np.array_equal(res["modq"], alone["modq"])
            assert res["large"].tolist() == alone["large"].tolist()
            check_invariants(res, a, b)
        else:
            assert np.array_equal(res, oracle.is_prime_batch(payload[k]))
