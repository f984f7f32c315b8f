"""Continuous-batching server, both schedulers (the device-resident loop and
the host-stepped one): requests that overlap on the device give exactly the
results of the same request run alone (tolerance 0)."""
import threading

import numpy as np
import pytest

import vossgpu

pytestmark = pytest.mark.gpu

MODES = [vossgpu.SERVE_DEVICE_LOOP, vossgpu.SERVE_HOST_STEPS]

LARGE = ("total_primes", "first_prime", "last_prime", "c1", "c3", "large_count")


def same(a, b):
    for k in LARGE:
        assert a[k] == b[k], (k, a[k], b[k])
    assert np.array_equal(a["gaps"], b["gaps"])
    assert np.array_equal(a["modq_gap"], b["modq_gap"])
    assert np.array_equal(a["modq"], b["modq"])
    assert a["large"].tolist() == b["large"].tolist()


def workload():
    rng = np.random.default_rng(0x564F5353 + 30)
    iv = [(0, 100), (7, 8), (10**9, 10**9 + 10**6), (10**13, 10**13 + 3 * 10**8),
          (10**15, 10**15 + 10**8), (10**16 - 10**7, 10**16 + 10**7)]
    for _ in range(14):
        lo = int(rng.integers(10**10, 10**15))
        iv.append((lo, lo + int(10 ** rng.uniform(4, 8.5))))
    mr = [rng.integers(2**40, 2**63, size=int(10 ** rng.uniform(2, 5)), dtype=np.uint64) | np.uint64(1)
          for _ in range(8)]
    mr.append(np.array([2, 3, 561, 2**61 - 1, 3825123056546413051], dtype=np.uint64))
    return iv, mr


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("cfg", [dict(max_active=4, windows_per_step=24),
                                 dict(max_active=64, windows_per_step=4096),
                                 dict(max_active=3, windows_per_step=7, hbm_budget_bytes=40000)])
def test_served_equals_alone(engine, cfg, mode):
    iv, mr = workload()
    cfg = dict(cfg, mode=mode)
    print(f"[gpu] serve {cfg}: alone", flush=True)
    alone_iv = [engine.run_batch([a], [b], large_cap=64)[0][0] for a, b in iv]
    alone_mr = [engine.mr(x) for x in mr]
    print("[gpu] serve: mixed", flush=True)
    with engine.serve(large_capacity=64, **cfg) as srv:
        ids = {}
        lock = threading.Lock()

        def client(items):
            for kind, k in items:
                rid = srv.submit_interval(*iv[k]) if kind == "iv" else srv.submit_mr(mr[k])
                with lock:
                    ids[(kind, k)] = rid

        items = [("iv", k) for k in range(len(iv))] + [("mr", k) for k in range(len(mr))]
        threads = [threading.Thread(target=client, args=(items[t::3],)) for t in range(3)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        got = {key: srv.wait(rid) for key, rid in ids.items()}
    for k in range(len(iv)):
        res, lat = got[("iv", k)]
        assert lat >= 0
        same(res, alone_iv[k])
    for k in range(len(mr)):
        v, _ = got[("mr", k)]
        assert np.array_equal(v, alone_mr[k])


def test_served_step_counters(engine):
    """The host-stepped server's counters: every window of every interval
    request is enqueued exactly once (no restarts here), in at least as many
    steps as the step size requires, and the steps' device time is within
    the wall time."""
    iv, _ = workload()
    alone_windows = sum(engine.run_batch([a], [b])[1]["windows"] for a, b in iv)
    with engine.serve(max_active=8, windows_per_step=64, large_capacity=64,
                      mode=vossgpu.SERVE_HOST_STEPS) as srv:
        ids = [srv.submit_interval(a, b) for a, b in iv]
        for rid in ids:
            srv.wait(rid)
        st = srv.stats()
    assert st["windows"] == alone_windows
    assert st["steps"] >= alone_windows / 64
    assert 0 < st["device_busy_s"] <= st["wall_s"]


def test_loop_counters(engine):
    """The device loop's counters: every window published once (no
    restarts here), at least one launch of the loop kernel, its resident
    time within the wall time."""
    iv, _ = workload()
    alone_windows = sum(engine.run_batch([a], [b])[1]["windows"] for a, b in iv)
    with engine.serve(max_active=8, large_capacity=64) as srv:
        ids = [srv.submit_interval(a, b) for a, b in iv]
        for rid in ids:
            srv.wait(rid)
        st = srv.stats()
    assert st["windows"] == alone_windows
    assert st["steps"] >= 1
    assert 0 < st["device_busy_s"] <= st["wall_s"]


@pytest.mark.parametrize("mode", MODES)
def test_served_bucket_overflow_restarts_exactly(engine, mode):
    """Bucket rooms at 50 % of their estimate: the first steps with bucketed
    windows overflow, their requests start over with exact rooms, and every
    result still equals the request run alone."""
    iv = [(10**15 + k * 10**9, 10**15 + k * 10**9 + 5 * 10**7) for k in range(4)]
    iv += [(10**16 - 10**8, 10**16 - 10**8 + 3 * 10**7), (10**12, 10**12 + 10**6)]
    alone = [engine.run_batch([a], [b], large_cap=64)[0][0] for a, b in iv]
    t0 = engine.tuning()
    engine.set_tuning(bucket_room_pct=50, bucket_limit=1 << 17, bucket_min_windows=2)
    try:
        print("[gpu] serve with undersized bucket rooms", flush=True)
        with engine.serve(max_active=4, windows_per_step=16, large_capacity=64, mode=mode) as srv:
            got = [srv.wait(srv.submit_interval(a, b))[0] for a, b in iv]
            st = srv.stats()
    finally:
        engine.set_tuning(**t0)
    for g, a in zip(got, alone):
        same(g, a)
    planned = sum(engine.run_batch([a], [b])[1]["windows"] for a, b in iv)
    assert st["windows"] > planned  # some windows ran twice (the restarts)
