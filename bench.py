"""Benchmark: sieved numbers/s of the plain [1, N] pipeline at N = 1e11
(SEG_NUM = 3e10) on one MI355X, with Miller-Rabin tests/s beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

One step = one full voss_run_plain(N=1e11, SEG_NUM=3e10): planning, the
sieve+extract+histogram kernel, the stitch kernel, the copy of the
accumulators and the host finalisation (the reference's TOTAL TIME window,
src/voss_v7_live.cu:432 -> 632).  The input is the number N itself, so
nothing crosses PCIe before the timed region.

Prints ONE JSON line (rank 0).  `roofline` describes the dominant kernel
(sieve_kernel): algorithmic LDS bytes per launch (8 bytes per dword
read-modify-write: one per (prime, word) for the presieved primes below 32
and one per mark for the rest, plus the bitmap written once and read once)
over its average duration from HIP events on the library's stream.
`cpu_baseline` times the CPU oracle (a port of the reference semantics) on
the whole workload, [1, 1e11], on every host CPU the process may use: the
median of 5 runs after a warm-up (about 20 s in all), with config 1's
[2, 1e9] run and its large-gap verification beside it and the host's CPU
model, clock and counts recorded.
"""
import argparse
import json
import math
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "voss-prime-gaps_v1_amd")
sys.path.insert(0, ROOT)
sys.path.insert(0, PKG)

import numpy as np  # noqa: E402

N_BENCH = 10**11
REF_NUMBERS_PER_S = 10**11 / 5.8  # README.md:124, A100-SXM4-40GB
SEG_NUM = 30000000000
CPU_SAMPLE_N = N_BENCH

# MI355X LDS: 160 KB per CU, 128 bytes/clock per CU, 256 CUs at 2.4 GHz.
LDS_PEAK_GBS = 128 * 256 * 2.4e9 / 1e9
# HBM bytes per sieve_kernel launch at N=1e11 from the two rocprofv3 --pmc
# runs (FETCH_SIZE x2.0 + WRITE_SIZE x1.0, scaled by a 1 GiB copy kernel);
# the commands and raw counters are in
# profiles/sieve_kernel_plain_1e11_hbm_traffic.md.
SIEVE_1E11_HBM_BYTES = (2 * 3731.6 + 18697.3) * 1024


def model_lds_bytes(N):
    """Algorithmic LDS bytes of the sieve for [1, N] (see module docstring).

    Marks: for every base prime 7 <= p <= sqrt(N), the multiples p*m, m >= p,
    m coprime to 30, that are <= N: sum_p (N/p - p) * 8/30 (plus rounding).
    Primes below 32 are charged one dword RMW per (prime, word) instead."""
    limit = int(math.isqrt(N)) + 1
    s = np.ones(limit + 1, dtype=bool)
    s[:2] = False
    for i in range(2, int(limit**0.5) + 1):
        if s[i]:
            s[i * i::i] = False
    ps = np.nonzero(s)[0]
    ps = ps[ps >= 7].astype(np.float64)
    words = N / 120.0
    small = ps[ps < 32]
    big = ps[ps >= 32]
    marks_big = float(np.sum(np.maximum(N / big - big, 0) * 8.0 / 30.0))
    m_eff = len(small) * words + marks_big
    return 8.0 * m_eff + 2.0 * (N / 30.0), m_eff


def measured_lds_peaks(copy_bytes=0):
    """LDS microbenchmarks from libvoss_probe.so (and the optional copy)."""
    import ctypes
    path = os.path.join(PKG, "lib", "libvoss_probe.so")
    if not os.path.exists(path):
        return None
    L = ctypes.CDLL(path)
    L.voss_probe_lds.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    L.voss_probe_copy.argtypes = [ctypes.c_size_t, ctypes.POINTER(ctypes.c_double)]
    out = {}
    if copy_bytes:
        ms = ctypes.c_double()
        L.voss_probe_copy(copy_bytes, ctypes.byref(ms))
        out["copy_bytes"] = copy_bytes
        out["copy_gbs"] = 2 * copy_bytes / (ms.value * 1e-3) / 1e9
    for k, name in ((0, "ds_read_b128_gbs"), (1, "ds_and_random_gops"), (2, "ds_and_linear_gops")):
        r = ctypes.c_double()
        if L.voss_probe_lds(k, 256, ctypes.byref(r)) == 0:
            out[name] = r.value / 1e9
    L.voss_probe_montmul.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    r = ctypes.c_double()
    if L.voss_probe_montmul(256, ctypes.byref(r)) == 0:
        out["montmul_per_s"] = r.value
    return out


MR_BASES = (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37)
def _tail_squarings(x, n, r):
    """Squarings the witness loop runs after x = b^d mod n, and its verdict."""
    if x == 1 or x == n - 1:
        return 0, True
    for j in range(r - 1):
        x = x * x % n
        if x == n - 1:
            return j + 1, True
        if x == 1:
            return j + 1, False
    return r - 1, False


def mr_mulmods_device(n):
    """Modular multiplications the device kernels run for n (csrc/mr.hip):
    base 2 by binary powering with doublings for the set bits (bitlen(d) - 1
    squarings), then bases 3..37 by fixed 3-bit windows: six products for
    the table of powers and four per base-8 digit below the top one (the
    lane's own digit count; a wave runs its largest); the witness-loop
    squarings that run are counted for every base.  Trial division and the
    modular additions that step from one base to the next cost 0."""
    if n < 2:
        return 0
    for b in MR_BASES:
        if n == b or n % b == 0:
            return 0
    d, r = n - 1, 0
    while d % 2 == 0:
        d //= 2
        r += 1
    work = d.bit_length() - 1
    t, ok = _tail_squarings(pow(2, d, n), n, r)
    work += t
    if not ok:
        return work
    digits = (d.bit_length() - 1) // 3
    for b in MR_BASES[1:]:
        t, ok = _tail_squarings(pow(b, d, n), n, r)
        work += 6 + 4 * digits + t
        if not ok:
            return work
    return work


def mr_mulmods(n):
    """Algorithmic work of one Miller-Rabin test (is_prime_u64,
    src/verify_cpp.cpp:38-57) in 64-bit modular multiplications: per base
    tried, the square-and-multiply of mod_pow (:26-34: bitlen(d) squarings
    and popcount(d) products) plus the squarings of the witness loop that
    run; a composite stops at its first witness; trial division costs 0."""
    if n < 2:
        return 0
    for b in MR_BASES:
        if n == b or n % b == 0:
            return 0
    d, r = n - 1, 0
    while d % 2 == 0:
        d //= 2
        r += 1
    per = d.bit_length() + bin(d).count("1")
    work = 0
    for a in MR_BASES:
        work += per
        x = pow(a, d, n)
        if x == 1 or x == n - 1:
            continue
        for _ in range(r - 1):
            x = x * x % n
            work += 1
            if x == n - 1:
                break
        else:
            return work
    return work


def host_cpu():
    """The host the CPU baselines run on: logical CPUs (nproc), the CPUs this
    process may use (its affinity set, the share of the host it was given),
    the model and clock from /proc/cpuinfo, OMP_NUM_THREADS."""
    info = {"nproc": os.cpu_count(), "usable_cpus": len(os.sched_getaffinity(0)),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}
    try:
        with open("/proc/cpuinfo") as f:
            lines = f.read().splitlines()
        model = [l.split(":", 1)[1].strip() for l in lines if l.startswith("model name")]
        mhz = [float(l.split(":", 1)[1]) for l in lines if l.startswith("cpu MHz")]
        info["model"] = model[0] if model else None
        info["mhz"] = round(statistics.median(mhz), 1) if mhz else None
    except OSError:
        pass
    return info


def cpu_threads():
    """Threads for the CPU baselines: every CPU this process may use (all
    host cores it was given), or OMP_NUM_THREADS when that is set lower."""
    n = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp and omp.isdigit() and 0 < int(omp) < n:
        n = int(omp)
    return n


def median_of(fn, runs=5, warmup=1):
    """fn() timed runs times after warmup untimed calls: (median s, all s, last result)."""
    out = None
    for _ in range(warmup):
        out = fn()
    times = []
    for _ in range(runs):
        t = time.perf_counter()
        out = fn()
        times.append(time.perf_counter() - t)
    return statistics.median(times), times, out


def cpu_baseline(threads):
    """The oracle (the reference's plain-mode semantics restated in C with
    OpenMP; the reference has no CPU sieve) on the whole workload [1, 1e11],
    and on config 1, [2, 1e9]: median of 5 runs after one warm-up."""
    from oracle import oracle as O
    med, times, r = median_of(lambda: O.run_plain(CPU_SAMPLE_N, SEG_NUM, threads=threads,
                                                  large_cap=16))
    assert r["total_primes"] == 4118054813  # pi(1e11)
    med9, times9, r9 = median_of(lambda: O.run_plain(10**9, SEG_NUM, threads=threads))
    assert r9["total_primes"] == 50847534  # pi(1e9)
    return {"value": CPU_SAMPLE_N / med, "unit": "numbers/s", "cores": threads,
            "kind": "port", "seconds": round(med, 3), "runs_s": [round(t, 3) for t in times],
            "sample": f"the whole workload, plain [1, {CPU_SAMPLE_N:.0e}] SEG_NUM=3e10, "
                      f"oracle C/OpenMP, {threads} threads, median of 5 after a warm-up",
            "config1_1e9": {"numbers_per_s": 10**9 / med9, "seconds": round(med9, 4),
                            "runs_s": [round(t, 4) for t in times9]},
            "host": host_cpu(), "large_1e9": r9["large"]}


def batch_leg(eng, W, with_plain=True, autotune=True, mr_count=10**7):
    """Config 4: 4096 seeded intervals (starts in [1e12, 1e16), lengths
    log-uniform in [1e4, 1e9]).  The autotuner picks window, workgroup and
    bucket thresholds on a 256-interval sample; then the batch runs as one
    captured hipGraph (one call, planning included), as replays of a
    prepared batch, eagerly (for the per-kernel split), and as one launch
    sequence per interval; all outputs are compared bit for bit."""
    import vossgpu
    iv = W.batch_intervals(4096)
    lo = [a for a, _ in iv]
    hi = [b for _, b in iv]
    numbers = sum(b - a for a, b in iv)
    eng.run_interval(10**16, 10**16 + 10**6)  # base primes to 1e8: outside the timing
    default = eng.tuning()
    out = {"intervals": len(iv), "numbers": numbers}
    if autotune:
        t = time.perf_counter()
        best, sec = eng.autotune(lo, hi, max_intervals=256, repeats=1)
        out["autotune"] = {"tuning": best, "sample_intervals": 256, "tuned_s": sec,
                           "tuner_s": time.perf_counter() - t}
    off = vossgpu.IntervalResult.large.offset  # compare all fields but the pointer

    def lg(r):
        return vossgpu.large_array(r.large, min(r.large_count, r.large_capacity)).tobytes()

    def same(ra, rb):
        return all(bytes(a)[:off] == bytes(b)[:off] and lg(a) == lg(b) for a, b in zip(ra, rb))

    # caller-owned output records, one set per form compared below (a C
    # caller keeps and reuses its voss_interval_result array); each set is
    # filled once before it is timed, so no timed call pays first-touch
    # page faults of fresh host memory
    outs = [vossgpu.Engine.batch_results(len(iv), 64) for _ in range(4)]
    try:
        eng.run_batch(lo, hi, large_cap=64, raw=True, out=outs[0])  # warm-up
        t = time.perf_counter()
        rb, lbuf, st = eng.run_batch(lo, hi, large_cap=64, raw=True, out=outs[0])
        t_batch = time.perf_counter() - t
        out.update({"batched_s": t_batch, "batched_numbers_per_s": numbers / t_batch,
                    "device_ms": st["device_ms"], "windows": st["windows"],
                    "bucket_room_entries": st["bucket_entries"]})
        with eng.prepare_batch(lo, hi) as pb:
            pb.run(large_cap=64, raw=True, out=outs[1])  # capture
            reps = []
            for _ in range(3):
                t = time.perf_counter()
                rr, lbuf3, _ = pb.run(large_cap=64, raw=True, out=outs[1])
                reps.append(time.perf_counter() - t)
            out.update({"replay_s": min(reps), "replay_numbers_per_s": numbers / min(reps),
                        "replay_bit_exact": bool(same(rb, rr))})
        # the same batch with config 4's Miller-Rabin set in the same graph
        # (MR on a second stream beside the sieve rounds)
        cand = W.mr_candidates(mr_count, large_endpoints(rb))
        out["with_mr"] = batch_with_mr(eng, lo, hi, cand, rb, same, min(reps), outs[2])
        eng.set_graph_mode(False)
        try:
            t = time.perf_counter()
            re_, lbuf4, se = eng.run_batch(lo, hi, large_cap=64, raw=True, out=outs[2])
            out.update({"eager_s": time.perf_counter() - t, "sieve_ms": se["sieve_ms"],
                        "bucket_ms": se["bucket_ms"], "eager_bit_exact": bool(same(rb, re_))})
        finally:
            eng.set_graph_mode(True)
        if with_plain:
            t = time.perf_counter()
            rp, lbuf2, sp = eng.run_batch_plain(lo, hi, large_cap=64, raw=True, out=outs[3])
            t_plain = time.perf_counter() - t
            out.update({"plain_s": t_plain, "batched_over_plain": t_plain / t_batch,
                        "bit_exact": bool(same(rb, rp))})
    finally:
        eng.set_tuning(**default)
    return out, large_endpoints(rb)


def large_endpoints(res):
    """Both ends of every large gap kept in a batch result (MR candidates)."""
    ends = []
    for r in res:
        for i in range(min(r.large_count, r.large_capacity)):
            g = r.large[i]
            ends += [g.position, g.position - g.gap]
    return np.array(ends, dtype=np.uint64)


def verify_leg(eng, large, no_cpu):
    """verify_gap (src/verify_cpp.cpp:59-93) on the GPU over the 1e11 run's
    whole large-gap list -- what run_voss.py hands to verify_cpp after the
    run -- in gaps/s (host buffers, both PCIe copies inside), the median of
    5 after a warm-up.  Beside it on the same list: the oracle's verify_gap
    restatement on the host cores.  (The reference's verify_cpp process is
    not timed: its progress thread sleeps 30 s before its first check and is
    joined before the end, src/verify_cpp.cpp:196-219, so a list that
    verifies in less takes 30 s whatever its size.)"""
    pa = np.ascontiguousarray(large["position"], dtype=np.uint64)
    g = np.ascontiguousarray(large["gap"], dtype=np.int32)
    eng.verify_gaps(pa[:1000], g[:1000])
    ts = []
    for _ in range(5):
        t = time.perf_counter()
        ok = eng.verify_gaps(pa, g)
        ts.append(time.perf_counter() - t)
    assert ok.all()
    sec = statistics.median(ts)
    out = {"gaps": int(len(pa)), "seconds": sec, "gaps_per_s": len(pa) / sec,
           "list": "large gaps of plain [1, 1e11] (gap >= 500, or > 100 with merit >= 10)"}
    if no_cpu:
        return out
    from oracle import oracle as O
    threads = cpu_threads()
    dt, _, ref = median_of(lambda: O.verify_gaps(pa, g, threads=threads))
    assert np.array_equal(ref, ok)
    out["cpu_baseline"] = {"value": len(pa) / dt, "unit": "gaps/s", "cores": threads,
                           "kind": "port", "sample": "the whole list, verify_gap restated "
                           "with __uint128_t (oracle, OpenMP), median of 5 after a warm-up"}
    return out


def verify_config1(eng, large, threads):
    """verify_cpp semantics on config 1's large-gap list ([2, 1e9]): the
    oracle's verify_gap on the host cores and voss_verify_gaps on the GPU,
    each the median of 5 after a warm-up."""
    from oracle import oracle as O
    pa = np.array([p for p, _, _ in large], dtype=np.uint64)
    g = np.array([x for _, x, _ in large], dtype=np.int32)
    cpu_s, _, ref = median_of(lambda: O.verify_gaps(pa, g, threads=threads))
    gpu_s, _, got = median_of(lambda: eng.verify_gaps(pa, g))
    assert ref.all() and np.array_equal(ref, got)
    return {"gaps": int(len(pa)), "cpu_gaps_per_s": len(pa) / cpu_s, "cpu_seconds": cpu_s,
            "gpu_gaps_per_s": len(pa) / gpu_s, "gpu_seconds": gpu_s, "cores": threads}


def batch_with_mr(eng, lo, hi, cand, rb, same, replay_s, out):
    """Prepared batch + device-resident MR batch in one graph: replay time,
    against the batch alone plus MR alone; verdicts against voss_mr_batch."""
    import ctypes
    import vossgpu
    L = vossgpu.lib()
    n = len(cand)
    dn, dv = ctypes.c_void_p(), ctypes.c_void_p()
    vossgpu._check(L.voss_device_alloc(eng._h, n * 8, ctypes.byref(dn)))
    vossgpu._check(L.voss_device_alloc(eng._h, n, ctypes.byref(dv)))
    try:
        vossgpu._check(L.voss_memcpy_h2d(eng._h, dn, cand.ctypes.data, n * 8))
        _, mr_ms = eng.mr_device(cand[: min(n, 10**6)])  # warm-up
        verdict, mr_ms = eng.mr_device(cand, repeats=5)
        with eng.prepare_batch(lo, hi) as pb:
            pb.add_mr(dn, n, dv)
            pb.run(large_cap=64, raw=True, out=out)  # capture
            reps = []
            for _ in range(3):
                t = time.perf_counter()
                rr, lbuf, _ = pb.run(large_cap=64, raw=True, out=out)
                reps.append(time.perf_counter() - t)
        v = np.zeros(n, dtype=np.uint8)
        vossgpu._check(L.voss_memcpy_d2h(eng._h, v.ctypes.data, dv, n))
    finally:
        L.voss_device_free(eng._h, dn)
        L.voss_device_free(eng._h, dv)
    separate = replay_s + statistics.median(mr_ms) / 1e3
    return {"mr_count": n, "replay_s": min(reps), "batch_plus_mr_separately_s": separate,
            "bit_exact": bool(same(rb, rr)), "verdicts_equal": bool(np.array_equal(v, verdict))}


def serve_leg(eng, W, n_req=300, n_cal=80, loads=(0.7, 1.3), windows_per_step=0):
    """Serving: a Poisson stream (half interval requests with start in
    [1e10, 1e16) and length log-uniform in [1e5, 1e10], half MR batches of
    1e3..1e6 candidates), offered at each of `loads` times the capacity of
    the plain path (which runs one request at a time in arrival order, as
    fast as it can): below capacity both paths keep up and latency shows
    the difference; above it the plain path's queue grows while continuous
    batching overlaps requests on the device.  Both servers run: the
    device-resident loop (the default) and the host-stepped scheduler.
    Every served result is compared with the same request run alone (all
    fields of an interval, every MR verdict).  Latency = completion -
    arrival; req/s = requests / (last completion - first arrival)."""
    import threading
    import vossgpu
    reqs = W.serve_requests(n_req + n_cal, 1.0)
    rng = np.random.default_rng(W.SEED0 + 4)
    payload = {}
    for k, (_, kind, p) in enumerate(reqs):
        if kind == "mr":
            payload[k] = rng.integers(2**40, 2**63, size=p, dtype=np.uint64) | np.uint64(1)

    def run_alone(k):
        _, kind, p = reqs[k]
        if kind == "interval":
            return eng.run_batch([p[0]], [p[1]], large_cap=64)[0][0]
        return eng.mr(payload[k])

    def same(a, b):
        if isinstance(a, np.ndarray):
            return bool(np.array_equal(a, b))
        return all(np.array_equal(a[f], b[f]) if isinstance(a[f], np.ndarray) else a[f] == b[f]
                   for f in ("total_primes", "first_prime", "last_prime", "c1", "c3",
                             "large_count", "gaps", "modq_gap", "modq", "large"))

    modes = {"served": vossgpu.SERVE_DEVICE_LOOP, "served_host_steps": vossgpu.SERVE_HOST_STEPS}
    eng.run_interval(10**16, 10**16 + 10**6)  # base primes to 1e8 outside the timing
    t = time.perf_counter()
    for k in range(n_req, n_req + n_cal):  # calibration requests (not in the stream)
        run_alone(k)
    cap = n_cal / (time.perf_counter() - t)
    # the calibration requests once through each server too, so that every
    # path starts with warm buffers (the servers' are kept by the context)
    for mode in modes.values():
        with eng.serve(max_active=64, windows_per_step=windows_per_step, large_capacity=64,
                       mode=mode) as srv:
            warm = [srv.submit_interval(*reqs[k][2]) if reqs[k][1] == "interval" else
                    srv.submit_mr(payload[k]) for k in range(n_req, n_req + n_cal)]
            for rid in warm:
                srv.wait(rid)

    def pct(x, q):
        return float(np.percentile(np.array(x) * 1e3, q))

    def summary(lat, span):
        return {"p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99), "makespan_s": span,
                "req_per_s": n_req / span}

    out = {"requests": n_req, "plain_capacity_req_per_s": cap,
           "interval_numbers": sum(p[1] - p[0] for _, kind, p in reqs[:n_req] if kind == "interval"),
           "mr_tests": sum(p for _, kind, p in reqs[:n_req] if kind == "mr"), "loads": {}}
    alone = {}
    for load in loads:
        lam = load * cap
        arrivals = [a / lam for a, _, _ in reqs[:n_req]]  # the stream at rate lam
        # plain path: one request at a time, FIFO
        lat_plain = []
        t0 = time.perf_counter()
        for k in range(n_req):
            now = time.perf_counter() - t0
            if now < arrivals[k]:
                time.sleep(arrivals[k] - now)
            alone[k] = run_alone(k)
            lat_plain.append(time.perf_counter() - t0 - arrivals[k])
        span_plain = time.perf_counter() - t0
        row = {"rate_req_per_s": lam, "plain": summary(lat_plain, span_plain)}
        # continuous batching
        for name, mode in modes.items():
            ids = [None] * n_req
            with eng.serve(max_active=64, windows_per_step=windows_per_step,
                           large_capacity=64, mode=mode) as srv:
                t0 = time.perf_counter()

                def submitter():
                    for k in range(n_req):
                        now = time.perf_counter() - t0
                        if now < arrivals[k]:
                            time.sleep(arrivals[k] - now)
                        _, kind, p = reqs[k]
                        ids[k] = srv.submit_interval(*p) if kind == "interval" else \
                            srv.submit_mr(payload[k])

                th = threading.Thread(target=submitter)
                th.start()
                th.join()
                got = [srv.wait(rid) for rid in ids]
                span_srv = time.perf_counter() - t0
                st = srv.stats()
                units = srv.unit_stats() if mode == vossgpu.SERVE_DEVICE_LOOP else None
            served = summary([g[1] for g in got], span_srv)
            served.update({"steps": st["steps"], "windows": st["windows"],
                           "device_busy_frac": st["device_busy_s"] / span_srv,
                           "equal_to_alone": all(same(alone[k], got[k][0]) for k in range(n_req))})
            if units:
                served["units"] = {k: {"count": n, "workgroup_s": t} for k, (n, t) in units.items()}
            row[name] = served
        out["loads"][f"{load:g}x"] = row
    return out


def interval_leg(eng):
    """Config 3: [1e13, 1e13 + 1e12), base primes to 3.32e6."""
    lo, hi = 10**13, 10**13 + 10**12
    eng.run_interval(lo, hi, large_cap=1 << 22)  # warm-up: the same call
    t = time.perf_counter()
    r = eng.run_interval(lo, hi, large_cap=1 << 22)
    dt = time.perf_counter() - t
    assert r["total_primes"] == 33353349498, r["total_primes"]
    return {"lo": lo, "hi": hi, "seconds": dt, "numbers_per_s": (hi - lo) / dt,
            "pi": r["total_primes"], "sieve_ms": r["sieve_ms"], "bucket_ms": r["bucket_ms"],
            "large_gaps": r["large_count"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mr-count", type=int, default=10**7)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="first run one plain device copy of BYTES (the HBM-counter "
                         "calibration kernel copy_b128 for rocprofv3 --pmc runs)")
    ap.add_argument("--headline-only", action="store_true",
                    help="only the 1e11 steps and MR (for profiling runs)")
    ap.add_argument("--no-plain", action="store_true",
                    help="skip the per-interval comparison form of the batch leg")
    ap.add_argument("--no-autotune", action="store_true",
                    help="run the batch leg with the default tuning")
    ap.add_argument("--legs", default="interval,batch,serve,verify",
                    help="comma-separated legs after the headline: interval, batch, serve, verify, "
                         "plain1e12, plain1e13, plain1e14 (not default: 0.2 s, 2.7 s, 35 s)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(the sieve would shard by range with one tiny histogram all-reduce)",
              file=sys.stderr)
        sys.exit(2)

    import vossgpu

    eng = vossgpu.Engine(0)
    probe = measured_lds_peaks(args.copy_calibration)
    for _ in range(args.warmup):
        eng.run_plain(N_BENCH, SEG_NUM)
    step_s, sieve_ms, dev_ms = [], [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        r = eng.run_plain(N_BENCH, SEG_NUM)
        step_s.append(time.perf_counter() - ts)
        sieve_ms.append(r["sieve_ms"])
        dev_ms.append(r["device_ms"])
    total = time.perf_counter() - t0
    assert r["total_primes"] == 4118054813, r["total_primes"]

    lds_bytes, m_eff = model_lds_bytes(N_BENCH)
    avg_sieve_s = statistics.mean(sieve_ms) / 1e3
    achieved = lds_bytes / avg_sieve_s / 1e9

    import workloads as W
    legs = {}
    want = set() if args.headline_only else set(args.legs.split(","))
    ends = np.array([10**12 + 39], dtype=np.uint64)
    if "interval" in want:
        legs["interval_1e13"] = interval_leg(eng)
    if "batch" in want:
        legs["batch_4096"], ends = batch_leg(eng, W, with_plain=not args.no_plain,
                                             autotune=not args.no_autotune,
                                             mr_count=args.mr_count)
    if "serve" in want:
        legs["serve"] = serve_leg(eng, W)
    if "verify" in want:
        legs["verify_1e11"] = verify_leg(eng, r["large"], args.no_cpu)
    # plain runs at the reference's larger published scales (not in the
    # default set: [1, 1e14] takes about 35 s); pi(N) from README.md:122-127
    for e, pi, ref_s in ((12, 37607912018, 68.5), (13, 346065536839, 954.0),
                         (14, 3204941750802, 21671.0)):
        if f"plain1e{e}" not in want:
            continue
        t = time.perf_counter()
        rN = eng.run_plain(10**e, SEG_NUM, large_cap=16)
        dt = time.perf_counter() - t
        assert rN["total_primes"] == pi, rN["total_primes"]
        legs[f"plain_1e{e}"] = {"seconds": dt, "numbers_per_s": 10**e / dt,
                                "reference_seconds": ref_s, "windows": rN["windows"],
                                "large_gaps": rN["large_count"]}

    # Miller-Rabin tests/s on device-resident candidates (config 4's set)
    cand = W.mr_candidates(args.mr_count, ends)
    verdict, mr_ms = eng.mr_device(cand, repeats=6)
    mr_ms = mr_ms[1:]  # the median of 5 runs after one warm-up
    mr_rate = len(cand) / (statistics.median(mr_ms) / 1e3)
    eng.mr(cand[:1000])
    t = time.perf_counter()
    v_host = eng.mr(cand)  # host buffers: both PCIe copies inside the timing
    t_host = time.perf_counter() - t
    assert np.array_equal(v_host, verdict)
    mr = {"tests_per_s": mr_rate, "count": int(len(cand)), "primes": int(verdict.sum()),
          "kernel_ms": mr_ms, "host_buffers_tests_per_s": len(cand) / t_host}
    # roofline: algorithmic modular multiplications (exact on a sample of the
    # same set, scaled) over the kernel time, against the measured rate of
    # independent 64-bit Montgomery products on all CUs
    n_sample = min(len(cand), 20000)
    work = sum(mr_mulmods(int(x)) for x in cand[:n_sample]) * len(cand) / n_sample
    executed = sum(mr_mulmods_device(int(x)) for x in cand[:n_sample]) * len(cand) / n_sample
    mr_achieved = work / (statistics.median(mr_ms) / 1e3)
    mr_executed = executed / (statistics.median(mr_ms) / 1e3)
    if probe and probe.get("montmul_per_s"):
        # achieved counts the products the kernels run (windowed powering
        # for bases 3..37); the reference algorithm's products (binary
        # powering for every base) over the same time are reported beside it
        mr["roofline"] = {"bound": "valu", "achieved": mr_executed,
                          "peak": probe["montmul_per_s"], "unit": "mulmod/s",
                          "frac": mr_executed / probe["montmul_per_s"],
                          "mulmods_per_test": executed / len(cand),
                          "reference_mulmods_per_test": work / len(cand),
                          "reference_work_per_s": mr_achieved,
                          "frac_reference_work": mr_achieved / probe["montmul_per_s"],
                          "sample": f"exact count on the first {n_sample} candidates"}
    if not args.no_cpu:
        from oracle import oracle as O
        threads = cpu_threads()
        dt, _, ref = median_of(lambda: O.is_prime_batch(cand, threads=threads))
        assert np.array_equal(ref, verdict)
        mr["cpu_baseline"] = {"value": len(cand) / dt, "unit": "tests/s", "cores": threads,
                              "kind": "port", "sample": f"all {len(cand)} candidates, "
                              "is_prime_u64 restated with __uint128_t (oracle, OpenMP), "
                              "median of 5 after a warm-up"}

    out = {
        "metric": "sieved numbers/s at N=1e11 (plain [1,N], SEG_NUM=3e10, bit-exact outputs)",
        "value": N_BENCH * args.steps / total,
        "unit": "numbers/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * total / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        # the reference's published pi(1e11) wall time: 5.8 s on one A100
        # (README.md:124) -> 1.72e10 numbers/s for the same [1, 1e11] pipeline
        "vs_baseline": (N_BENCH * args.steps / total) / REF_NUMBERS_PER_S,
        "dtype": "u32/u64 integer",
        "data": "synthetic: the range [1, 1e11] itself; batch intervals and MR candidates "
                "from seeded SplitMix64 streams (seed 0x564F5353+k)",
        "config": {"workload": "plain sieve+gaps [1, 1e11], SEG_NUM=3e10 (4 logical segments), "
                               "131072-byte LDS windows", "N": N_BENCH, "SEG_NUM": SEG_NUM,
                   "windows": r["windows"], "device": eng.name},
        "roofline": {"bound": "lds", "kernel": "sieve_kernel",
                     "achieved": achieved, "peak": LDS_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / LDS_PEAK_GBS, "traffic": SIEVE_1E11_HBM_BYTES,
                     "algorithmic_bytes": lds_bytes, "m_eff": m_eff,
                     "avg_kernel_ms": statistics.mean(sieve_ms),
                     # the same work as dword read-modify-writes (M_eff per
                     # launch) against the measured rate of random-address
                     # ds_and_b32 on all CUs (lds_probe), the op the marks are
                     "marks_per_s": m_eff / avg_sieve_s,
                     "frac_of_measured_ds_and_rate": (
                         m_eff / avg_sieve_s / (probe["ds_and_random_gops"] * 1e9)
                         if probe.get("ds_and_random_gops") else None)},
        "device_ms_per_step": statistics.mean(dev_ms),
        "step_ms": [round(1e3 * x, 3) for x in step_s],
        "lds_probe": probe,
        "mr": mr,
        "legs": legs,
    }
    if not args.no_cpu:
        threads = cpu_threads()
        cb = cpu_baseline(threads)
        cb["config1_verify"] = verify_config1(eng, cb.pop("large_1e9"), threads)
        out["cpu_baseline"] = cb
    print(json.dumps(out), flush=True)
    eng.close()


if __name__ == "__main__":
    main()
