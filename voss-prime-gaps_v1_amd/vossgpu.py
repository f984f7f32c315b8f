"""ctypes binding of libvoss_amd.so (include/voss/voss.h).

The product path: every call goes to the gfx950 HIP library.  There is no
CPU fallback -- if the library is missing or no GPU is present, the calls
raise.  Tests, bench.py and the drop-in driver use this module.
"""
import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "libvoss_amd.so")
HEADER = os.path.join(os.path.dirname(PKG), "include", "voss", "voss.h")

MAX_GAP = 100000
MODQ_MAX_GAP = 100
MODQ_ROWS = 10
Q_SLOTS = 40
GAP_SEQ_SIZE = 5000000

VOSS_OK = 0
ERRORS = {-1: "EINVAL", -2: "EDEVICE", -3: "ENOMEM", -4: "ERANGE", -5: "EBUSY",
          -6: "ENOTFOUND"}


class VossError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class LargeGap(ctypes.Structure):
    _fields_ = [("position", ctypes.c_uint64), ("gap", ctypes.c_int32),
                ("merit", ctypes.c_float)]


LARGE_DTYPE = np.dtype([("position", "<u8"), ("gap", "<i4"), ("merit", "<f4")])


def large_array(ptr, n):
    """View n voss_large_gap records at ptr as a numpy structured array (copy)."""
    out = np.empty(n, dtype=LARGE_DTYPE)
    if n:  # one memmove (a structured-dtype copy goes record by record)
        ctypes.memmove(out.ctypes.data, ptr, 16 * n)
    return out


class Result(ctypes.Structure):
    _fields_ = [
        ("total_primes", ctypes.c_uint64),
        ("first_prime", ctypes.c_uint64),
        ("last_prime", ctypes.c_uint64),
        ("c1", ctypes.c_uint64),
        ("c3", ctypes.c_uint64),
        ("modq", ctypes.c_uint64 * Q_SLOTS),
        ("gaps", ctypes.c_int64 * MAX_GAP),
        ("modq_gap", ctypes.c_int64 * (MODQ_ROWS * MODQ_MAX_GAP)),
        ("large_count", ctypes.c_uint64),
        ("large_capacity", ctypes.c_uint64),
        ("large", ctypes.POINTER(LargeGap)),
        ("seconds", ctypes.c_double),
        ("device_ms", ctypes.c_double),
        ("sieve_ms", ctypes.c_double),
        ("windows", ctypes.c_uint64),
        ("bucket_ms", ctypes.c_double),
        ("bucket_entries", ctypes.c_uint64),
    ]


GAP_BINS = 1024
MODQ_BINS = 50


class IntervalResult(ctypes.Structure):
    """voss_interval_result: compact per-interval result of the batch APIs."""
    _fields_ = [
        ("lo", ctypes.c_uint64), ("hi", ctypes.c_uint64),
        ("total_primes", ctypes.c_uint64),
        ("first_prime", ctypes.c_uint64),
        ("last_prime", ctypes.c_uint64),
        ("c1", ctypes.c_uint64),
        ("c3", ctypes.c_uint64),
        ("modq", ctypes.c_uint64 * Q_SLOTS),
        ("gaps", ctypes.c_uint64 * GAP_BINS),
        ("modq_gap", ctypes.c_uint64 * (MODQ_ROWS * MODQ_BINS)),
        ("large_count", ctypes.c_uint64),
        ("large_capacity", ctypes.c_uint64),
        ("large", ctypes.POINTER(LargeGap)),
    ]


def interval_to_dict(r):
    """Expands a compact result to the layout of result_to_dict."""
    gaps = np.zeros(MAX_GAP, dtype=np.int64)
    bins = np.ctypeslib.as_array(r.gaps).astype(np.int64)
    gaps[1] = bins[0]
    gaps[2:2 * GAP_BINS:2] = bins[1:]
    mq = np.zeros((MODQ_ROWS, MODQ_MAX_GAP), dtype=np.int64)
    mb = np.ctypeslib.as_array(r.modq_gap).astype(np.int64).reshape(MODQ_ROWS, MODQ_BINS)
    mq[:, 0:2 * MODQ_BINS:2] = mb
    n = min(int(r.large_count), int(r.large_capacity))
    return {
        "lo": int(r.lo), "hi": int(r.hi),
        "total_primes": int(r.total_primes),
        "first_prime": int(r.first_prime),
        "last_prime": int(r.last_prime),
        "c1": int(r.c1),
        "c3": int(r.c3),
        "modq": np.ctypeslib.as_array(r.modq).astype(np.uint64).copy(),
        "gaps": gaps,
        "modq_gap": mq,
        "large_count": int(r.large_count),
        "large": large_array(r.large, n),
    }


class ServeConfig(ctypes.Structure):
    _fields_ = [("max_active", ctypes.c_uint32), ("windows_per_step", ctypes.c_uint32),
                ("hbm_budget_bytes", ctypes.c_uint64), ("large_capacity", ctypes.c_uint32),
                ("mr_per_step", ctypes.c_uint32), ("mode", ctypes.c_uint32)]


SERVE_DEVICE_LOOP = 0  # persistent loop kernel fed from a pinned ring (default)
SERVE_HOST_STEPS = 1   # host-planned sieve steps enqueued back to back


class Tuning(ctypes.Structure):
    _fields_ = [("window_bytes", ctypes.c_uint32), ("wg_per_cu", ctypes.c_uint32),
                ("tier_a_limit", ctypes.c_uint32), ("bucket_limit", ctypes.c_uint32),
                ("bucket_min_windows", ctypes.c_uint32), ("bucket_room_pct", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32 * 2)]


EXPORTS = [
    "voss_create", "voss_destroy", "voss_last_error", "voss_device_name",
    "voss_set_tuning", "voss_get_tuning", "voss_autotune", "voss_run_plain", "voss_run_plain_range",
    "voss_sample_plain",
    "voss_run_interval", "voss_run_batch", "voss_run_batch_plain", "voss_mr_batch",
    "voss_mr_batch_device", "voss_mr_batch_device_noalloc", "voss_verify_gaps", "voss_device_alloc", "voss_device_free",
    "voss_memcpy_h2d", "voss_memcpy_d2h", "voss_serve_start", "voss_serve_submit_interval",
    "voss_serve_submit_mr", "voss_serve_poll", "voss_serve_wait", "voss_serve_stats",
    "voss_serve_unit_stats", "voss_serve_stop",
    "voss_set_graph_mode", "voss_batch_prepare", "voss_batch_run", "voss_batch_free",
    "voss_batch_add_mr",
]

_lib = None


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libvoss_amd.so not built: {LIB_PATH} "
                               "(python voss-prime-gaps_v1_amd/build.py)")
        L = ctypes.CDLL(LIB_PATH)
        vp, u64, sz, i = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_size_t, ctypes.c_int
        P = ctypes.POINTER
        L.voss_create.argtypes = [i, P(vp)]
        L.voss_destroy.argtypes = [vp]
        L.voss_destroy.restype = None
        L.voss_last_error.restype = ctypes.c_char_p
        L.voss_device_name.argtypes = [vp, ctypes.c_char_p, sz, P(i), P(ctypes.c_double)]
        L.voss_set_tuning.argtypes = [vp, P(Tuning)]
        L.voss_get_tuning.argtypes = [vp, P(Tuning)]
        L.voss_autotune.argtypes = [vp, vp, vp, sz, sz, i, P(Tuning), P(ctypes.c_double)]
        L.voss_run_plain.argtypes = [vp, u64, u64, P(Result)]
        L.voss_sample_plain.argtypes = [vp, u64, u64, vp, vp, P(u64)]
        L.voss_run_plain_range.argtypes = [vp, u64, u64, u64, u64, P(Result)]
        L.voss_run_interval.argtypes = [vp, u64, u64, P(Result)]
        L.voss_run_batch.argtypes = [vp, vp, vp, sz, vp, P(Result)]
        L.voss_run_batch_plain.argtypes = [vp, vp, vp, sz, vp, P(Result)]
        L.voss_set_graph_mode.argtypes = [vp, i]
        L.voss_batch_prepare.argtypes = [vp, vp, vp, sz, P(vp)]
        L.voss_batch_run.argtypes = [vp, vp, P(Result)]
        L.voss_batch_free.argtypes = [vp]
        L.voss_batch_add_mr.argtypes = [vp, vp, sz, vp]
        L.voss_batch_free.restype = None
        L.voss_mr_batch.argtypes = [vp, vp, sz, vp]
        L.voss_mr_batch_device.argtypes = [vp, vp, sz, vp, P(ctypes.c_double)]
        L.voss_mr_batch_device_noalloc.argtypes = [vp, vp, sz, vp, P(ctypes.c_double)]
        L.voss_verify_gaps.argtypes = [vp, vp, vp, sz, vp]
        L.voss_device_alloc.argtypes = [vp, sz, P(vp)]
        L.voss_device_free.argtypes = [vp, vp]
        L.voss_memcpy_h2d.argtypes = [vp, vp, vp, sz]
        L.voss_memcpy_d2h.argtypes = [vp, vp, vp, sz]
        L.voss_serve_start.argtypes = [vp, P(ServeConfig), P(vp)]
        L.voss_serve_submit_interval.argtypes = [vp, u64, u64, P(u64)]
        L.voss_serve_submit_mr.argtypes = [vp, vp, sz, P(u64)]
        L.voss_serve_poll.argtypes = [vp, u64, P(i), vp, vp, P(ctypes.c_double)]
        L.voss_serve_wait.argtypes = [vp, u64, vp, vp, P(ctypes.c_double)]
        L.voss_serve_stop.argtypes = [vp]
        L.voss_serve_unit_stats.argtypes = [vp, P(u64), P(ctypes.c_double)]
        L.voss_serve_stats.argtypes = [vp, P(u64), P(u64), P(ctypes.c_double),
                                       P(ctypes.c_double)]
        for name in EXPORTS:
            if name not in ("voss_destroy", "voss_last_error", "voss_batch_free"):
                getattr(L, name).restype = ctypes.c_int
        _lib = L
    return _lib


def _check(rc):
    if rc != VOSS_OK:
        raise VossError(rc, lib().voss_last_error().decode(errors="replace"))


def result_to_dict(r, n_large=None):
    n = min(int(r.large_count), int(r.large_capacity)) if n_large is None else n_large
    return {
        "total_primes": int(r.total_primes),
        "first_prime": int(r.first_prime),
        "last_prime": int(r.last_prime),
        "c1": int(r.c1),
        "c3": int(r.c3),
        "modq": np.array(np.ctypeslib.as_array(r.modq), dtype=np.uint64),
        "gaps": np.array(np.ctypeslib.as_array(r.gaps), dtype=np.int64),
        "modq_gap": np.array(np.ctypeslib.as_array(r.modq_gap), dtype=np.int64)
                      .reshape(MODQ_ROWS, MODQ_MAX_GAP),
        "large_count": int(r.large_count),
        "large": large_array(r.large, n),
        "seconds": float(r.seconds),
        "device_ms": float(r.device_ms),
        "sieve_ms": float(r.sieve_ms),
        "windows": int(r.windows),
        "bucket_ms": float(r.bucket_ms),
        "bucket_entries": int(r.bucket_entries),
    }


def _new_result(large_cap):
    r = Result()
    buf = None
    if large_cap:
        buf = (LargeGap * large_cap)()
        r.large = ctypes.cast(buf, ctypes.POINTER(LargeGap))
        r.large_capacity = large_cap
    return r, buf


class Engine:
    """One GPU context (voss_ctx)."""

    def __init__(self, device=0):
        self._h = ctypes.c_void_p()
        _check(lib().voss_create(device, ctypes.byref(self._h)))
        name = ctypes.create_string_buffer(256)
        cus = ctypes.c_int()
        gb = ctypes.c_double()
        _check(lib().voss_device_name(self._h, name, 256, ctypes.byref(cus), ctypes.byref(gb)))
        self.name, self.num_cus, self.total_gb = name.value.decode(), cus.value, gb.value

    def close(self):
        if self._h:
            lib().voss_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def tuning(self):
        t = Tuning()
        _check(lib().voss_get_tuning(self._h, ctypes.byref(t)))
        return {"window_bytes": t.window_bytes, "wg_per_cu": t.wg_per_cu,
                "tier_a_limit": t.tier_a_limit, "bucket_limit": t.bucket_limit,
                "bucket_min_windows": t.bucket_min_windows, "bucket_room_pct": t.bucket_room_pct}

    def set_tuning(self, window_bytes=0, wg_per_cu=0, tier_a_limit=0, bucket_limit=0,
                   bucket_min_windows=0, bucket_room_pct=0):
        """Zero fields keep the current value."""
        t = Tuning(window_bytes, wg_per_cu, tier_a_limit, bucket_limit, bucket_min_windows,
                   bucket_room_pct)
        _check(lib().voss_set_tuning(self._h, ctypes.byref(t)))

    def _result(self, large_cap):
        """A Result with a large-gap buffer of at least large_cap entries,
        reused across calls (allocating and zeroing it per call costs more
        than the host side of a small run)."""
        if getattr(self, "_res_cap", -1) < large_cap:
            self._res, self._res_buf = _new_result(large_cap)
            self._res_cap = large_cap
        r = self._res
        r.large_capacity = large_cap
        return r

    def autotune(self, lo, hi, max_intervals=512, repeats=1):
        """Coordinate-descent autotuner on a sample of the intervals; the best
        tuning is installed and returned with its time."""
        lo = np.ascontiguousarray(lo, dtype=np.uint64)
        hi = np.ascontiguousarray(hi, dtype=np.uint64)
        t = Tuning()
        sec = ctypes.c_double()
        _check(lib().voss_autotune(self._h, lo.ctypes.data, hi.ctypes.data, lo.shape[0],
                                   max_intervals, repeats, ctypes.byref(t), ctypes.byref(sec)))
        return self.tuning(), sec.value

    def run_plain(self, N, seg_num, large_cap=1 << 20):
        r = self._result(large_cap)
        _check(lib().voss_run_plain(self._h, N, seg_num, ctypes.byref(r)))
        return result_to_dict(r)

    def run_plain_range(self, N, seg_num, seg_begin, seg_end, large_cap=1 << 20):
        r = self._result(large_cap)
        _check(lib().voss_run_plain_range(self._h, N, seg_num, seg_begin, seg_end,
                                          ctypes.byref(r)))
        return result_to_dict(r)

    def sample_plain(self, N, seg_num):
        """gap_sequence sample: (gap_seq int32[5e6], pos_seq uint64[5e6], rows)."""
        gs = np.zeros(GAP_SEQ_SIZE, dtype=np.int32)
        ps = np.zeros(GAP_SEQ_SIZE, dtype=np.uint64)
        rows = ctypes.c_uint64()
        _check(lib().voss_sample_plain(self._h, N, seg_num, gs.ctypes.data, ps.ctypes.data,
                                       ctypes.byref(rows)))
        return gs, ps, int(rows.value)

    def run_interval(self, lo, hi, large_cap=1 << 16):
        r = self._result(large_cap)
        _check(lib().voss_run_interval(self._h, lo, hi, ctypes.byref(r)))
        return result_to_dict(r)

    @staticmethod
    def _interval_results(n, large_cap):
        """n IntervalResult records with large_cap large-gap slots each.  The
        memory comes uninitialised from numpy (the library clears every
        record it fills) and the per-record pointers are written as one
        strided array, not field by field from Python."""
        size = ctypes.sizeof(IntervalResult)
        rbuf = np.empty(max(n, 1) * size, dtype=np.uint8)
        res = (IntervalResult * n).from_buffer(rbuf)
        lbuf = np.empty(max(large_cap * n, 1) * ctypes.sizeof(LargeGap), dtype=np.uint8)
        if n:
            u64 = rbuf[:n * size].view(np.uint64).reshape(n, size // 8)
            u64[:, IntervalResult.large_capacity.offset // 8] = large_cap
            u64[:, IntervalResult.large.offset // 8] = (
                (lbuf.ctypes.data + 16 * large_cap * np.arange(n, dtype=np.uint64))
                if large_cap else 0)
        res._keep = rbuf  # the records live in rbuf
        return res, lbuf

    @staticmethod
    def _stats(stats):
        return {"seconds": stats.seconds, "device_ms": stats.device_ms,
                "sieve_ms": stats.sieve_ms, "windows": int(stats.windows),
                "bucket_ms": stats.bucket_ms, "bucket_entries": int(stats.bucket_entries)}

    @staticmethod
    def batch_results(n, large_cap=64):
        """Caller-owned output records for n intervals (pass as out= to
        run_batch / run_batch_plain / PreparedBatch.run to reuse them, as a
        C caller reuses its voss_interval_result array)."""
        return Engine._interval_results(n, large_cap)

    @staticmethod
    def _outputs(n, large_cap, out):
        if out is None:
            return Engine._interval_results(n, large_cap)
        res, lbuf = out
        if len(res) != n or (n and res[0].large_capacity != large_cap):
            raise ValueError("out= holds records of another batch shape")
        return res, lbuf

    def _batch(self, fn, lo, hi, large_cap, raw, out=None):
        lo = np.ascontiguousarray(lo, dtype=np.uint64)
        hi = np.ascontiguousarray(hi, dtype=np.uint64)
        n = lo.shape[0]
        res, lbuf = self._outputs(n, large_cap, out)
        stats = Result()
        _check(fn(self._h, lo.ctypes.data, hi.ctypes.data, n, ctypes.cast(res, ctypes.c_void_p),
                  ctypes.byref(stats)))
        st = self._stats(stats)
        if raw:
            return res, lbuf, st
        return [interval_to_dict(res[k]) for k in range(n)], st

    def set_graph_mode(self, enable):
        """True: voss_run_batch runs as one captured hipGraph (the default)."""
        _check(lib().voss_set_graph_mode(self._h, 1 if enable else 0))

    def prepare_batch(self, lo, hi):
        """A prepared batch: planned once, captured on its first run and
        replayed by every later run."""
        return PreparedBatch(self, lo, hi)

    def run_batch(self, lo, hi, large_cap=64, raw=False, out=None):
        """All intervals in one persistent launch.  Returns (results, stats);
        raw=True returns the ctypes IntervalResult array instead of dicts.
        out: records from batch_results to fill instead of new ones."""
        return self._batch(lib().voss_run_batch, lo, hi, large_cap, raw, out)

    def run_batch_plain(self, lo, hi, large_cap=64, raw=False, out=None):
        """One launch sequence per interval (the comparison form)."""
        return self._batch(lib().voss_run_batch_plain, lo, hi, large_cap, raw, out)

    def mr(self, ns):
        ns = np.ascontiguousarray(ns, dtype=np.uint64)
        out = np.zeros(ns.shape[0], dtype=np.uint8)
        _check(lib().voss_mr_batch(self._h, ns.ctypes.data, ns.shape[0], out.ctypes.data))
        return out

    def mr_device(self, ns, repeats=1):
        """MR on device-resident candidates; returns (verdicts, [kernel ms])."""
        ns = np.ascontiguousarray(ns, dtype=np.uint64)
        n = ns.shape[0]
        dn, dv = ctypes.c_void_p(), ctypes.c_void_p()
        _check(lib().voss_device_alloc(self._h, max(n * 8, 8), ctypes.byref(dn)))
        _check(lib().voss_device_alloc(self._h, max(n, 8), ctypes.byref(dv)))
        try:
            _check(lib().voss_memcpy_h2d(self._h, dn, ns.ctypes.data, n * 8))
            times = []
            for _ in range(repeats):
                ms = ctypes.c_double()
                _check(lib().voss_mr_batch_device(self._h, dn, n, dv, ctypes.byref(ms)))
                times.append(ms.value)
            out = np.zeros(n, dtype=np.uint8)
            _check(lib().voss_memcpy_d2h(self._h, out.ctypes.data, dv, n))
        finally:
            lib().voss_device_free(self._h, dn)
            lib().voss_device_free(self._h, dv)
        return out, times

    def serve(self, max_active=0, windows_per_step=0, hbm_budget_bytes=0, large_capacity=0,
              mr_per_step=0, mode=SERVE_DEVICE_LOOP):
        """A continuous-batching Server on this engine (which it owns until
        the server stops); mode SERVE_DEVICE_LOOP or SERVE_HOST_STEPS."""
        return Server(self, ServeConfig(max_active, windows_per_step, hbm_budget_bytes,
                                        large_capacity, mr_per_step, mode))

    def verify_gaps(self, p_after, gaps):
        pa = np.ascontiguousarray(p_after, dtype=np.uint64)
        g = np.ascontiguousarray(gaps, dtype=np.int32)
        ok = np.zeros(pa.shape[0], dtype=np.uint8)
        _check(lib().voss_verify_gaps(self._h, pa.ctypes.data, g.ctypes.data, pa.shape[0],
                                      ok.ctypes.data))
        return ok


class PreparedBatch:
    """voss_batch_prepare / voss_batch_run / voss_batch_free."""

    def __init__(self, engine, lo, hi):
        self.engine = engine
        self.lo = np.ascontiguousarray(lo, dtype=np.uint64)
        self.hi = np.ascontiguousarray(hi, dtype=np.uint64)
        self._b = ctypes.c_void_p()
        _check(lib().voss_batch_prepare(engine._h, self.lo.ctypes.data, self.hi.ctypes.data,
                                        self.lo.shape[0], ctypes.byref(self._b)))

    def run(self, large_cap=64, raw=False, out=None):
        n = self.lo.shape[0]
        res, lbuf = Engine._outputs(n, large_cap, out)
        stats = Result()
        _check(lib().voss_batch_run(self._b, ctypes.cast(res, ctypes.c_void_p), ctypes.byref(stats)))
        st = Engine._stats(stats)
        if raw:
            return res, lbuf, st
        return [interval_to_dict(res[k]) for k in range(n)], st

    def add_mr(self, d_n, count, d_verdict):
        """Miller-Rabin over device buffers (voss_device_alloc) in every run."""
        _check(lib().voss_batch_add_mr(self._b, d_n, count, d_verdict))

    def close(self):
        if self._b:
            lib().voss_batch_free(self._b)
            self._b = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Server:
    """voss_serve_*: submit interval / MR requests from any thread; results
    are identical to running each request alone."""

    def __init__(self, engine, cfg):
        self._h = ctypes.c_void_p()
        self.large_capacity = cfg.large_capacity or 64
        _check(lib().voss_serve_start(engine.handle, ctypes.byref(cfg), ctypes.byref(self._h)))
        self._mr_sizes = {}

    def submit_interval(self, lo, hi):
        rid = ctypes.c_uint64()
        _check(lib().voss_serve_submit_interval(self._h, lo, hi, ctypes.byref(rid)))
        return rid.value

    def submit_mr(self, ns):
        ns = np.ascontiguousarray(ns, dtype=np.uint64)
        rid = ctypes.c_uint64()
        _check(lib().voss_serve_submit_mr(self._h, ns.ctypes.data, ns.shape[0], ctypes.byref(rid)))
        self._mr_sizes[rid.value] = ns.shape[0]
        return rid.value

    def _out(self, rid):
        if rid in self._mr_sizes:
            v = np.zeros(max(self._mr_sizes[rid], 1), dtype=np.uint8)
            return None, v, None
        r = IntervalResult()
        lbuf = (LargeGap * self.large_capacity)()
        r.large = ctypes.cast(lbuf, ctypes.POINTER(LargeGap))
        r.large_capacity = self.large_capacity
        return r, None, lbuf

    def _conv(self, rid, r, v):
        if rid in self._mr_sizes:
            return v[: self._mr_sizes.pop(rid)]
        return interval_to_dict(r)

    def wait(self, rid):
        """(result, latency_s): dict for intervals, verdict array for MR."""
        r, v, lbuf = self._out(rid)
        lat = ctypes.c_double()
        _check(lib().voss_serve_wait(self._h, rid, ctypes.byref(r) if r is not None else None,
                                     v.ctypes.data if v is not None else None, ctypes.byref(lat)))
        return self._conv(rid, r, v), lat.value

    def poll(self, rid):
        """None while running, else (result, latency_s)."""
        r, v, lbuf = self._out(rid)
        done = ctypes.c_int()
        lat = ctypes.c_double()
        _check(lib().voss_serve_poll(self._h, rid, ctypes.byref(done),
                                     ctypes.byref(r) if r is not None else None,
                                     v.ctypes.data if v is not None else None, ctypes.byref(lat)))
        if not done.value:
            return None
        return self._conv(rid, r, v), lat.value

    def stats(self):
        """{steps, windows, device_busy_s, wall_s} of the running server."""
        st, win = ctypes.c_uint64(), ctypes.c_uint64()
        busy, wall = ctypes.c_double(), ctypes.c_double()
        _check(lib().voss_serve_stats(self._h, ctypes.byref(st), ctypes.byref(win),
                                      ctypes.byref(busy), ctypes.byref(wall)))
        return {"steps": st.value, "windows": win.value, "device_busy_s": busy.value,
                "wall_s": wall.value}

    def unit_stats(self):
        """Device loop: {kind: (units, workgroup_seconds)} for windows,
        bucket tasks and Miller-Rabin chunks."""
        n = (ctypes.c_uint64 * 3)()
        b = (ctypes.c_double * 3)()
        _check(lib().voss_serve_unit_stats(self._h, n, b))
        return {k: (int(n[i]), float(b[i])) for i, k in enumerate(("window", "bucket", "mr"))}

    def stop(self):
        if self._h:
            rc = lib().voss_serve_stop(self._h)
            self._h = ctypes.c_void_p()
            _check(rc)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.stop()
