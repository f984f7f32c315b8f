"""ctypes binding of the C oracle (``oracle/libvoss_oracle.so``).

TEST INFRASTRUCTURE ONLY: used by tests/, ``__graft_entry__.smoke()`` and the
cpu_baseline leg of ``bench.py`` to compare results with; the product path
never imports it.  Each wrapper names the reference function it restates (paths in
the reference tree).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvoss_oracle.so")
REF_VERIFY = os.path.join(HERE, "_ref", "verify_cpp")

MAX_GAP = 100000
MODQ_MAX_GAP = 100
MODQ_ROWS = 10
Q_SLOTS = 40
GAP_SEQ_SIZE = 5000000


class LargeGap(ctypes.Structure):
    """struct LargeGap (src/voss_v7_live.cu:40)."""
    _fields_ = [("position", ctypes.c_uint64), ("gap", ctypes.c_int32),
                ("merit", ctypes.c_float)]


class OracleResult(ctypes.Structure):
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
        ("sample_rows", ctypes.c_uint64),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built: {LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        u64, i32, i = ctypes.c_uint64, ctypes.c_int32, ctypes.c_int
        vp = ctypes.c_void_p
        L.oracle_base_primes.argtypes = [u64, vp, ctypes.c_size_t]
        L.oracle_base_primes.restype = ctypes.c_size_t
        L.oracle_run_plain.argtypes = [u64, u64, i, ctypes.POINTER(OracleResult),
                                       vp, ctypes.c_size_t, vp, vp]
        L.oracle_run_plain.restype = i
        L.oracle_run_interval.argtypes = [u64, u64, i, ctypes.POINTER(OracleResult),
                                          vp, ctypes.c_size_t]
        L.oracle_run_interval.restype = i
        L.oracle_is_prime_u64.argtypes = [u64]
        L.oracle_is_prime_u64.restype = i
        L.oracle_is_prime_batch.argtypes = [vp, ctypes.c_size_t, vp, i]
        L.oracle_is_prime_batch.restype = None
        L.oracle_mr_mulmods.argtypes = [u64]
        L.oracle_mr_mulmods.restype = u64
        L.oracle_verify_gap.argtypes = [u64, i32]
        L.oracle_verify_gap.restype = i
        L.oracle_verify_gaps.argtypes = [vp, vp, ctypes.c_size_t, vp, i]
        L.oracle_verify_gaps.restype = None
        L.oracle_merit.argtypes = [u64, i32]
        L.oracle_merit.restype = ctypes.c_float
        _lib = L
    return _lib


def _to_dict(r, large, n_large):
    return {
        "total_primes": int(r.total_primes),
        "first_prime": int(r.first_prime),
        "last_prime": int(r.last_prime),
        "c1": int(r.c1),
        "c3": int(r.c3),
        "modq": np.ctypeslib.as_array(r.modq).astype(np.uint64).copy(),
        "gaps": np.ctypeslib.as_array(r.gaps).astype(np.int64).copy(),
        "modq_gap": np.ctypeslib.as_array(r.modq_gap).astype(np.int64)
                      .reshape(MODQ_ROWS, MODQ_MAX_GAP).copy(),
        "large_count": int(r.large_count),
        "large": [(int(large[i].position), int(large[i].gap), float(large[i].merit))
                  for i in range(n_large)],
    }


def base_primes(limit):
    """generate_base_primes (src/voss_v7_live.cu:96-108): primes 7..limit."""
    n = lib().oracle_base_primes(limit, None, 0)
    out = np.zeros(n, dtype=np.uint32)
    lib().oracle_base_primes(limit, out.ctypes.data, n)
    return out


def run_plain(N, seg_num, threads=0, large_cap=1 << 20, sample=False):
    """[1,N] with logical SEG_NUM segments (src/voss_v7_live.cu:337-705)."""
    r = OracleResult()
    large = (LargeGap * large_cap)()
    gs = ps = None
    if sample:
        gs = np.zeros(GAP_SEQ_SIZE, dtype=np.int32)
        ps = np.zeros(GAP_SEQ_SIZE, dtype=np.uint64)
    rc = lib().oracle_run_plain(N, seg_num, threads, ctypes.byref(r), large, large_cap,
                                gs.ctypes.data if sample else None,
                                ps.ctypes.data if sample else None)
    if rc != 0:
        raise ValueError(f"oracle_run_plain rejected N={N} SEG_NUM={seg_num}")
    d = _to_dict(r, large, min(int(r.large_count), large_cap))
    if sample:
        d["gap_seq"], d["pos_seq"] = gs, ps
        d["sample_rows"] = int(r.sample_rows)
    return d


def run_interval(lo, hi, threads=0, large_cap=1 << 20):
    """Primes in [lo, hi): every consecutive pair inside the interval."""
    r = OracleResult()
    large = (LargeGap * large_cap)()
    rc = lib().oracle_run_interval(lo, hi, threads, ctypes.byref(r), large, large_cap)
    if rc != 0:
        raise ValueError(f"oracle_run_interval rejected [{lo},{hi})")
    return _to_dict(r, large, min(int(r.large_count), large_cap))


def is_prime(n):
    """is_prime_u64 (src/verify_cpp.cpp:38-57)."""
    return bool(lib().oracle_is_prime_u64(int(n)))


def is_prime_batch(ns, threads=0):
    ns = np.ascontiguousarray(ns, dtype=np.uint64)
    out = np.zeros(ns.shape[0], dtype=np.uint8)
    lib().oracle_is_prime_batch(ns.ctypes.data, ns.shape[0], out.ctypes.data, threads)
    return out


def mr_mulmods(n):
    """Modular multiplications is_prime_u64 performs on n (work unit of MR)."""
    return int(lib().oracle_mr_mulmods(int(n)))


def verify_gap(p_after, gap):
    """verify_gap (src/verify_cpp.cpp:59-93)."""
    return bool(lib().oracle_verify_gap(int(p_after), int(gap)))


def verify_gaps(p_after, gaps, threads=0):
    pa = np.ascontiguousarray(p_after, dtype=np.uint64)
    g = np.ascontiguousarray(gaps, dtype=np.int32)
    ok = np.zeros(pa.shape[0], dtype=np.uint8)
    lib().oracle_verify_gaps(pa.ctypes.data, g.ctypes.data, pa.shape[0], ok.ctypes.data,
                             threads)
    return ok


def merit(p_before, gap):
    """(float)((double)g / log((double)p_before))  (src/voss_v7_live.cu:290-297)."""
    return float(lib().oracle_merit(int(p_before), int(gap)))
