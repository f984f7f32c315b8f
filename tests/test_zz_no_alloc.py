"""voss_mr_batch_device_noalloc through the C ABI: the same verdicts as
voss_mr_batch_device on two calls in a row (the second smaller, so stale
survivor lists and counters from the first must not leak into it), no
change in the device's free memory across calls larger than any before on
this context, and a clear error beyond VOSS_MR_NOALLOC_MAX."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX = 4194304


def hip_free_bytes():
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def candidates(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.integers(2**40, 2**63, size=n, dtype=np.uint64) | np.uint64(1)
    x[::7] = np.uint64(2**61 - 1)  # primes among them, so every stage has work
    x[3::7] = np.uint64(3825123056546413051)  # rejected by base 37 only
    return x


def test_noalloc_matches_device_entry_point():
    import vossgpu
    L = vossgpu.lib()
    L.voss_mr_batch_device_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                               ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
    eng = vossgpu.Engine(0)
    try:
        n1, n2 = MAX, 300001
        dn, dv1, dv2 = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        vossgpu._check(L.voss_device_alloc(eng.handle, n1 * 8, ctypes.byref(dn)))
        vossgpu._check(L.voss_device_alloc(eng.handle, n1, ctypes.byref(dv1)))
        vossgpu._check(L.voss_device_alloc(eng.handle, n1, ctypes.byref(dv2)))
        try:
            ms = ctypes.c_double()
            # a first small call (the runtime loads the kernels); each later
            # call, larger than any before on this context, must leave the
            # device's free memory as it found it (the comparison calls of
            # voss_mr_batch_device in between do grow their context buffers)
            warm = candidates(1000, 1)
            vossgpu._check(L.voss_memcpy_h2d(eng.handle, dn, warm.ctypes.data, warm.nbytes))
            vossgpu._check(L.voss_mr_batch_device_noalloc(eng.handle, dn, 1000, dv1,
                                                          ctypes.byref(ms)))
            for call, (n, seed) in enumerate(((n1, 2), (n2, 3))):
                print(f"[gpu] no-alloc MR call {call}: {n} candidates", flush=True)
                x = candidates(n, seed)
                vossgpu._check(L.voss_memcpy_h2d(eng.handle, dn, x.ctypes.data, x.nbytes))
                free0 = hip_free_bytes()
                vossgpu._check(L.voss_mr_batch_device_noalloc(eng.handle, dn, n, dv1,
                                                              ctypes.byref(ms)))
                assert hip_free_bytes() == free0, "the no-allocation entry point allocated"
                got = np.zeros(n, dtype=np.uint8)
                vossgpu._check(L.voss_memcpy_d2h(eng.handle, got.ctypes.data, dv1, n))
                vossgpu._check(L.voss_mr_batch_device(eng.handle, dn, n, dv2, ctypes.byref(ms)))
                want = np.zeros(n, dtype=np.uint8)
                vossgpu._check(L.voss_memcpy_d2h(eng.handle, want.ctypes.data, dv2, n))
                assert np.array_equal(got, want)
                assert want[::7].all() and not want[3::7].any()
            # beyond the maximum: a clear error, nothing launched
            rc = L.voss_mr_batch_device_noalloc(eng.handle, dn, MAX + 1, dv1, None)
            assert rc == -1
            assert b"VOSS_MR_NOALLOC_MAX" in L.voss_last_error()
        finally:
            for p in (dn, dv1, dv2):
                L.voss_device_free(eng.handle, p)
    finally:
        eng.close()
