// mr.hip -- deterministic 12-base Miller-Rabin on gfx950.
//
// Same verdict as the reference's is_prime_u64 (src/verify_cpp.cpp:38-57):
// n < 2 is composite, n equal to one of the bases {2,...,37} is prime, n
// divisible by one of them is composite, otherwise n must be a strong
// probable prime to all 12 bases.  The reference multiplies with a 128-bit
// `%` (mul_mod, :24); here every product is a 64-bit Montgomery product
// (R = 2^64) built from 32-bit multiplies, so no division runs in the loop.
//
// Three stages keep the waves coherent, each on a compact list of what the
// previous one left open: trial division by the 12 bases for every
// candidate, the base-2 test for those no base divides (most composites
// stop here), then bases 3..37 on the survivors only, which are almost all
// primes and so all take the same number of steps.
#include <hip/hip_runtime.h>

#include <mutex>

#include "montgomery.h"
#include "mr_device.h"
#include "mr_launch.h"
#include "sieve_launch.h"

namespace voss {

// Survivors are appended with one atomic per workgroup pass (wave counts by
// ballot, their prefix in LDS): a single counter hit once per wave
// serialises at the L2 and bounded the first stage.  Every thread of the
// block calls this once per pass.
// With `from_end` the list is filled downwards from list[cap - 1] instead.
__device__ __forceinline__ void append_block(bool more, uint32_t value, uint32_t* list,
                                             uint32_t* counter, uint32_t* s_wave,
                                             uint32_t* s_base, bool from_end = false,
                                             uint32_t cap = 0) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nwaves = (int)(blockDim.x >> 6);
    const uint64_t mask = __ballot(more);
    if (lane == 0) s_wave[wave] = (uint32_t)__popcll(mask);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < nwaves; w++) {
            const uint32_t c = s_wave[w];
            s_wave[w] = t;
            t += c;
        }
        *s_base = t ? atomicAdd(counter, t) : 0u;
    }
    __syncthreads();
    if (more) {
        const uint32_t at = *s_base + s_wave[wave] + __popcll(mask & ((1ull << lane) - 1ull));
        list[from_end ? cap - 1 - at : at] = value;
    }
    __syncthreads();  // s_wave and s_base are rewritten by the next pass
}

// Stage 1a: trial division for every candidate; the verdict of the ones it
// decides is written, the others (index) go to list 1.  Candidates that a
// small base divides are most of a random set, so the powering stages
// below run on a compact list instead of waves with those lanes idle.
__global__ __launch_bounds__(1024) void mr_trial(const uint64_t* __restrict__ cand,
                                                 uint64_t count, uint8_t* __restrict__ verdict,
                                                 uint32_t* __restrict__ list1,
                                                 uint32_t* __restrict__ n1) {
    __shared__ uint32_t s_wave[16];
    __shared__ uint32_t s_base;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    // the loop bound depends on blockIdx only, so every thread of the
    // block takes every pass and reaches every barrier
    for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < count;
         base += stride) {
        const uint64_t i = base + threadIdx.x;
        bool more = false;
        if (i < count) {
            const int v = trial(cand[i]);
            more = (v == 2);
            verdict[i] = (uint8_t)(v == 1);
        }
        append_block(more, (uint32_t)i, list1, n1, s_wave, &s_base);
    }
}

// The windowed powering of stage 2 runs as many base-8 digits as the
// longest exponent of the wave, so its list is kept in two parts: d < 2^57
// from the front, longer exponents from the back (the two never meet: both
// together hold at most `cap` entries).
constexpr int kLongExponentBits = 57;

// Stage 1b: the base-2 strong test on list 1; the candidates that pass go to
// list 2 (their verdict is already 0 from stage 1a and is set by stage 2).
__global__ __launch_bounds__(1024) void mr_base2(const uint64_t* __restrict__ cand,
                                                 const uint32_t* __restrict__ list1,
                                                 const uint32_t* __restrict__ n1,
                                                 uint32_t* __restrict__ list2, uint32_t cap,
                                                 uint32_t* __restrict__ n2) {
    __shared__ uint32_t s_wave[16];
    __shared__ uint32_t s_base;
    const uint32_t count = *n1;
    const uint32_t stride = gridDim.x * blockDim.x;
    // the loop bound is uniform over the block (blocks past the list exit)
    for (uint32_t base = blockIdx.x * blockDim.x; base < count; base += stride) {
        const uint32_t k = base + threadIdx.x;
        bool more = false, longer = false;
        uint32_t i = 0;
        if (k < count) {
            i = list1[k];
            const uint64_t n = cand[i];
            const Mont M = mont_setup(n);
            uint64_t d = n - 1;
            const int s = __ffsll((long long)d) - 1;
            d >>= s;
            more = sprp2(M, d, s);
            longer = (d >> kLongExponentBits) != 0;
        }
        append_block(more && !longer, i, list2, n2, s_wave, &s_base);
        append_block(more && longer, i, list2, n2 + 1, s_wave, &s_base, true, cap);
    }
}

__global__ __launch_bounds__(256) void mr_stage2(const uint64_t* __restrict__ cand,
                                                 const uint32_t* __restrict__ surv,
                                                 uint32_t cap,
                                                 const uint32_t* __restrict__ nsurv,
                                                 uint8_t* __restrict__ verdict) {
    __shared__ uint64_t tbl[8 * 256];
    // the front part of the list, then the back part read downwards
    const uint32_t nfront = nsurv[0];
    const uint32_t ns = nfront + nsurv[1];
    const uint32_t stride = gridDim.x * blockDim.x;
    // the loop bound is uniform over the block, so the whole wave is present
    // for the digit-count reduction below
    for (uint32_t base = blockIdx.x * blockDim.x; base < ns; base += stride) {
        const uint32_t k = base + threadIdx.x;
        const bool active = k < ns;
        uint64_t n = 3, d = 1;
        int s = 1;
        uint32_t i = 0;
        if (active) {
            i = surv[k < nfront ? k : cap - 1 - (k - nfront)];
            n = cand[i];
            d = n - 1;
            s = __ffsll((long long)d) - 1;
            d >>= s;
        }
        // highest base-8 digit of d that any lane of this wave needs (d < 2^63)
        int ktop = active ? (63 - __clzll((long long)d)) / 3 : 0;
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) ktop = max(ktop, __shfl_xor(ktop, o));
        ktop = min(ktop, 20);
        if (!active) continue;
        const Mont M = mont_setup(n);
        // each base in Montgomery form by modular additions from the last
        // (n > 37 here, so no base is 0 mod n)
        uint64_t bm = add_mod(M.one, M.one, n);  // base 2
        uint32_t b = 2;
        bool prime = true;
        for (int t = 0; t < 11 && prime; t++) {
            for (const uint32_t nb = c_bases[t]; b < nb; b++) bm = add_mod(bm, M.one, n);
            prime = sprp_window<256>(M, bm, d, s, ktop, tbl);
        }
        verdict[i] = prime ? 1 : 0;
    }
}

// Strong probable-prime test to base b by binary powering (the rare path of
// verify_interior: only interior numbers that pass base 2 reach it).
__device__ __forceinline__ bool sprp_binary(const Mont& M, uint64_t bm, uint64_t d, int s) {
    uint64_t x = bm;
    for (int k = 62 - __clzll((long long)d); k >= 0; k--) {
        x = mont_sqr(x, M);
        if ((d >> k) & 1) x = mont_mul(x, bm, M);
    }
    return sprp_tail(M, x, s);
}

// is_prime_u64 of the reference (src/verify_cpp.cpp:38-57), one thread.
__device__ __forceinline__ bool is_prime_one(uint64_t n) {
    const int v = trial(n);
    if (v != 2) return v == 1;
    const Mont M = mont_setup(n);
    uint64_t d = n - 1;
    const int s = __ffsll((long long)d) - 1;
    d >>= s;
    if (!sprp2(M, d, s)) return false;
    uint64_t bm = add_mod(M.one, M.one, n);  // base 2
    uint32_t b = 2;
    for (int t = 0; t < 11; t++) {
        for (const uint32_t nb = c_bases[t]; b < nb; b++) bm = add_mod(bm, M.one, n);
        if (!sprp_binary(M, bm, d, s)) return false;
    }
    return true;
}

// The interior of verify_gap (src/verify_cpp.cpp:63-92): with both ends
// prime, the reference's odd sieve by the primes up to isqrt(p_after) marks
// every number of (p_before, p_after) exactly when none of them is prime
// (a composite n < p_after has an odd prime factor <= isqrt(p_after), whose
// first multiple after p_before the sieve reaches; the only other marks, of
// a prime q <= isqrt(p_after) inside the interval, need p_after < 7, where
// both agree).  So one wave per gap tests the interior numbers coprime to
// 30 (and 3, 5 themselves) with the same Miller-Rabin, its lanes on 64
// consecutive odd numbers at a time, and clears ok[i] at the first prime
// (a wave vote ends the loop for every lane at once).  Every gap between
// primes below 2^64 is at most 1550, so a wave stops within 13 rounds
// whatever gap it is given.
// A claimed gap larger than p_after (p_before wraps below 0 to near 2^64,
// both ends prime) fails in the reference: its sieve runs over the numbers
// p_before + 1 .. 2^64 - 1, 0, 1, .., p_after - 1 in that order, and the
// number 1 there sits at the odd index gap - p_after; only the multiples of
// an odd q <= isqrt(p_after) < 46341 whose first multiple after p_before
// passes 2^64 are marked, starting from (that multiple - 2^64) or that plus
// q, which is 1 only if q divides 2^64 + 1 = 274177 * 67280421310721.  So
// such rows are cleared here.  ok[i] enters as "both ends prime".
__global__ __launch_bounds__(256) void verify_interior_kernel(
    const uint64_t* __restrict__ p_after, const int32_t* __restrict__ gap,
    uint8_t* __restrict__ ok, uint32_t count) {
    const uint32_t i = blockIdx.x * 4u + (threadIdx.x >> 6);  // one gap per wave
    const uint32_t lane = threadIdx.x & 63u;
    if (i >= count || !ok[i]) return;  // uniform over the wave
    const int32_t g = gap[i];
    const uint64_t pa = p_after[i];
    if (g < 2) return;
    if (pa < (uint64_t)g) {
        if (lane == 0) ok[i] = 0;
        return;
    }
    const uint64_t pb = pa - (uint64_t)g;
    // interior numbers as offsets d from p_before (d < g < 2^31, so no sum
    // wraps even next to 2^64); the first odd one is pb + 1 or pb + 2.  The
    // loop bound is uniform over the wave, so every lane reaches the vote.
    for (uint32_t off = (pb & 1u) ? 2u : 1u; off < (uint32_t)g; off += 128u) {
        const uint32_t d = off + 2u * lane;
        bool prime = false;
        if (d < (uint32_t)g) {
            const uint64_t n = pb + d;
            if (n <= 5 || (n % 3ull != 0 && n % 5ull != 0)) prime = is_prime_one(n);
        }
        if (__any(prime)) {
            if (lane == 0) ok[i] = 0;
            return;
        }
    }
}

hipError_t launch_verify_interior(const uint64_t* d_p_after, const int32_t* d_gap,
                                  uint8_t* d_ok, uint32_t count, hipStream_t s) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(verify_interior_kernel, dim3((count + 3) / 4), dim3(256), 0, s,
                       d_p_after, d_gap, d_ok, count);
    return hipGetLastError();
}

// Scratch of the Miller-Rabin entry point that allocates nothing: the two
// survivor lists of up to kMrNoAllocMax candidates, then the three list
// counters.  One copy per GPU, held until the process ends (32 MiB).  Host
// code never passes the array's name to a kernel: that is a host-side
// stand-in, not the device address; mr_scratch_address() gives the address.
static __device__ __align__(256) uint32_t scratch[2 * kMrNoAllocMax + 4];

hipError_t mr_scratch_address(int device, uint32_t** out) {
    // looked up once per device, with that device current (the caller's
    // context setup), and kept: each GPU holds its own copy of the array
    static std::mutex mu;
    static uint32_t* addr[kMaxDevices] = {};
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> g(mu);
    if (!addr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        addr[device] = static_cast<uint32_t*>(p);
    }
    *out = addr[device];
    return hipSuccess;
}

hipError_t launch_mr(const uint64_t* d_cand, uint64_t count, uint8_t* d_verdict,
                     uint32_t* d_surv, uint32_t* d_nsurv, int num_cus,
                     hipStream_t s) {
    if (count == 0) return hipSuccess;
    // a fill kernel, not hipMemsetAsync, so the sequence can be captured
    hipError_t e = launch_fill_u32(d_nsurv, 3, 0, s);
    if (e != hipSuccess) return e;
    // one thread per candidate (the kernels grid-stride only beyond 2^32
    // threads): a capped grid with a few candidates per thread leaves its
    // last pass half empty; blocks of the later stages beyond their list's
    // length exit at once
    (void)num_cus;
    uint32_t* list1 = d_surv;
    uint32_t* list2 = d_surv + count;
    uint64_t blocks1 = (count + 1023) / 1024;
    if (blocks1 > (1u << 22)) blocks1 = 1u << 22;
    hipLaunchKernelGGL(mr_trial, dim3((uint32_t)blocks1), dim3(1024), 0, s, d_cand, count,
                       d_verdict, list1, d_nsurv);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(mr_base2, dim3((uint32_t)blocks1), dim3(1024), 0, s, d_cand, list1,
                       d_nsurv, list2, (uint32_t)count, d_nsurv + 1);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    uint64_t blocks = (count + 255) / 256;
    if (blocks > (1u << 24)) blocks = 1u << 24;
    hipLaunchKernelGGL(mr_stage2, dim3((uint32_t)blocks), dim3(256), 0, s, d_cand,
                       list2, (uint32_t)count, d_nsurv + 1, d_verdict);
    return hipGetLastError();
}

}  // namespace voss
