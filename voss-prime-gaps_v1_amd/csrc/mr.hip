// mr.hip -- deterministic 12-base Miller-Rabin on gfx950.
//
// Same verdict as the reference's is_prime_u64 (src/verify_cpp.cpp:38-57):
// n < 2 is composite, n equal to one of the bases {2,...,37} is prime, n
// divisible by one of them is composite, otherwise n must be a strong
// probable prime to all 12 bases.  The reference multiplies with a 128-bit
// `%` (mul_mod, :24); here every product is a 64-bit Montgomery product
// (R = 2^64) built from 32-bit multiplies, so no division runs in the loop.
//
// Two stages keep the waves coherent: stage 1 does trial division and the
// base-2 test for every candidate (most composites stop here) and appends
// the survivors to a compact list with one atomic per wave; stage 2 runs
// bases 3..37 on the survivors only, which are almost all primes and so
// all take the same number of steps.
#include <hip/hip_runtime.h>

#include "montgomery.h"
#include "mr_launch.h"

namespace voss {

// Strong probable-prime test to the base whose Montgomery form is bm.
__device__ __forceinline__ bool sprp(const Mont& M, uint64_t bm, uint64_t d,
                                     int s) {
    uint64_t x = bm;
    const int top = 63 - __clzll((long long)d);
    for (int k = top - 1; k >= 0; k--) {
        x = mont_mul(x, x, M);
        if ((d >> k) & 1) x = mont_mul(x, bm, M);
    }
    const uint64_t mone = M.n - M.one;
    if (x == M.one || x == mone) return true;
    for (int j = 0; j < s - 1; j++) {
        x = mont_mul(x, x, M);
        if (x == mone) return true;
        if (x == M.one) return false;
    }
    return false;
}

__constant__ uint32_t c_bases[12] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};

// 0 = composite, 1 = prime, 2 = needs the witness loop.
__device__ __forceinline__ int trial(uint64_t n) {
    if (n < 2) return 0;
#pragma unroll
    for (int i = 0; i < 12; i++) {
        const uint64_t b = c_bases[i];
        if (n == b) return 1;
        if (n % b == 0) return 0;
    }
    return 2;
}

__global__ __launch_bounds__(256) void mr_stage1(const uint64_t* __restrict__ cand,
                                                 uint64_t count,
                                                 uint8_t* __restrict__ verdict,
                                                 uint32_t* __restrict__ surv,
                                                 uint32_t* __restrict__ nsurv) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const int lane = threadIdx.x & 63;
    for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < count;
         base += stride) {
        const uint64_t i = base + threadIdx.x;
        bool more = false;
        if (i < count) {
            const uint64_t n = cand[i];
            int v = trial(n);
            if (v == 2) {
                const Mont M = mont_setup(n);
                uint64_t d = n - 1;
                const int s = __ffsll((long long)d) - 1;
                d >>= s;
                const uint64_t b2 = add_mod(M.one, M.one, n);
                v = sprp(M, b2, d, s) ? 2 : 0;
            }
            more = (v == 2);
            verdict[i] = (uint8_t)(v == 1);
        }
        const uint64_t mask = __ballot(more);
        if (mask) {
            const int leader = __ffsll((long long)mask) - 1;
            uint32_t slot = 0;
            if (lane == leader) slot = atomicAdd(nsurv, (uint32_t)__popcll(mask));
            slot = __shfl(slot, leader);
            if (more)
                surv[slot + __popcll(mask & ((1ull << lane) - 1ull))] = (uint32_t)i;
        }
    }
}

__global__ __launch_bounds__(256) void mr_stage2(const uint64_t* __restrict__ cand,
                                                 const uint32_t* __restrict__ surv,
                                                 const uint32_t* __restrict__ nsurv,
                                                 uint8_t* __restrict__ verdict) {
    const uint32_t ns = *nsurv;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < ns; k += stride) {
        const uint32_t i = surv[k];
        const uint64_t n = cand[i];
        const Mont M = mont_setup(n);
        uint64_t d = n - 1;
        const int s = __ffsll((long long)d) - 1;
        d >>= s;
        uint64_t bm = add_mod(M.one, M.one, n);  // base 2
        uint32_t b = 2;
        bool prime = true;
        for (int t = 1; t < 12 && prime; t++) {
            const uint32_t nb = c_bases[t];
            for (; b < nb; b++) bm = add_mod(bm, M.one, n);
            prime = sprp(M, bm, d, s);
        }
        verdict[i] = prime ? 1 : 0;
    }
}

hipError_t launch_mr(const uint64_t* d_cand, uint64_t count, uint8_t* d_verdict,
                     uint32_t* d_surv, uint32_t* d_nsurv, int num_cus,
                     hipStream_t s) {
    if (count == 0) return hipSuccess;
    hipError_t e = hipMemsetAsync(d_nsurv, 0, sizeof(uint32_t), s);
    if (e != hipSuccess) return e;
    uint64_t blocks = (count + 255) / 256;
    const uint64_t cap = (uint64_t)num_cus * 32;
    if (blocks > cap) blocks = cap;
    hipLaunchKernelGGL(mr_stage1, dim3((uint32_t)blocks), dim3(256), 0, s, d_cand,
                       count, d_verdict, d_surv, d_nsurv);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(mr_stage2, dim3((uint32_t)blocks), dim3(256), 0, s, d_cand,
                       d_surv, d_nsurv, d_verdict);
    return hipGetLastError();
}

}  // namespace voss
