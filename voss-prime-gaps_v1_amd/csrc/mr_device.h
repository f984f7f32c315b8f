// mr_device.h -- device pieces of the deterministic 12-base Miller-Rabin
// (is_prime_u64, src/verify_cpp.cpp:38-57), shared by the Miller-Rabin
// kernels (mr.hip) and the serving loop (serve_kernel.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "montgomery.h"

namespace voss {

// The witness loop of the strong probable-prime test after x = b^d: true
// iff x = 1 or x reaches -1 within s - 1 squarings.
__device__ __forceinline__ bool sprp_tail(const Mont& M, uint64_t x, int s) {
    const uint64_t mone = M.n - M.one;
    if (x == M.one || x == mone) return true;
    for (int j = 0; j < s - 1; j++) {
        x = mont_sqr(x, M);
        if (x == mone) return true;
        if (x == M.one) return false;
    }
    return false;
}

// Base 2: left-to-right binary powering where "multiply by the base" is a
// modular doubling, so the lanes whose exponent bit is set pay 4 integer
// ops instead of a Montgomery product.
__device__ __forceinline__ bool sprp2(const Mont& M, uint64_t d, int s) {
    uint64_t x = add_mod(M.one, M.one, M.n);  // 2 in Montgomery form
    const int top = 63 - __clzll((long long)d);
    for (int k = top - 1; k >= 0; k--) {
        x = mont_sqr(x, M);
        if ((d >> k) & 1) x = add_mod(x, x, M.n);
    }
    return sprp_tail(M, x, s);
}

// Other bases: fixed 3-bit windows, so every lane of a wave runs the same
// squarings and products (a binary ladder runs a product at every bit where
// any lane of the wave has a 1).  ktop = highest digit index any lane of
// the wave needs (d < 2^63: at most 20); tbl holds b^0..b^7 in Montgomery
// form, 8 rows of the block's BLOCK lanes.
template <int BLOCK>
__device__ __forceinline__ bool sprp_window(const Mont& M, uint64_t bm, uint64_t d, int s,
                                            int ktop, uint64_t* tbl) {
    const uint32_t t = threadIdx.x;
    tbl[t] = M.one;
    tbl[BLOCK + t] = bm;
    uint64_t p = bm;
#pragma unroll
    for (int k = 2; k < 8; k++) {
        p = mont_mul(p, bm, M);
        tbl[k * BLOCK + t] = p;
    }
    uint64_t x = tbl[((uint32_t)(d >> (3 * ktop)) & 7u) * BLOCK + t];
    for (int k = ktop - 1; k >= 0; k--) {
        x = mont_sqr(x, M);
        x = mont_sqr(x, M);
        x = mont_sqr(x, M);
        x = mont_mul(x, tbl[((uint32_t)(d >> (3 * k)) & 7u) * BLOCK + t], M);
    }
    return sprp_tail(M, x, s);
}

// the bases after 2
static __constant__ uint32_t c_bases[11] = {3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};

// Divisibility of a 64-bit n by an odd constant b without a division:
// b | n iff n * b^-1 mod 2^64 <= (2^64 - 1) / b (the multiples of b are
// exactly the images of 0..(2^64-1)/b under multiplication by b^-1).
constexpr uint64_t inv_mod64(uint64_t b) {
    uint64_t x = b;  // correct to 3 bits for odd b
    for (int i = 0; i < 5; i++) x *= 2 - b * x;
    return x;
}
template <uint64_t B>
__device__ __forceinline__ bool divisible(uint64_t n) {
    return n * inv_mod64(B) <= ~0ull / B;
}

// 0 = composite, 1 = prime, 2 = needs the witness loop (the reference's
// order: n < 2, n equal to a base, n divisible by a base).
__device__ __forceinline__ int trial(uint64_t n) {
    // the primes up to 37 as bits: 2,3,5,7,11,13,17,19,23,29,31,37
    constexpr uint64_t kSmallPrimes = (1ull << 2) | (1ull << 3) | (1ull << 5) | (1ull << 7) |
                                      (1ull << 11) | (1ull << 13) | (1ull << 17) |
                                      (1ull << 19) | (1ull << 23) | (1ull << 29) |
                                      (1ull << 31) | (1ull << 37);
    if (n < 38) return (int)((kSmallPrimes >> n) & 1u);  // composites <= 37 have a base factor
    if ((n & 1) == 0 || divisible<3>(n) || divisible<5>(n) || divisible<7>(n) ||
        divisible<11>(n) || divisible<13>(n) || divisible<17>(n) || divisible<19>(n) ||
        divisible<23>(n) || divisible<29>(n) || divisible<31>(n) || divisible<37>(n))
        return 0;
    return 2;
}

}  // namespace voss
