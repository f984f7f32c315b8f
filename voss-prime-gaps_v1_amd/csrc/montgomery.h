// montgomery.h -- 64-bit Montgomery arithmetic (R = 2^64) for gfx950,
// shared by the Miller-Rabin kernels and the multiply-rate probe.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace voss {

struct Mont {
    uint64_t n, ninv, one;  // ninv = n^-1 mod 2^64, one = 2^64 mod n
};

__device__ __forceinline__ uint64_t mont_mul(uint64_t a, uint64_t b,
                                             const Mont& M) {
    const uint64_t lo = a * b;
    const uint64_t hi = __umul64hi(a, b);
    const uint64_t m = lo * M.ninv;
    const uint64_t mh = __umul64hi(m, M.n);
    const uint64_t r = hi - mh;
    return hi < mh ? r + M.n : r;
}

// a*a in Montgomery form: the 128-bit square from three 32x32 products
// (a = ah*2^32 + al: a^2 = ah^2*2^64 + al*ah*2^33 + al^2), then the same
// reduction as mont_mul.
__device__ __forceinline__ uint64_t mont_sqr(uint64_t a, const Mont& M) {
    const uint32_t al = (uint32_t)a, ah = (uint32_t)(a >> 32);
    const uint64_t ll = (uint64_t)al * al;
    const uint64_t lh = (uint64_t)al * ah;
    const uint64_t hh = (uint64_t)ah * ah;
    const uint64_t lo = ll + (lh << 33);
    const uint64_t hi = hh + (lh >> 31) + (lo < ll ? 1u : 0u);
    const uint64_t m = lo * M.ninv;
    const uint64_t mh = __umul64hi(m, M.n);
    const uint64_t r = hi - mh;
    return hi < mh ? r + M.n : r;
}

__device__ __forceinline__ uint64_t add_mod(uint64_t a, uint64_t b,
                                            uint64_t n) {
    const uint64_t s = a + b;
    return (s < a || s >= n) ? s - n : s;
}

// 2^64 mod n for odd n > 1 without a 64-bit division in the common case.
//   n >= 2^63:        2^64 - n.
//   2^62 <= n < 2^63: 2^64 - 2n, then at most two subtractions (2^64/n <= 4).
//   2^13 <= n < 2^62: q = trunc(2^64 / n) in double precision is within 1 of
//                     floor(2^64 / n) (relative error <= 2^-52, so the
//                     absolute error is <= 2^12 / n < 1/2); r = -q*n mod 2^64
//                     is then the remainder, or the remainder + n, or the
//                     remainder - n wrapped to [2^64 - n, 2^64) -- disjoint
//                     ranges because 2n < 2^63.
//   n < 2^13:         the 64-bit remainder (rare: small odd candidates).
__device__ __forceinline__ uint64_t r64_mod(uint64_t n) {
    if (n >> 63) return 0 - n;
    if (n >> 62) {
        uint64_t r = 0 - 2 * n;
        if (r >= n) r -= n;
        if (r >= n) r -= n;
        return r;
    }
    if (n >= (1u << 13)) {
        const uint64_t q = (uint64_t)(18446744073709551616.0 / (double)n);
        uint64_t r = 0 - q * n;
        if (r >= (0 - n)) r += n;   // q was one too large
        else if (r >= n) r -= n;    // q was one too small
        return r;
    }
    return (0 - n) % n;
}

__device__ __forceinline__ Mont mont_setup(uint64_t n) {
    Mont M;
    M.n = n;
    uint64_t x = n;  // correct to 3 bits for odd n
#pragma unroll
    for (int i = 0; i < 5; i++) x *= 2 - n * x;
    M.ninv = x;
    M.one = r64_mod(n);
    return M;
}

}  // namespace voss
