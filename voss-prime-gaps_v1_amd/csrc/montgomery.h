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

__device__ __forceinline__ uint64_t add_mod(uint64_t a, uint64_t b,
                                            uint64_t n) {
    const uint64_t s = a + b;
    return (s < a || s >= n) ? s - n : s;
}

__device__ __forceinline__ Mont mont_setup(uint64_t n) {
    Mont M;
    M.n = n;
    uint64_t x = n;  // correct to 3 bits for odd n
#pragma unroll
    for (int i = 0; i < 5; i++) x *= 2 - n * x;
    M.ninv = x;
    M.one = (n >> 63) ? (0 - n) : (0 - n) % n;
    return M;
}

}  // namespace voss
