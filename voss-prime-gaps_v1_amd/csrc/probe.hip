// probe.hip -- roofline microbenchmarks for gfx950 (used by bench.py to
// state measured peaks beside the published ones): LDS bandwidth and
// atomic rates, the HBM copy used to calibrate the traffic counters, and
// the 64-bit Montgomery multiply rate that bounds Miller-Rabin; and the
// error bound of the double-precision reciprocal the bucket fill uses.  The only
// global write of the compute probes is one word per thread, so the work
// cannot be optimised away.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstring>

#include "fp_recip.h"
#include "montgomery.h"

namespace {

constexpr int kWords = 32768;  // 128 KB of LDS per workgroup

// Streaming 16-byte reads, conflict-free (consecutive lanes, consecutive
// 16-byte slots).
__global__ __launch_bounds__(1024) void lds_read_b128(uint32_t iters, uint32_t stride,
                                                     uint32_t* out) {
    __shared__ uint4 buf[kWords / 4];
    for (int i = threadIdx.x; i < kWords / 4; i += blockDim.x)
        buf[i] = make_uint4(i, i * 3u, i * 5u, i * 7u);
    __syncthreads();
    uint4 acc = make_uint4(0, 0, 0, 0);
    uint32_t idx = threadIdx.x;
    for (uint32_t it = 0; it < iters; it++) {
        const uint4 v = buf[idx];
        acc.x ^= v.x; acc.y += v.y; acc.z ^= v.z; acc.w += v.w;
        idx = (idx + stride + (acc.x & 1u)) & (kWords / 4 - 1);
    }
    out[blockIdx.x * blockDim.x + threadIdx.x] = acc.x ^ acc.y ^ acc.z ^ acc.w;
}

// ds_and_b32 at pseudo-random word addresses (the sieve's marking pattern).
__global__ __launch_bounds__(1024) void lds_and_random(uint32_t iters, uint32_t* out) {
    __shared__ uint32_t buf[kWords];
    for (int i = threadIdx.x; i < kWords; i += blockDim.x) buf[i] = 0xFFFFFFFFu;
    __syncthreads();
    uint32_t x = threadIdx.x * 2654435761u + blockIdx.x;
    for (uint32_t it = 0; it < iters; it++) {
        x = x * 1664525u + 1013904223u;
        atomicAnd(&buf[x >> 17], ~(1u << (x & 31u)));
    }
    __syncthreads();
    out[blockIdx.x * blockDim.x + threadIdx.x] = buf[threadIdx.x];
}

// ds_and_b32 at consecutive word addresses (no bank conflicts).
__global__ __launch_bounds__(1024) void lds_and_linear(uint32_t iters, uint32_t* out) {
    __shared__ uint32_t buf[kWords];
    for (int i = threadIdx.x; i < kWords; i += blockDim.x) buf[i] = 0xFFFFFFFFu;
    __syncthreads();
    uint32_t idx = threadIdx.x;
    for (uint32_t it = 0; it < iters; it++) {
        atomicAnd(&buf[idx], ~(1u << (it & 31u)));
        idx = (idx + 1024u) & (kWords - 1);
    }
    __syncthreads();
    out[blockIdx.x * blockDim.x + threadIdx.x] = buf[threadIdx.x];
}

// Plain device copy: reads and writes `words` 16-byte words once each (the
// HBM-traffic calibration for the FETCH_SIZE / WRITE_SIZE counters).
__global__ __launch_bounds__(256) void copy_b128(const uint4* __restrict__ src,
                                                 uint4* __restrict__ dst, size_t words) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += stride)
        dst[i] = src[i];
}

// Four independent chains of Montgomery products per lane, each lane with
// its own odd 64-bit modulus (as in a Miller-Rabin batch): the latency of
// one chain is hidden by the other three and by the other waves.
__global__ __launch_bounds__(256) void montmul_chains(uint32_t iters, unsigned long long* out) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t n = (0x9E3779B97F4A7C15ull * (t + 1)) | 0x8000000000000001ull;
    const voss::Mont M = voss::mont_setup(n);
    uint64_t x0 = M.one, x1 = M.one + 1, x2 = M.one + 2, x3 = M.one + 3;
    const uint64_t y = (0xD1B54A32D192ED03ull * (t + 7)) % n;
    for (uint32_t i = 0; i < iters; i++) {
        x0 = voss::mont_mul(x0, y, M);
        x1 = voss::mont_mul(x1, y, M);
        x2 = voss::mont_mul(x2, y, M);
        x3 = voss::mont_mul(x3, y, M);
    }
    out[t] = x0 ^ x1 ^ x2 ^ x3;
}

// The largest relative error |1 - p*recip_newton(p)| over the odd p in
// [lo, hi), as the bits of a non-negative double (they order like the
// values), one atomic max per thread.
__global__ void recip_error(uint32_t lo, uint32_t hi, unsigned long long* out) {
    double m = 0.0;
    const uint32_t n = (hi - lo) / 2u;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double pd = (double)(lo + 2u * i);
        const double e = __builtin_fabs(__builtin_fma(-pd, voss::recip_newton(pd), 1.0));
        m = e > m ? e : m;
    }
    atomicMax(out, (unsigned long long)__double_as_longlong(m));
}

}  // namespace

// 64-bit Montgomery products per second on the whole device.
extern "C" int voss_probe_montmul(int num_cus, double* rate) {
    const uint32_t blocks = (uint32_t)num_cus * 8, threads = 256, iters = 4096;
    unsigned long long* out = nullptr;
    if (hipMalloc((void**)&out, (size_t)blocks * threads * 8) != hipSuccess) return -1;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    float best = 1e30f;
    for (int rep = 0; rep < 4; rep++) {
        (void)hipEventRecord(e0, 0);
        hipLaunchKernelGGL(montmul_chains, dim3(blocks), dim3(threads), 0, 0, iters, out);
        (void)hipEventRecord(e1, 0);
        (void)hipEventSynchronize(e1);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (rep > 0 && ms < best) best = ms;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(out);
    *rate = (double)blocks * threads * iters * 4.0 / (best * 1e-3);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// One copy of `bytes` (multiple of 16) between two fresh device buffers.
extern "C" int voss_probe_copy(size_t bytes, double* ms_out) {
    uint4 *src = nullptr, *dst = nullptr;
    if (hipMalloc((void**)&src, bytes) != hipSuccess) return -1;
    if (hipMalloc((void**)&dst, bytes) != hipSuccess) {
        (void)hipFree(src);
        return -1;
    }
    (void)hipMemset(src, 1, bytes);
    (void)hipDeviceSynchronize();
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventRecord(e0, 0);
    hipLaunchKernelGGL(copy_b128, dim3(4096), dim3(256), 0, 0, src, dst, bytes / 16);
    (void)hipEventRecord(e1, 0);
    (void)hipEventSynchronize(e1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (ms_out) *ms_out = ms;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(src);
    (void)hipFree(dst);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Max relative error of recip_newton over the odd p in [lo, hi) (lo odd).
extern "C" int voss_probe_recip(uint32_t lo, uint32_t hi, double* max_err) {
    if (!(lo & 1u) || hi <= lo) return -1;
    unsigned long long* d = nullptr;
    if (hipMalloc((void**)&d, 8) != hipSuccess) return -1;
    unsigned long long h = 0;
    int rc = -1;
    if (hipMemcpy(d, &h, 8, hipMemcpyHostToDevice) == hipSuccess) {
        hipLaunchKernelGGL(recip_error, dim3(4096), dim3(256), 0, 0, lo, hi, d);
        if (hipDeviceSynchronize() == hipSuccess && hipMemcpy(&h, d, 8, hipMemcpyDeviceToHost) == hipSuccess) {
            std::memcpy(max_err, &h, 8);
            rc = 0;
        }
    }
    (void)hipFree(d);
    return rc;
}

// kind 0: read b128 -> bytes/s; 1: ds_and random -> ops/s; 2: ds_and linear.
extern "C" int voss_probe_lds(int kind, int num_cus, double* rate) {
    const uint32_t iters = kind == 0 ? 4096 : 8192;
    const uint32_t grid = (uint32_t)num_cus;
    uint32_t* out = nullptr;
    if (hipMalloc((void**)&out, (size_t)grid * 1024 * 4) != hipSuccess) return -1;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    float best = 1e30f;
    for (int rep = 0; rep < 4; rep++) {
        (void)hipEventRecord(e0, 0);
        if (kind == 0)
            hipLaunchKernelGGL(lds_read_b128, dim3(grid), dim3(1024), 0, 0, iters, 1024u, out);
        else if (kind == 1) hipLaunchKernelGGL(lds_and_random, dim3(grid), dim3(1024), 0, 0, iters, out);
        else hipLaunchKernelGGL(lds_and_linear, dim3(grid), dim3(1024), 0, 0, iters, out);
        (void)hipEventRecord(e1, 0);
        (void)hipEventSynchronize(e1);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (rep > 0 && ms < best) best = ms;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(out);
    const double ops = (double)grid * 1024.0 * iters;
    *rate = (kind == 0 ? ops * 16.0 : ops) / (best * 1e-3);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
