// sieve.hip -- the wheel-30 window sieve kernels for gfx950, launched per
// plan round (the batch, interval and plain forms).
//
// Replaces, in one persistent kernel, the reference's per-segment chain
// (src/voss_v7_live.cu): one sieve_w30_seg_kernel launch per base prime
// over a 1 GB HBM bitmap (:120-143, :480-502), extract_w30_seg_kernel
// (:145-179), thrust::sort (:522-524), mod4_count_kernel /
// modq_count_kernel (:181-232) and gaps_w30_seg_kernel (:255-312).
//
// One workgroup = one LDS window at a time (sieve_window in
// sieve_device.h); a workgroup's first window is blockIdx.x, its next ones
// come from a device counter.  One wave fetches the next window's index and
// descriptor during marking, so consecutive windows are not separated by
// global-memory round trips.  HBM traffic per window is the descriptor, one
// 24-byte record, and the (rare) large-gap records; histograms are merged
// into HBM once per workgroup per job.
#include <hip/hip_runtime.h>

#include "sieve_device.h"
#include "sieve_launch.h"
#include "wheel.h"

namespace voss {

template <int WIN, int BLOCK, bool EMIT>
__global__ __launch_bounds__(BLOCK) void sieve_kernel(SieveArgs a) {
    __shared__ SieveShared<WIN, BLOCK, EMIT> S;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    sieve_init_lds(a, S);
    uint32_t cur_job = blockIdx.x < a.ndesc ? a.desc[blockIdx.x].job : 0;

    // windows are taken from a counter, not by a fixed stride: window costs
    // vary (bucketed or not, partial, base-prime counts), and a fixed share
    // per workgroup left a long tail in batches of mixed intervals.  A
    // workgroup's windows still come in increasing order.
    // The next window's index, descriptor and presieve table offsets are
    // fetched by the last wave during this window's marking: the global
    // atomic, the descriptor load and the 64-bit remainders are off the path
    // from one window to the next.
    bool prefetched = false;  // uniform: D of windows after the first is in S.desc
    for (uint32_t d = blockIdx.x; d < a.ndesc;) {
        SegDesc D;
        if (prefetched) {
            uint32_t* dd = reinterpret_cast<uint32_t*>(&D);
#pragma unroll
            for (int k = 0; k < (int)(sizeof(SegDesc) / 4); k++)
                dd[k] = __builtin_amdgcn_readfirstlane(S.desc[k]);
        } else {
            D = a.desc[d];
            sieve_set_patbase(a, S, D.w0, tid);
        }
        if (D.job != cur_job) {
            sieve_flush(S, &a.acc[cur_job]);
            cur_job = D.job;
        }
        WinIO io;
        io.acc = &a.acc[D.job];
        io.large = a.large;
        io.large_n = a.large_n;
        io.large_cap = a.large_cap;
        io.job = D.job;
        io.bent = nullptr;
        io.nbent = 0;
        if (D.bucket != kNoBucket) {
            // buckets start 16-byte aligned; a count beyond the room (an
            // overflowed fill) is clamped
            const unsigned long long e0 = a.bucket_off[D.bucket];
            const unsigned long long cap = a.bucket_off[D.bucket + 1] - e0;
            io.bent = a.bucket_ent + e0;
            io.nbent = (uint32_t)min((unsigned long long)a.bucket_cnt[D.bucket], cap);
        }
        io.rec = &a.rec[d];
        io.emit_at = EMIT ? a.emit_off[d] : 0ull;
        sieve_window(a, S, D, io, [&] {
            uint32_t nx = 0;
            if (lane == 0) nx = gridDim.x + atomicAdd(a.win_next, 1u);
            nx = __builtin_amdgcn_readfirstlane(nx);
            if (lane == 0) S.next = nx;
            if (nx < a.ndesc) {
                const uint32_t* src = reinterpret_cast<const uint32_t*>(a.desc + nx);
                if (lane < (int)(sizeof(SegDesc) / 4)) S.desc[lane] = src[lane];
                sieve_set_patbase(a, S, a.desc[nx].w0, lane);
            }
        });
        d = S.next;
        prefetched = true;
    }
    sieve_flush(S, &a.acc[cur_job]);
}

// Joins the last prime of one window with the first prime of the next
// non-empty window of the same job (the host boundary gap of the
// reference, src/voss_v7_live.cu:551-562, generalised to every window
// edge).  Such a pair counts in-segment only when both windows carry the
// same logical segment.
// When every window of the workgroup belongs to one job (always in a plain
// run), its pair counts are summed in LDS and flushed with one global
// atomic per nonzero bin: in a plain run every thread's pair lands on the
// few hundred bins of one job, and global atomics on the same addresses
// serialise.  The flag is uniform over the workgroup, so are its barriers.
__global__ void stitch_kernel(SieveArgs a) {
    __shared__ uint32_t s_hist[kHistBins];
    __shared__ uint32_t s_mq[kModqRows * kModqBins];
    const uint32_t dlo = blockIdx.x * blockDim.x;  // < ndesc: one block per 256
    const uint32_t dhi = min(dlo + blockDim.x, a.ndesc) - 1u;
    const uint32_t job0 = a.desc[dlo].job;
    const bool local = a.desc[dhi].job == job0;
    if (local) {
        for (uint32_t i = threadIdx.x; i < (uint32_t)kHistBins; i += blockDim.x) s_hist[i] = 0;
        for (uint32_t i = threadIdx.x; i < (uint32_t)(kModqRows * kModqBins); i += blockDim.x)
            s_mq[i] = 0;
        __syncthreads();
    }
    const uint32_t d = dlo + threadIdx.x;
    do {  // one pass; break = this window adds nothing
        if (d >= a.ndesc) break;
        const SegRec rd = a.rec[d];
        if (rd.count == 0) break;
        const uint32_t job = a.desc[d].job;
        int e = (int)d - 1;
        for (; e >= 0 && a.desc[e].job == job; e--)
            if (a.rec[e].count) break;
        if (e < 0 || a.desc[e].job != job) break;
        const unsigned long long p = a.rec[e].last;
        const unsigned long long g64 = rd.first - p;
        JobAcc* acc = &a.acc[job];
        if (g64 >= 2ull * kHistBins) {
            atomicOr(&acc->error, 1ull);
            break;
        }
        const uint32_t g = (uint32_t)g64;
        if (local) atomicAdd(&s_hist[g >> 1], 1u);
        else atomicAdd(&acc->hist[g >> 1], 1ull);
        if (a.desc[e].lseg != a.desc[d].lseg) break;
        if (g < 100u) {
            const uint32_t o = (uint32_t)c_residx[p % 30ull];
            const uint32_t i6 = q6row_of(o) * kModqBins + (g >> 1);
            const uint32_t i30 = (2u + o) * kModqBins + (g >> 1);
            if (local) {
                atomicAdd(&s_mq[i6], 1u);
                atomicAdd(&s_mq[i30], 1u);
            } else {
                atomicAdd(&acc->modq[i6], 1ull);
                atomicAdd(&acc->modq[i30], 1ull);
            }
        } else if (g > 100u) {
            if (g >= (uint32_t)kLargeAlways || p <= a.lg_thresh[g >> 1]) {
                unsigned long long slot = atomicAdd(a.large_n, 1ull);
                atomicAdd(&acc->large_count, 1ull);
                if (slot < a.large_cap) {
                    LargeRaw r;
                    r.p_before = p;
                    r.gap = g;
                    r.job = job;
                    a.large[slot] = r;
                }
            }
        }
    } while (false);
    if (local) {
        __syncthreads();
        JobAcc* acc = &a.acc[job0];
        for (uint32_t i = threadIdx.x; i < (uint32_t)kHistBins; i += blockDim.x)
            if (s_hist[i]) atomicAdd(&acc->hist[i], (unsigned long long)s_hist[i]);
        for (uint32_t i = threadIdx.x; i < (uint32_t)(kModqRows * kModqBins); i += blockDim.x)
            if (s_mq[i]) atomicAdd(&acc->modq[i], (unsigned long long)s_mq[i]);
    }
}

__global__ void init_acc_kernel(JobAcc* acc, uint32_t njobs) {
    const uint32_t job = blockIdx.x;
    if (job >= njobs) return;
    JobAcc* a = &acc[job];
    unsigned long long* w = reinterpret_cast<unsigned long long*>(a);
    const uint32_t nw = sizeof(JobAcc) / 8;
    for (uint32_t i = threadIdx.x; i < nw; i += blockDim.x) w[i] = 0;
    __syncthreads();
    if (threadIdx.x == 0) a->first = ~0ull;
}


// Large-prime bucket passes (bucket_task in sieve_device.h), one task per
// workgroup at a time, tasks taken from a device counter (largest first:
// the host orders them).
template <bool FILL>
__global__ __launch_bounds__(1024) void bucket_kernel(const BucketTask* __restrict__ tasks,
                                                      uint32_t ntask,
                                                      const PrimeRec* __restrict__ primes,
                                                      const uint32_t* __restrict__ p32,
                                                      uint32_t* __restrict__ bucket_cnt,
                                                      const unsigned long long* __restrict__ bucket_off,
                                                      uint32_t* __restrict__ entries,
                                                      uint32_t* __restrict__ overflow,
                                                      uint32_t* __restrict__ task_next, int log2w) {
    __shared__ uint32_t cnt[kBucketTileMax];
    __shared__ uint32_t lim[FILL ? kBucketTileMax : 1];
    __shared__ uint32_t s_task;
    while (true) {
        if (threadIdx.x == 0) s_task = atomicAdd(task_next, 1u);
        __syncthreads();
        const uint32_t t = s_task;
        if (t >= ntask) break;  // uniform: every thread read the same s_task
        const BucketTask T = tasks[t];
        bucket_task<FILL>(T, primes, p32, bucket_cnt, bucket_off, entries, overflow, log2w, cnt, lim);
    }
}

__global__ void add_acc_kernel(JobAcc* __restrict__ dst, const JobAcc* __restrict__ src) {
    JobAcc& d = dst[blockIdx.x];
    const JobAcc& s = src[blockIdx.x];
    constexpr uint32_t kSums = offsetof(JobAcc, first) / sizeof(unsigned long long);
    unsigned long long* dp = reinterpret_cast<unsigned long long*>(&d);
    const unsigned long long* sp = reinterpret_cast<const unsigned long long*>(&s);
    for (uint32_t i = threadIdx.x; i < kSums; i += blockDim.x) dp[i] += sp[i];
    if (threadIdx.x == 0) {
        d.first = min(d.first, s.first);
        d.last = max(d.last, s.last);
        d.large_count += s.large_count;
        d.error |= s.error;
    }
}

__global__ void fill_u32_kernel(uint32_t* p, size_t n, uint32_t v) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}
// ------------------------------------------------------------------ host

hipError_t launch_init_acc(JobAcc* acc, uint32_t njobs, hipStream_t s) {
    if (njobs == 0) return hipSuccess;
    hipLaunchKernelGGL(init_acc_kernel, dim3(njobs), dim3(256), 0, s, acc, njobs);
    return hipGetLastError();
}

hipError_t launch_sieve(const SieveArgs& a, int window_bytes, uint32_t grid,
                        hipStream_t s) {
    if (a.ndesc == 0) return hipSuccess;
    if (grid > a.ndesc) grid = a.ndesc;
    switch (window_bytes) {
        case 131072:
            hipLaunchKernelGGL((sieve_kernel<131072, 1024, false>), dim3(grid), dim3(1024), 0, s, a);
            break;
        case 65536:
            hipLaunchKernelGGL((sieve_kernel<65536, 512, false>), dim3(grid), dim3(512), 0, s, a);
            break;
        case 32768:
            hipLaunchKernelGGL((sieve_kernel<32768, 256, false>), dim3(grid), dim3(256), 0, s, a);
            break;
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_sieve_emit(const SieveArgs& a, int window_bytes, uint32_t grid,
                             hipStream_t s) {
    if (a.ndesc == 0) return hipSuccess;
    if (grid > a.ndesc) grid = a.ndesc;
    switch (window_bytes) {
        case 131072:
            hipLaunchKernelGGL((sieve_kernel<131072, 1024, true>), dim3(grid), dim3(1024), 0, s, a);
            break;
        case 65536:
            hipLaunchKernelGGL((sieve_kernel<65536, 512, true>), dim3(grid), dim3(512), 0, s, a);
            break;
        case 32768:
            hipLaunchKernelGGL((sieve_kernel<32768, 256, true>), dim3(grid), dim3(256), 0, s, a);
            break;
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

int sieve_block_threads(int window_bytes) {
    switch (window_bytes) {
        case 131072: return 1024;
        case 65536: return 512;
        case 32768: return 256;
        default: return 0;
    }
}

hipError_t launch_stitch(const SieveArgs& a, hipStream_t s) {
    if (a.ndesc == 0) return hipSuccess;
    const uint32_t blocks = (a.ndesc + 255) / 256;
    hipLaunchKernelGGL(stitch_kernel, dim3(blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace voss

namespace voss {

hipError_t launch_bucket_count(const BucketTask* tasks, uint32_t ntask, const PrimeRec* primes,
                               const uint32_t* p32, uint32_t* bucket_cnt, uint32_t* task_next,
                               int log2w, uint32_t grid, hipStream_t s) {
    if (ntask == 0) return hipSuccess;
    if (grid > ntask) grid = ntask;
    hipLaunchKernelGGL(bucket_kernel<false>, dim3(grid), dim3(1024), 0, s, tasks, ntask, primes, p32,
                       bucket_cnt, nullptr, nullptr, nullptr, task_next, log2w);
    return hipGetLastError();
}

hipError_t launch_bucket_fill(const BucketTask* tasks, uint32_t ntask, const PrimeRec* primes,
                              const uint32_t* p32, const unsigned long long* bucket_off,
                              uint32_t* bucket_cursor, uint32_t* entries, uint32_t* overflow,
                              uint32_t* task_next, int log2w, uint32_t grid, hipStream_t s) {
    if (ntask == 0) return hipSuccess;
    if (grid > ntask) grid = ntask;
    hipLaunchKernelGGL(bucket_kernel<true>, dim3(grid), dim3(1024), 0, s, tasks, ntask, primes, p32,
                       bucket_cursor, bucket_off, entries, overflow, task_next, log2w);
    return hipGetLastError();
}

hipError_t launch_add_acc(JobAcc* dst, const JobAcc* src, uint32_t njobs, hipStream_t s) {
    if (njobs == 0) return hipSuccess;
    hipLaunchKernelGGL(add_acc_kernel, dim3(njobs), dim3(256), 0, s, dst, src);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t s) {
    if (n == 0) return hipSuccess;
    size_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(fill_u32_kernel, dim3((uint32_t)blocks), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

}  // namespace voss
