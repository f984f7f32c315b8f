// serve_kernel.hip -- the device-resident serving loop.
//
// One persistent workgroup per CU (1024 threads, the sieve's 128 KiB LDS
// window).  Each workgroup claims the next unit of the ring (serve_launch.h)
// and runs it:
//   window  one LDS window of an interval request (sieve_window, the same
//           code as the batch kernels), after its tile's bucket tasks have
//           completed; its sums go to the request slot's HBM accumulators;
//           the workgroup that completes the request's last window joins
//           the pairs across its windows (the stitch), copies the sums into
//           the request's mailbox and clears the slot;
//   bucket  one bucket-fill task of an interval request (bucket_task);
//   mr      a chunk of 8192 candidates of a Miller-Rabin request: trial
//           division, the base-2 test on what it leaves, the other 11
//           bases on what passes (each stage over an LDS list, so the waves
//           stay coherent), verdicts written to the request's host buffer;
//           the workgroup that completes the last chunk publishes it.
// A workgroup exits after idle_limit empty polls of the ring (about 30 us
// apart) or when the host sets ctl->stop; no workgroup waits for another
// except a window for its tile's bucket tasks, which precede it in the ring
// and so are held by running workgroups (that wait gives up after
// dep_limit polls and marks the request failed).
#include <hip/hip_runtime.h>

#include "mr_device.h"
#include "serve_launch.h"
#include "sieve_device.h"

namespace voss {

constexpr unsigned long long kNoUnit = ~0ull;
constexpr int kServeWin = 131072;
constexpr int kServeBlock = 1024;
using ServeLds = SieveShared<kServeWin, kServeBlock, false>;

// The claimed unit and the host records it refers to, copied into LDS.
struct ServeExtra {
    unsigned long long unit_idx;
    ServeUnit unit;
    ServeReqInfo req;
    union {
        SegDesc desc;
        BucketTask task;
        ServeMrInfo mr;
    };
    unsigned long long bent_off;
    uint32_t nbent;
    uint32_t has_bucket;
    uint32_t flag;
    uint32_t n1, n2;
    unsigned long long t0;        // wall clock at the claim
    unsigned long long st[6];     // units run and ticks spent, per kind
};

// Host memory that the host writes while the loop runs is read with
// system-scope loads: they bypass the GPU caches, which would otherwise keep
// serving a line read before the host wrote it.
template <typename T>
__device__ __forceinline__ T host_ld(const T* p) {
    return __hip_atomic_load(const_cast<T*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Copies n dwords (n even, src 8-byte aligned) from src (host memory) to
// LDS in 8-byte loads; the first wave calls it.
__device__ __forceinline__ void fetch_dwords(void* dst, const void* src, int n) {
    const int lane = threadIdx.x & 63;
    if (2 * lane < n)
        static_cast<unsigned long long*>(dst)[lane] =
            host_ld(static_cast<const unsigned long long*>(src) + lane);
}

// Thread 0: ring entry u (three 8-byte loads).
__device__ __forceinline__ ServeUnit read_unit(const ServeArgs& sv, unsigned long long u) {
    const unsigned long long* q =
        reinterpret_cast<const unsigned long long*>(sv.ring + (u & sv.ring_mask));
    static_assert(sizeof(ServeUnit) == 24, "ring entry");
    ServeUnit e;
    unsigned long long* d = reinterpret_cast<unsigned long long*>(&e);
    d[0] = host_ld(q);
    d[1] = host_ld(q + 1);
    d[2] = host_ld(q + 2);
    return e;
}

// Thread 0: the next unit's index, or kNoUnit when the workgroup should exit.
// Workgroups claim against a device copy of the host's tail; when it shows
// no unit, one workgroup at a time (a ticket in device memory) reads the
// host's tail and stop flag over the bus and updates the copies.  Idle
// workgroups therefore poll device memory, not the bus.
__device__ __forceinline__ unsigned long long serve_claim(const ServeArgs& sv) {
    unsigned long long* const claim = sv.claim;
    unsigned long long* const tail_seen = sv.claim + 16;
    unsigned long long* const stop_seen = sv.claim + 17;
    unsigned long long* const ticket = sv.claim + 32;
    auto dld = [](unsigned long long* p) {
        return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    };
    uint32_t idle = 0;
    for (;;) {
        unsigned long long c = dld(claim);
        const unsigned long long stop = dld(stop_seen);
        if (stop >= 2) return kNoUnit;
        if (c < dld(tail_seen)) {
            // a failed exchange means another workgroup took unit c
            if (__hip_atomic_compare_exchange_strong(claim, &c, c + 1, __ATOMIC_RELAXED,
                                                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                __hip_atomic_store(&sv.claimed_by[blockIdx.x], c + 1, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_SYSTEM);
                return c;
            }
            continue;
        }
        unsigned long long free_ticket = 0;
        if ((idle & 3u) == 0u &&
            __hip_atomic_compare_exchange_strong(ticket, &free_ticket, 1ull, __ATOMIC_RELAXED,
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
            // relaxed: the ring and the records it points to are uncached
            // host memory (system-scope loads), so no cache needs
            // invalidating to see them; a ring entry's sequence stamp is
            // checked when it is read
            const unsigned long long t =
                __hip_atomic_load(&sv.ctl->tail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            const unsigned long long st =
                __hip_atomic_load(&sv.ctl->stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            __hip_atomic_fetch_max(tail_seen, t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(stop_seen, st, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(ticket, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (c < t) continue;
            if (st == 1) return kNoUnit;
        }
        if (++idle > sv.idle_limit) return kNoUnit;
        __builtin_amdgcn_s_sleep(16);
    }
}

// The pairs across the request's windows (each window's first prime with
// the last prime of the previous non-empty window: every pair of an
// interval counts in full), the sums into the mailbox, the slot cleared,
// the generation published.
__device__ void serve_finish(const SieveArgs& a, const ServeArgs& sv, ServeLds& S, ServeExtra& X,
                             uint32_t slot) {
    const int tid = threadIdx.x;
    const ServeReqInfo& R = X.req;
    JobAcc* acc = &sv.acc[slot];
    ServeSlot* st = &sv.slot[slot];
    // the windows' records and sums come from other workgroups: coherent
    // loads (no L2 invalidation)
    auto ld = [](const unsigned long long* p) {
        return __hip_atomic_load(const_cast<unsigned long long*>(p), __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_AGENT);
    };
    for (uint32_t i = tid; i < R.nwin; i += kServeBlock) {
        if (ld(&R.rec[i].count) == 0) continue;
        int e = (int)i - 1;
        while (e >= 0 && ld(&R.rec[e].count) == 0) e--;
        if (e < 0) continue;
        const unsigned long long p = ld(&R.rec[e].last);
        const unsigned long long g64 = ld(&R.rec[i].first) - p;
        if (g64 >= 2ull * kHistBins) {
            atomicOr(&acc->error, 1ull);
            continue;
        }
        const uint32_t g = (uint32_t)g64;
        if (g < 100u) {
            // the flush adds the q=30 row's count to the q=6 row and the
            // main histogram
            atomicAdd(&S.mq[((g >> 1) << 3) + (uint32_t)c_residx[p % 30ull]], 1u);
        } else {
            atomicAdd(&S.hist[g >> 1], 1u);
            if (g > 100u && (g >= (uint32_t)kLargeAlways || p <= a.lg_thresh[g >> 1])) {
                const unsigned long long k = atomicAdd(&st->large_n, 1ull);
                atomicAdd(&acc->large_count, 1ull);
                if (k < R.lcap) {
                    LargeRaw r;
                    r.p_before = p;
                    r.gap = g;
                    r.job = slot;
                    R.large[k] = r;
                }
            }
        }
    }
    sieve_flush(S, acc);
    // this workgroup's atomics complete, then read the sums of all of them
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    ServeMailbox* mb = &sv.mail[slot];
    {
        const unsigned long long* src = reinterpret_cast<const unsigned long long*>(acc);
        unsigned long long* dst = reinterpret_cast<unsigned long long*>(&mb->acc);
        for (uint32_t k = tid; k < sizeof(JobAcc) / 8; k += kServeBlock) dst[k] = ld(&src[k]);
    }
    if (tid == 0) {
        mb->large_n = __hip_atomic_load(&st->large_n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        mb->error = __hip_atomic_load(&st->error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        mb->overflow = __hip_atomic_load(&st->overflow, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    // the slot's device state back to its free state
    {
        unsigned long long* w = reinterpret_cast<unsigned long long*>(acc);
        for (uint32_t k = tid; k < sizeof(JobAcc) / 8; k += kServeBlock) w[k] = 0;
        for (uint32_t k = tid; k < R.nbkt; k += kServeBlock) R.cursor[k] = 0;
        for (uint32_t k = tid; k < R.ntile; k += kServeBlock) R.tile_done[k] = 0;
    }
    __syncthreads();
    if (tid == 0) {
        acc->first = ~0ull;
        __hip_atomic_store(&st->large_n, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&st->windows_done, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&st->error, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&st->overflow, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
    __syncthreads();
    if (tid == 0)
        __hip_atomic_store(&mb->done_gen, (unsigned long long)R.gen, __ATOMIC_RELEASE,
                           __HIP_MEMORY_SCOPE_SYSTEM);
}

__device__ void serve_window(const SieveArgs& a, const ServeArgs& sv, ServeLds& S, ServeExtra& X,
                             uint32_t slot, uint32_t index) {
    const int tid = threadIdx.x;
    if (tid < 64) fetch_dwords(&X.req, &sv.req[slot], sizeof(ServeReqInfo) / 4);
    __syncthreads();
    if (tid < 64) fetch_dwords(&X.desc, X.req.desc + index, sizeof(SegDesc) / 4);
    __syncthreads();
    SegDesc D;
    {
        uint32_t* dd = reinterpret_cast<uint32_t*>(&D);
        const uint32_t* sd = reinterpret_cast<const uint32_t*>(&X.desc);
#pragma unroll
        for (int k = 0; k < (int)(sizeof(SegDesc) / 4); k++)
            dd[k] = __builtin_amdgcn_readfirstlane(sd[k]);
    }
    if (tid == 0) {
        X.has_bucket = 0;
        X.nbent = 0;
        X.bent_off = 0;
        if (D.bucket != kNoBucket) {
            // the tile's bucket tasks were claimed before this window, by
            // workgroups that are running them
            const uint32_t tile = D.bucket >> 8;
            const uint32_t need = host_ld(&X.req.tile_ntask[tile]);
            uint32_t spins = 0;
            while (__hip_atomic_load(&X.req.tile_done[tile], __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_AGENT) < need) {
                if (++spins > sv.dep_limit) {
                    atomicOr(&sv.slot[slot].error, 1u);
                    break;
                }
                __builtin_amdgcn_s_sleep(8);
            }
            const unsigned long long e0 = host_ld(&X.req.boff[D.bucket]);
            const unsigned long long e1 = host_ld(&X.req.boff[D.bucket + 1]);
            const uint32_t c = __hip_atomic_load(&X.req.cursor[D.bucket], __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT);
            X.bent_off = e0;
            X.nbent = (uint32_t)min((unsigned long long)c, e1 - e0);
            X.has_bucket = 1;
        }
    }
    sieve_set_patbase(a, S, D.w0, tid);
    __syncthreads();
    const bool has_bucket = X.has_bucket != 0;
    WinIO io;
    io.acc = &sv.acc[slot];
    io.large = X.req.large;
    io.large_n = &sv.slot[slot].large_n;
    io.large_cap = X.req.lcap;
    io.job = slot;
    io.bent = has_bucket ? sv.entries + X.bent_off : nullptr;
    io.nbent = X.nbent;
    io.rec = X.req.rec + index;
    io.emit_at = 0;
    // bucket entries written by other workgroups: read with coherent loads
    // (no L2 invalidation, which would evict the prime tables of the CUs
    // beside this one)
    sieve_window<kServeWin, kServeBlock, false, true>(a, S, D, io, [] {});
    sieve_flush(S, io.acc);
    // the window's record, sums and large-gap records before its count
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
    __syncthreads();
    if (tid == 0) {
        const uint32_t done = atomicAdd(&sv.slot[slot].windows_done, 1u);
        X.flag = done + 1u == X.req.nwin ? 1u : 0u;
    }
    __syncthreads();
    if (X.flag) serve_finish(a, sv, S, X, slot);
}

__device__ void serve_bucket(const SieveArgs& a, const ServeArgs& sv, ServeLds& S, ServeExtra& X,
                             uint32_t slot, uint32_t index) {
    const int tid = threadIdx.x;
    if (tid < 64) fetch_dwords(&X.req, &sv.req[slot], sizeof(ServeReqInfo) / 4);
    __syncthreads();
    if (tid < 64) fetch_dwords(&X.task, X.req.tasks + index, sizeof(BucketTask) / 4);
    __syncthreads();
    const BucketTask T = X.task;
    uint32_t* cnt = S.bm;
    uint32_t* lim = S.bm + kBucketTileMax;
    // one workgroup per CU here (the sieve's LDS): eight records in flight
    // per thread make up for the occupancy of the batch form's fill kernel
    bucket_task<true, true, 8>(T, a.primes, a.p32, X.req.cursor, X.req.boff, sv.entries, &sv.slot[slot].overflow,
                      sv.log2w, cnt, lim);
    // the entries and cursors before the tile's count; the tile's last task
    // tells the host, which publishes the tile's windows after that
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (tid == 0) {
        const uint32_t tile = T.b0 >> 8;
        const uint32_t done = atomicAdd(&X.req.tile_done[tile], 1u);
        if (done + 1u == host_ld(&X.req.tile_ntask[tile]))
            __hip_atomic_store(&X.req.tile_ready[tile], 1u, __ATOMIC_RELEASE,
                               __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

__device__ void serve_mr(const ServeArgs& sv, ServeLds& S, ServeExtra& X, uint32_t slot,
                         uint32_t index) {
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    if (tid < 64) fetch_dwords(&X.mr, &sv.mrq[slot], sizeof(ServeMrInfo) / 4);
    if (tid == 0) {
        X.n1 = 0;
        X.n2 = 0;
    }
    __syncthreads();
    const ServeMrInfo M = X.mr;
    const unsigned long long base = (unsigned long long)index * kServeMrChunk;
    const uint32_t n = (uint32_t)min((unsigned long long)kServeMrChunk, M.count - base);
    const uint64_t* cand = M.cand + base;
    // LDS (the window bitmap's space): the powering table of the other
    // bases, the two lists, the chunk's verdict bytes
    uint64_t* tbl = reinterpret_cast<uint64_t*>(S.bm);                     // 64 KiB
    uint16_t* list1 = reinterpret_cast<uint16_t*>(S.bm + 16384);          // 16 KiB
    uint16_t* list2 = reinterpret_cast<uint16_t*>(S.bm + 16384 + 4096);   // 16 KiB
    uint8_t* vb = reinterpret_cast<uint8_t*>(S.bm + 16384 + 8192);        // 8 KiB
    // trial division by the 12 bases decides most candidates
    for (uint32_t k = tid; k < n; k += kServeBlock) {
        const int v = trial(host_ld(&cand[k]));
        vb[k] = (uint8_t)(v == 1);
        if (v == 2) list1[atomicAdd(&X.n1, 1u)] = (uint16_t)k;
    }
    __syncthreads();
    // the base-2 strong test on the rest
    const uint32_t n1 = X.n1;
    for (uint32_t q = tid; q < n1; q += kServeBlock) {
        const uint32_t k = list1[q];
        const uint64_t x = host_ld(&cand[k]);
        const Mont Mt = mont_setup(x);
        uint64_t d = x - 1;
        const int s = __ffsll((long long)d) - 1;
        d >>= s;
        if (sprp2(Mt, d, s)) list2[atomicAdd(&X.n2, 1u)] = (uint16_t)k;
    }
    __syncthreads();
    // bases 3..37 on the base-2 probable primes (almost all prime, so the
    // lanes of a wave run the same steps); the loop bound is uniform, so
    // the whole wave is present for the digit-count reduction
    const uint32_t n2 = X.n2;
    for (uint32_t q0 = 0; q0 < n2; q0 += kServeBlock) {
        const uint32_t q = q0 + tid;
        const bool active = q < n2;
        uint32_t k = 0;
        uint64_t x = 3, d = 1;
        int s = 1;
        if (active) {
            k = list2[q];
            x = host_ld(&cand[k]);
            d = x - 1;
            s = __ffsll((long long)d) - 1;
            d >>= s;
        }
        int ktop = active ? (63 - __clzll((long long)d)) / 3 : 0;
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) ktop = max(ktop, __shfl_xor(ktop, o));
        ktop = min(ktop, 20);
        (void)lane;
        if (active) {
            const Mont Mt = mont_setup(x);
            uint64_t bm = add_mod(Mt.one, Mt.one, x);  // base 2 in Montgomery form
            uint32_t b = 2;
            bool prime = true;
            for (int t = 0; t < 11 && prime; t++) {
                for (const uint32_t nb = c_bases[t]; b < nb; b++) bm = add_mod(bm, Mt.one, x);
                prime = sprp_window<kServeBlock>(Mt, bm, d, s, ktop, tbl);
            }
            vb[k] = prime ? 1 : 0;
        }
    }
    __syncthreads();
    // verdicts out in 4-byte words (base is a multiple of 8192; the host
    // buffer is rounded up to 4 bytes)
    {
        uint32_t* dst = reinterpret_cast<uint32_t*>(M.verdict + base);
        const uint32_t* src = reinterpret_cast<const uint32_t*>(vb);
        for (uint32_t w = tid; w < (n + 3u) / 4u; w += kServeBlock) dst[w] = src[w];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
    __syncthreads();
    if (tid == 0) {
        const uint32_t done = atomicAdd(&sv.mr_done[slot], 1u);
        if (done + 1u == M.nchunk) {
            __hip_atomic_store(&sv.mr_done[slot], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&sv.mr_mail[slot], (unsigned long long)M.gen, __ATOMIC_RELEASE,
                               __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
}

__global__ __launch_bounds__(kServeBlock) void serve_kernel(SieveArgs a, ServeArgs sv) {
    __shared__ ServeLds S;
    __shared__ ServeExtra X;
    const int tid = threadIdx.x;
    sieve_init_lds(a, S);
    if (tid < 6) X.st[tid] = 0;
    for (;;) {
        if (tid == 0) {
            const unsigned long long u = serve_claim(sv);
            X.unit_idx = u;
            if (u != kNoUnit) {
                // the host stamps an entry before publishing it; an older
                // entry seen here is read again (bounded: a stale entry
                // marks the unit failed)
                ServeUnit e = read_unit(sv, u);
                for (uint32_t k = 0; k < 1024u && e.seq != u; k++) {
                    __builtin_amdgcn_s_sleep(4);
                    e = read_unit(sv, u);
                }
                if (e.seq != u) {
                    e.kind = 0;
                    __hip_atomic_store(&sv.ctl->fault, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                }
                X.unit = e;
            }
            X.t0 = __builtin_amdgcn_s_memrealtime();
        }
        __syncthreads();
        if (X.unit_idx == kNoUnit) break;  // uniform: every thread read the same value
        const ServeUnit unit = X.unit;
        if (unit.kind == kUnitWindow) serve_window(a, sv, S, X, unit.slot, unit.index);
        else if (unit.kind == kUnitBucket) serve_bucket(a, sv, S, X, unit.slot, unit.index);
        else if (unit.kind == kUnitMr) serve_mr(sv, S, X, unit.slot, unit.index);
        if (tid == 0) {
            atomicAdd(sv.units_done, 1ull);
            // this workgroup's units and wall-clock ticks of this kind
            const uint32_t k = unit.kind - 1u;
            if (k < 3u) {
                const unsigned long long n = ++X.st[2u * k];
                const unsigned long long t = X.st[2u * k + 1u] += __builtin_amdgcn_s_memrealtime() - X.t0;
                unsigned long long* w = sv.wg_stats + 8ull * blockIdx.x + 2u * k;
                __hip_atomic_store(&w[0], n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                __hip_atomic_store(&w[1], t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            }
        }
        __syncthreads();  // X is rewritten by the next claim
    }
}

hipError_t launch_serve(const SieveArgs& a, const ServeArgs& sv, uint32_t grid, hipStream_t s) {
    hipLaunchKernelGGL(serve_kernel, dim3(grid), dim3(kServeBlock), 0, s, a, sv);
    return hipGetLastError();
}

}  // namespace voss
