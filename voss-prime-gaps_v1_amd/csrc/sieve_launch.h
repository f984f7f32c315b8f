// sieve_launch.h -- host-side view of the sieve kernels (sieve.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "wheel.h"

namespace voss {

struct SieveArgs {
    const SegDesc* desc;
    uint32_t ndesc;
    uint32_t tier_a_begin;  // first prime index not covered by presieve
    uint32_t tier_a_end;    // first prime index marked one-lane-per-prime
    uint32_t ib_frac[3];    // first prime indexes >= window_bytes/4, /3, /2
                            // (tier B2's ranges in a full window)
    uint32_t restore0;      // word 0 bits of the presieved primes 7..37
    const PrimeRec* primes; // base primes >= 7, ascending
    const uint32_t* p32;    // the same primes as 32-bit values
    const uint32_t* pat;                 // all tables, kPresieveWordsPadded words
    uint32_t pat_off[kPresieveGroups];   // table start in pat
    uint32_t pat_len[kPresieveGroups];   // period in words
    uint32_t pat_step[kPresieveGroups];  // block size mod period
    const unsigned long long* lg_thresh; // [kHistBins]: largest p_before
                                         // with float merit >= 10, per g/2
    const unsigned long long* bucket_off;  // first entry of each bucket, and
                                           // one more: the end of the last
    const uint32_t* bucket_cnt;            // entries of each bucket (values
                                           // beyond the bucket are clamped)
    const uint32_t* bucket_ent;            // (offset << 3) | bit
    unsigned long long* emit_out;          // EMIT mode: primes, in order
    const unsigned long long* emit_off;    // EMIT mode: first slot per window
    SegRec* rec;
    JobAcc* acc;
    LargeRaw* large;
    unsigned long long large_cap;
    unsigned long long* large_n;
    uint32_t* win_next;  // window counter, 0 at launch: a workgroup takes
                         // window blockIdx.x, then gridDim.x + atomicAdd
};

hipError_t launch_init_acc(JobAcc* acc, uint32_t njobs, hipStream_t s);
hipError_t launch_sieve(const SieveArgs& a, int window_bytes, uint32_t grid,
                        hipStream_t s);
hipError_t launch_stitch(const SieveArgs& a, hipStream_t s);
// The sieve in EMIT mode: every counted prime of window d is written, in
// increasing order, to emit_out[emit_off[d] + i] (no histograms).
hipError_t launch_sieve_emit(const SieveArgs& a, int window_bytes, uint32_t grid,
                             hipStream_t s);

// Bucket passes over `ntask` tasks (one workgroup owns a task's tile; tasks
// are taken from *task_next, which must be 0 at launch).  count:
// bucket_cnt[b] = hits of window b.  fill: writes entries using bucket_off
// (nbuckets + 1 offsets) and sets bucket_cursor[b] to the hits of window b;
// a bucket that would exceed its room sets *overflow instead.
hipError_t launch_bucket_count(const BucketTask* tasks, uint32_t ntask,
                               const PrimeRec* primes, const uint32_t* p32, uint32_t* bucket_cnt,
                               uint32_t* task_next, int log2w, uint32_t grid,
                               hipStream_t s);
hipError_t launch_bucket_fill(const BucketTask* tasks, uint32_t ntask,
                              const PrimeRec* primes, const uint32_t* p32,
                              const unsigned long long* bucket_off,
                              uint32_t* bucket_cursor, uint32_t* entries,
                              uint32_t* overflow, uint32_t* task_next, int log2w,
                              uint32_t grid, hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t s);
// dst[j] += src[j] for j < njobs (sums; min first, max last, or error).
hipError_t launch_add_acc(JobAcc* dst, const JobAcc* src, uint32_t njobs, hipStream_t s);
int sieve_block_threads(int window_bytes);

}  // namespace voss
