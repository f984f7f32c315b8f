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
    uint32_t restore0;      // word 0 bits of the presieved primes 7..73
    const PrimeRec* primes; // base primes >= 7, ascending
    const uint32_t* pat;                 // all tables, kPresieveWords words
    uint32_t pat_off[kPresieveGroups];   // table start in pat
    uint32_t pat_len[kPresieveGroups];   // period in words
    uint32_t pat_step[kPresieveGroups];  // block size mod period
    const unsigned long long* lg_thresh; // [kHistBins]: largest p_before
                                         // with float merit >= 10, per g/2
    SegRec* rec;
    JobAcc* acc;
    LargeRaw* large;
    unsigned long long large_cap;
    unsigned long long* large_n;
};

hipError_t upload_sieve_constants(int device);
hipError_t launch_init_acc(JobAcc* acc, uint32_t njobs, hipStream_t s);
hipError_t launch_sieve(const SieveArgs& a, int window_bytes, uint32_t grid,
                        hipStream_t s);
hipError_t launch_stitch(const SieveArgs& a, hipStream_t s);
int sieve_block_threads(int window_bytes);

}  // namespace voss
