// mr_launch.h -- host-side view of the Miller-Rabin kernels (mr.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace voss {

// verdict[i] = 1 iff cand[i] is prime.  d_surv must hold 2 * count entries
// (two survivor lists); d_nsurv holds their three counters.  count must be
// below 2^32.
hipError_t launch_mr(const uint64_t* d_cand, uint64_t count, uint8_t* d_verdict,
                     uint32_t* d_surv, uint32_t* d_nsurv, int num_cus,
                     hipStream_t s);

// Largest batch of the no-allocation entry point, whose survivor lists and
// counters live in a fixed device array of mr.hip.
constexpr uint64_t kMrNoAllocMax = 1ull << 22;

constexpr int kMaxDevices = 64;

// Device address of that array on `device`, which must be the current
// device the first time it is asked for (hipGetSymbolAddress, once per
// device): d_surv = *out (2 * kMrNoAllocMax entries), d_nsurv = *out +
// 2 * kMrNoAllocMax.
hipError_t mr_scratch_address(int device, uint32_t** out);

// verify_gap's interior check for the gaps >= 2 whose ok[i] is set (both
// ends prime): ok[i] is cleared when a prime lies strictly between
// p_after[i] - gap[i] and p_after[i], or when gap[i] > p_after[i] (the
// reference's verdict on such rows).  count must be below 2^32.
hipError_t launch_verify_interior(const uint64_t* d_p_after, const int32_t* d_gap,
                                  uint8_t* d_ok, uint32_t count, hipStream_t s);

}  // namespace voss
