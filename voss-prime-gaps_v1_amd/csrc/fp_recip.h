// fp_recip.h -- 1/p in double precision for the window/tile start remainder
// (start_mod in sieve_device.h), shared with the probe that bounds its error
// (voss_probe_recip in probe.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace voss {

// The hardware reciprocal (v_rcp_f64, relative error ~2^-24.4 on gfx950)
// refined by one Newton step.  Over every odd p in [3, 2^30) the relative
// error |1 - p*r| of the result stays below 2^-48 (voss_probe_recip, run by
// the GPU tests).
__device__ __forceinline__ double recip_newton(double pd) {
    const double r = __builtin_amdgcn_rcp(pd);
    return __builtin_fma(__builtin_fma(-pd, r, 1.0), r, r);
}

}  // namespace voss
