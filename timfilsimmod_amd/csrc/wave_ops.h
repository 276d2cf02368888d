// wave_ops.h -- cross-lane helpers for 64-wide wavefronts.
#pragma once
#include <hip/hip_runtime.h>

namespace tfs {

// Lane `src` of v on every lane (src must be uniform across the wave): two
// v_readlane_b32, no LDS.
__device__ __forceinline__ double lane_bcast(double v, int src) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), src);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), src);
    return __hiloint2double(hi, lo);
}

}  // namespace tfs
