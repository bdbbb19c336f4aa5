// srf_internal.hpp -- helpers shared by the translation units of libsrf_hip.so.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "srf_hip.h"

namespace srf {

// Status of the launches just issued (hipGetLastError clears the sticky error
// of a failed launch, so each entry point reports its own).
inline int launch_status()
{
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? SRF_OK : SRF_ERR_LAUNCH;
}

// dst[i] = v for i < n.  A kernel, not hipMemsetAsync, so the call may be
// captured into a HIP graph and replayed.
int fill_f32(float *dst, float v, int64_t n, hipStream_t s);

struct LaunchGeom;
// Validates a grid (and TOF parameters) and derives the per-launch constants.
int make_geom(const srf_grid_t &g, const srf_tof_t *tof, LaunchGeom &G);

}  // namespace srf
