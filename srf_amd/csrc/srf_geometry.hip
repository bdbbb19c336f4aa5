// srf_geometry.hip -- crystal-level geometry on the GPU: LORs of crystal pairs
// (for the sensitivity image over all pairs) and listmode snapping to crystal
// centres.
//
// Crystal centres are a table (nb_crystals x 3 float32, index order) built on
// the host from the scanner description; these kernels only index it, so the
// per-pair / per-event work never leaves the GPU and no host LOR array is
// materialised.  All are HBM-bound streaming kernels.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "srf_hip.h"
#include "srf_internal.hpp"

namespace srf {

// LORs of the ordered crystal pairs p in [p0, p0 + n): first crystal p / N,
// second p % N (the order of srfnef's kron construction,
// emap_generator.py:77-81 and :94-104), six floats per LOR.
__global__ __launch_bounds__(256) void k_pair_lors(const float *__restrict__ pos, int64_t ncrys, int64_t p0,
                                                   int64_t n, float *__restrict__ lors)
{
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    const int64_t p = p0 + k;
    const int64_t a = p / ncrys, b = p - a * ncrys;
    float *o = lors + k * 6;
    o[0] = pos[a * 3 + 0];
    o[1] = pos[a * 3 + 1];
    o[2] = pos[a * 3 + 2];
    o[3] = pos[b * 3 + 0];
    o[4] = pos[b * 3 + 1];
    o[5] = pos[b * 3 + 2];
}

}  // namespace srf

using namespace srf;

extern "C" int srf_pair_lors(const float *crystal_pos, int64_t n_crystals, int64_t pair_start, int64_t n_pairs,
                             float *lors, void *stream)
{
    if (n_pairs == 0) return SRF_OK;
    if (!crystal_pos || !lors || n_crystals <= 0 || n_pairs < 0 || pair_start < 0) return SRF_ERR_INVALID_ARG;
    if (pair_start > n_crystals * n_crystals - n_pairs) return SRF_ERR_INVALID_ARG;
    hipLaunchKernelGGL(k_pair_lors, dim3((unsigned)((n_pairs + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       crystal_pos, n_crystals, pair_start, n_pairs, lors);
    return launch_status();
}
