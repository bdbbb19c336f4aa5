// srf_sss.hip -- single-scatter simulation (SSS) on a coarse virtual-detector
// sinogram, and its C-ABI entry points (include/srf_hip.h, srf_sss_*).
//
// Physics of srfnef/corrections/scattering/scatter_fraction/scatter_fraction.py:
// 70-131 and kn_formula.py:6-46 (per LOR (A,B), a Klein-Nishina weighted sum
// over scatter points S of the emission and attenuation line integrals from S
// to both detectors).  The line integrals come from the projector kernels
// (srf_project / plans) over the (detector, scatter point) LOR set; this file
// holds
//   k_sss_pairs   per (detector, scatter point): unit vector S->D, the
//                 obliquity / distance factor g = ((D-S).n_D)/|D-S|^3, and the
//                 emission and attenuation integrals packed as one 32-byte row;
//   k_sss_bins    per sinogram bin: the sum over scatter points, one bin per
//                 lane.  Bins are ordered by ring and detector A, so the lanes
//                 of a wavefront share A's row (a broadcast read) and stream B's.
//   k_sss_bins_wave  the same with one wavefront per bin (small sinograms).
// Energy-dependent factors depend only on the integer scattering angle
// (the reference quantises it to whole degrees, kn_formula.py:41), so E', the
// window efficiency and the KN factor are 181-entry tables.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "srf_hip.h"
#include "srf_internal.hpp"

namespace srf {

struct __align__(16) PairRow {
    float ux, uy, uz, g;    // unit vector S -> D, obliquity / |D-S|^2 factor
    float em, mu, pad0, pad1;
};

// det: (nd, 4) x, y, z, normal angle; sp: (ns, 4) x, y, z, mu(S);
// em, mu: (nd, ns) line integrals (row-major by detector).
__global__ __launch_bounds__(256) void k_sss_pairs(const float *__restrict__ det, const float *__restrict__ sp,
                                                   const float *__restrict__ em, const float *__restrict__ mu,
                                                   int nd, int ns, PairRow *__restrict__ rows)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (int64_t)nd * ns) return;
    const int d = (int)(i / ns), s = (int)(i % ns);
    const float vx = det[4 * d] - sp[4 * s], vy = det[4 * d + 1] - sp[4 * s + 1], vz = det[4 * d + 2] - sp[4 * s + 2];
    const float r2 = vx * vx + vy * vy + vz * vz;
    const float r = sqrtf(r2);
    const float inv = 1.0f / r;
    const float a = det[4 * d + 3];
    PairRow R;
    R.ux = vx * inv;
    R.uy = vy * inv;
    R.uz = vz * inv;
    R.g = (vx * cosf(a) + vy * sinf(a)) * inv / r2;
    R.em = em[i];
    R.mu = mu[i];
    R.pad0 = R.pad1 = 0.0f;
    rows[i] = R;
}

// tables: [0,181) 511/E', [181,362) window efficiency, [362,543) KN / sigma_KN
__global__ __launch_bounds__(256) void k_sss_bins(const PairRow *__restrict__ rows, const float *__restrict__ sp,
                                                  int nd, int ns, const int *__restrict__ bins, int nb,
                                                  const float *__restrict__ tables, float scale,
                                                  float *__restrict__ out)
{
    __shared__ float tab[543];
    for (int i = threadIdx.x; i < 543; i += 256) tab[i] = tables[i];
    __syncthreads();
    const int b = blockIdx.x * 256 + threadIdx.x;
    if (b >= nb) return;
    const int da = bins[2 * b], db = bins[2 * b + 1];
    if (da < 0 || da >= nd || db < 0 || db >= nd) {   // detector index outside the table
        out[b] = __builtin_nanf("");
        return;
    }
    const PairRow *ra = rows + (int64_t)da * ns;
    const PairRow *rb = rows + (int64_t)db * ns;
    float acc = 0.0f;
    for (int s = 0; s < ns; ++s) {
        const float m = sp[4 * s + 3];
        if (m == 0.0f) continue;   // kn term is proportional to mu(S) (kn_formula.py:16)
        const PairRow A = ra[s], B = rb[s];
        // scattering angle = pi - angle(A S B) (law of cosines, kn_formula.py:36-41)
        const float x = -(A.ux * B.ux + A.uy * B.uy + A.uz * B.uz);
        int k = (int)(acosf(fminf(1.0f, fmaxf(-1.0f, x))) * 57.295779513082323f);
        k = min(k, 180);
        const float rr = tab[k];
        const float Ia = __expf(-(A.mu + B.mu * rr)) * A.em;
        const float Ib = __expf(-(A.mu * rr + B.mu)) * B.em;
        acc += A.g * B.g * tab[181 + k] * tab[362 + k] * m * (Ia + Ib);
    }
    out[b] = acc * scale;
}

// The same sum with one wavefront per bin, for sinograms too small to fill
// the GPU one bin per lane (a few rings: a check subset).  Lane j takes the
// points s = j, j + 64, ...; the A and B rows are then read contiguously.
// The 64 partial sums are combined by a fixed butterfly, so the result is
// deterministic (it differs from the serial order of k_sss_bins only by
// float rounding).
constexpr int SSS_WAVE = 64;

__global__ __launch_bounds__(256) void k_sss_bins_wave(const PairRow *__restrict__ rows,
                                                       const float *__restrict__ sp, int nd, int ns,
                                                       const int *__restrict__ bins, int nb,
                                                       const float *__restrict__ tables, float scale,
                                                       float *__restrict__ out)
{
    __shared__ float tab[543];
    for (int i = threadIdx.x; i < 543; i += 256) tab[i] = tables[i];
    __syncthreads();
    const int b = blockIdx.x * (256 / SSS_WAVE) + (int)(threadIdx.x / SSS_WAVE);
    const int lane = (int)(threadIdx.x % SSS_WAVE);
    if (b >= nb) return;   // whole wavefronts leave together
    const int da = bins[2 * b], db = bins[2 * b + 1];
    if (da < 0 || da >= nd || db < 0 || db >= nd) {
        if (lane == 0) out[b] = __builtin_nanf("");
        return;
    }
    const PairRow *ra = rows + (int64_t)da * ns;
    const PairRow *rb = rows + (int64_t)db * ns;
    float acc = 0.0f;
    for (int s = lane; s < ns; s += SSS_WAVE) {
        const float m = sp[4 * s + 3];
        if (m == 0.0f) continue;
        const PairRow A = ra[s], B = rb[s];
        const float x = -(A.ux * B.ux + A.uy * B.uy + A.uz * B.uz);
        int k = (int)(acosf(fminf(1.0f, fmaxf(-1.0f, x))) * 57.295779513082323f);
        k = min(k, 180);
        const float rr = tab[k];
        const float Ia = __expf(-(A.mu + B.mu * rr)) * A.em;
        const float Ib = __expf(-(A.mu * rr + B.mu)) * B.em;
        acc += A.g * B.g * tab[181 + k] * tab[362 + k] * m * (Ia + Ib);
    }
#pragma unroll
    for (int off = SSS_WAVE / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, SSS_WAVE);
    if (lane == 0) out[b] = acc * scale;
}

// Sinograms below this many bins use k_sss_bins_wave.
constexpr int SSS_WAVE_BINS = 32768;

// work[0] += sum_e sino[bins[e]], work[1] += sum_e y[e] (double, one atomic per block)
__device__ __forceinline__ float bin_value(const float *__restrict__ sino, const int32_t *__restrict__ bins,
                                           int64_t e, int64_t nbins)
{
    const int32_t b = bins[e];
    return (b >= 0 && b < nbins) ? sino[b] : 0.0f;
}

__global__ __launch_bounds__(256) void k_scatter_sums(const float *__restrict__ y, const int32_t *__restrict__ bins,
                                                      const float *__restrict__ sino, int64_t nbins, int64_t n,
                                                      double *__restrict__ work)
{
    __shared__ double rs[256], ry[256];
    double a = 0.0, b = 0.0;
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256) {
        a += (double)bin_value(sino, bins, e, nbins);
        b += (double)(y ? y[e] : 1.0f);
    }
    rs[threadIdx.x] = a;
    ry[threadIdx.x] = b;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) {
            rs[threadIdx.x] += rs[threadIdx.x + k];
            ry[threadIdx.x] += ry[threadIdx.x + k];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomicAdd(work, rs[0]);
        atomicAdd(work + 1, ry[0]);
    }
}

__global__ void k_zero2(double *__restrict__ work)
{
    if (threadIdx.x < 2) work[threadIdx.x] = 0.0;
}

// out[e] = (y[e] - sino[bins[e]] / S * Y * fraction) * w[e]   (S = 0: no subtraction)
__global__ __launch_bounds__(256) void k_scatter_apply(const float *__restrict__ y, const float *__restrict__ w,
                                                       const int32_t *__restrict__ bins,
                                                       const float *__restrict__ sino, int64_t nbins, int64_t n,
                                                       float fraction,
                                                       const double *__restrict__ work, float *__restrict__ out)
{
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    const double S = work[0];
    const float c = S > 0.0 ? (float)(work[1] * (double)fraction / S) : 0.0f;
    const float ye = y ? y[e] : 1.0f;
    float v = ye - bin_value(sino, bins, e, nbins) * c;
    if (w) v *= w[e];
    out[e] = v;
}

}  // namespace srf

using namespace srf;

extern "C" {

size_t srf_sss_rows_bytes(int32_t n_det, int32_t n_points)
{
    if (n_det <= 0 || n_points <= 0) return 0;
    return (size_t)n_det * (size_t)n_points * sizeof(PairRow);
}

int srf_sss_prepare(const float *det, int32_t n_det, const float *points, int32_t n_points, const float *em_int,
                    const float *mu_int, void *rows, size_t rows_bytes, void *stream)
{
    if (n_det <= 0 || n_points <= 0) return SRF_ERR_INVALID_ARG;
    if (!det || !points || !em_int || !mu_int || !rows) return SRF_ERR_INVALID_ARG;
    if (rows_bytes < srf_sss_rows_bytes(n_det, n_points)) return SRF_ERR_WORKSPACE;
    if ((reinterpret_cast<uintptr_t>(rows) & 15) != 0) return SRF_ERR_INVALID_ARG;
    const int64_t n = (int64_t)n_det * n_points;
    hipLaunchKernelGGL(k_sss_pairs, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, det,
                       points, em_int, mu_int, n_det, n_points, (PairRow *)rows);
    return launch_status();
}

int srf_sss_sinogram(const void *rows, int32_t n_det, const float *points, int32_t n_points, const int32_t *bins,
                     int32_t n_bins, const float *tables, float scale, float *out, void *stream)
{
    if (n_bins < 0 || n_det <= 0 || n_points <= 0) return SRF_ERR_INVALID_ARG;
    if (n_bins == 0) return SRF_OK;
    if (!rows || !points || !bins || !tables || !out) return SRF_ERR_INVALID_ARG;
    if (n_bins < SSS_WAVE_BINS) {
        constexpr int per_block = 256 / SSS_WAVE;
        hipLaunchKernelGGL(k_sss_bins_wave, dim3((unsigned)((n_bins + per_block - 1) / per_block)), dim3(256), 0,
                           (hipStream_t)stream, (const PairRow *)rows, points, n_det, n_points, bins, n_bins, tables,
                           scale, out);
    } else {
        hipLaunchKernelGGL(k_sss_bins, dim3((unsigned)((n_bins + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                           (const PairRow *)rows, points, n_det, n_points, bins, n_bins, tables, scale, out);
    }
    return launch_status();
}

int srf_scatter_correct(const float *y, const float *weight, const int32_t *bins, const float *sino, int64_t n_bins,
                        int64_t n, float fraction, double *work, float *out, void *stream)
{
    if (n < 0 || n_bins < 1 || !bins || !sino || !work || !out) return SRF_ERR_INVALID_ARG;
    if (n == 0) return SRF_OK;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_zero2, dim3(1), dim3(64), 0, s, work);
    const unsigned nb = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_scatter_sums, dim3(nb < 1024 ? nb : 1024), dim3(256), 0, s, y, bins, sino, n_bins, n, work);
    hipLaunchKernelGGL(k_scatter_apply, dim3(nb), dim3(256), 0, s, y, weight, bins, sino, n_bins, n, fraction,
                       (const double *)work, out);
    return launch_status();
}

}  // extern "C"
