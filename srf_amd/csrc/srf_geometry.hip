// srf_geometry.hip -- crystal-level geometry on the GPU: LORs of crystal pairs
// (for the sensitivity image over all pairs) and listmode snapping to crystal
// centres.
//
// Crystal centres are a table (nb_crystals x 3 float32, index order) built on
// the host from the scanner description; these kernels only index it, so the
// per-pair / per-event work never leaves the GPU and no host LOR array is
// materialised.  All are HBM-bound streaming kernels.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <limits.h>
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

// Python's floor division and modulo (the reference's index arithmetic).
__device__ __forceinline__ long long pymod(long long a, long long m)
{
    const long long r = a % m;
    return r < 0 ? r + m : r;
}

// A float rounded / floored to an integer as NumPy's round (half to even) /
// floor then astype(int); non-finite or huge values give INT64_MIN.
__device__ __forceinline__ long long to_index(double v)
{
    return (v == v && fabs(v) < 4.0e18) ? (long long)v : LLONG_MIN;
}

// Crystal index of a position on an ECAT scanner, in the reference's
// precision (functions/geometry/pet_ecat_scanner.py:15-62 on float32
// positions): the block angle and the ring in float32, the in-block
// coordinates in float64.
__device__ long long ecat_index(float x, float y, float z, const srf_ecat_t &P)
{
    const float q = atan2f(y, x) / (float)P.angle_per_block;
    const long long ib0 = to_index((double)rintf(q));
    const float rz = floorf((z + (float)(P.axial_length / 2)) / (float)(P.gap + P.block_size_z));
    const long long ir = to_index((double)rz);
    if (ib0 == LLONG_MIN || ir == LLONG_MIN) return LLONG_MIN;
    const long long ib = pymod(ib0, P.nb_blocks_per_ring);
    const double ang = (double)ib * P.angle_per_block;
    const double yc = -(double)x * sin(ang) + (double)y * cos(ang) + P.block_size_y / 2;
    const long long iy = to_index(rint(yc / P.unit_y - 0.5));
    const double zc = (double)(z + (float)(P.axial_length / 2)) - (double)ir * (P.block_size_z + P.gap);
    const long long iz = to_index(rint(zc / P.unit_z - 0.5));
    if (iy == LLONG_MIN || iz == LLONG_MIN) return LLONG_MIN;
    return iy + iz * P.shape_y + ib * P.nb_crystals_per_block + ir * P.nb_crystals_per_ring;
}

// Crystal index on a cylindrical scanner (functions/geometry/
// pet_cylindrical_scanner.py:15-110): the sector angle in float32, the rest in
// float64; the module z index strided by nb_module[1], as the reference does.
__device__ long long cyl_index(float x, float y, float z, const srf_cyl_t &P)
{
    const float q = atan2f(y, x) / (float)P.angle_per_rsector;
    const long long rs0 = to_index((double)rintf(q));
    if (rs0 == LLONG_MIN) return LLONG_MIN;
    const long long rs = pymod(rs0, P.nb_rsector);
    const double ang = (double)rs * P.angle_per_rsector;
    double Y = -(double)x * sin(ang) + (double)y * cos(ang);
    double Z = (double)z;
    const long long my = to_index(floor((Y + P.sz_rsector[0] / 2) / P.mv_module[0]));
    const long long mz = to_index(floor((Z + P.sz_rsector[1] / 2) / P.mv_module[1]));
    if (my == LLONG_MIN || mz == LLONG_MIN) return LLONG_MIN;
    Y = Y - ((double)my * P.mv_module[0] - P.sz_rsector[0] / 2 + P.sz_module[0] / 2);
    Z = Z - ((double)mz * P.mv_module[1] - P.sz_rsector[1] / 2 + P.sz_module[1] / 2);
    const long long sy = to_index(floor((Y + P.sz_module[0] / 2) / P.mv_submodule[0]));
    const long long sz = to_index(floor((Z + P.sz_module[1] / 2) / P.mv_submodule[1]));
    if (sy == LLONG_MIN || sz == LLONG_MIN) return LLONG_MIN;
    Y = Y - ((double)sy * P.mv_submodule[0] - P.sz_module[0] / 2 + P.sz_submodule[0] / 2);
    Z = Z - ((double)sz * P.mv_submodule[1] - P.sz_module[1] / 2 + P.sz_submodule[1] / 2);
    const long long cy = to_index(floor((Y + P.sz_submodule[0] / 2) / P.mv_crystal[0]));
    const long long cz = to_index(floor((Z + P.sz_submodule[1] / 2) / P.mv_crystal[1]));
    if (cy == LLONG_MIN || cz == LLONG_MIN) return LLONG_MIN;
    return cy + cz * P.nb_crystal[0] + (sy + sz * P.nb_submodule[0]) * P.nb_crystal_per_submodule +
           (my + mz * P.nb_module[1]) * P.nb_crystal_per_module + rs * P.nb_crystal_per_rsector;
}

template <int KIND>
__global__ __launch_bounds__(256) void k_crystal_index(const float *__restrict__ lors, int64_t n, int64_t stride,
                                                       srf_ecat_t E, srf_cyl_t C, int64_t *__restrict__ ind)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= 2 * n) return;
    const float *p = lors + (i >> 1) * stride + (i & 1) * 3;
    ind[i] = KIND == 0 ? ecat_index(p[0], p[1], p[2], E) : cyl_index(p[0], p[1], p[2], C);
}

// LOR e with both endpoints moved to their crystals' centres (TOF and extra
// columns copied); an endpoint whose index is not a crystal makes the LOR NaN.
__global__ __launch_bounds__(256) void k_snap(const float *__restrict__ lors, int64_t n, int64_t stride,
                                              const int64_t *__restrict__ ind, const float *__restrict__ pos,
                                              int64_t ncrys, float *__restrict__ out,
                                              unsigned long long *__restrict__ outside)
{
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    const int64_t a = ind[2 * e], b = ind[2 * e + 1];
    float *o = out + e * stride;
    if (a >= 0 && a < ncrys && b >= 0 && b < ncrys) {
        o[0] = pos[a * 3];
        o[1] = pos[a * 3 + 1];
        o[2] = pos[a * 3 + 2];
        o[3] = pos[b * 3];
        o[4] = pos[b * 3 + 1];
        o[5] = pos[b * 3 + 2];
    } else {
        for (int k = 0; k < 6; ++k) o[k] = __builtin_nanf("");
        atomicAdd(outside, 1ull);
    }
    for (int64_t k = 6; k < stride; ++k) o[k] = lors[e * stride + k];
}

// Crystal-pair sinogram (listmode2sinogram.py:17-24, 76-85): the pair is
// ordered so that i >= j and lands in bin (i - 1) i / 2 + j of the
// N (N - 1) / 2 pairs i > j (get_all_lors_id's order, :27-35), weighted by
// the event value.  i == j lands in bin i (i + 1) / 2, the first pair of the
// next row, as in the reference; bins outside the table are counted instead.
__global__ __launch_bounds__(256) void k_pair_sinogram(const int64_t *__restrict__ ind, const float *__restrict__ w,
                                                       int64_t n, int64_t ncrys, double *__restrict__ sino,
                                                       unsigned long long *__restrict__ outside)
{
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    int64_t i = ind[2 * e], j = ind[2 * e + 1];
    if (i < j) {
        const int64_t t = i;
        i = j;
        j = t;
    }
    const int64_t nbins = ncrys * (ncrys - 1) / 2;
    if (j < 0 || i >= ncrys) {
        atomicAdd(outside, 1ull);
        return;
    }
    const int64_t bin = (i - 1) * i / 2 + j;   // 0 <= j <= i < ncrys
    if (bin >= nbins) {
        atomicAdd(outside, 1ull);
        return;
    }
    atomicAdd(sino + bin, w ? (double)w[e] : 1.0);
}

// Bin b of the pair table -> (i, j), i > j, b = (i - 1) i / 2 + j.
__device__ __forceinline__ void pair_of_bin(int64_t b, int64_t &i, int64_t &j)
{
    i = (int64_t)((1.0 + sqrt(1.0 + 8.0 * (double)b)) * 0.5);
    while ((i - 1) * i / 2 > b) --i;
    while (i * (i + 1) / 2 <= b) ++i;
    j = b - (i - 1) * i / 2;
}

struct SinoKeep {
    const double *sino;
    __device__ __forceinline__ bool operator()(const int64_t &b) const
    {
        const double v = sino[b];
        return v > 0.0 && v < 1e8;
    }
};

__global__ __launch_bounds__(256) void k_sino_lors(const int64_t *__restrict__ bins, int64_t count,
                                                   const double *__restrict__ sino, const float *__restrict__ pos,
                                                   float *__restrict__ lors, double *__restrict__ values)
{
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= count) return;
    const int64_t b = bins[k];
    int64_t i, j;
    pair_of_bin(b, i, j);
    float *o = lors + k * 6;
    o[0] = pos[i * 3];
    o[1] = pos[i * 3 + 1];
    o[2] = pos[i * 3 + 2];
    o[3] = pos[j * 3];
    o[4] = pos[j * 3 + 1];
    o[5] = pos[j * 3 + 2];
    values[k] = sino[b];
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

extern "C" int srf_crystal_index(const float *lors, int64_t n, int64_t lor_stride, int32_t kind,
                                 const srf_ecat_t *ecat, const srf_cyl_t *cyl, int64_t *ind, void *stream)
{
    if (n == 0) return SRF_OK;
    if (!lors || !ind || n < 0 || lor_stride < 6) return SRF_ERR_INVALID_ARG;
    srf_ecat_t E{};
    srf_cyl_t C{};
    if (kind == 0) {
        if (!ecat || ecat->nb_blocks_per_ring <= 0 || !(ecat->unit_y > 0) || !(ecat->unit_z > 0)) return SRF_ERR_INVALID_ARG;
        E = *ecat;
    } else if (kind == 1) {
        if (!cyl || cyl->nb_rsector <= 0 || !(cyl->mv_module[0] > 0) || !(cyl->mv_module[1] > 0) ||
            !(cyl->mv_submodule[0] > 0) || !(cyl->mv_submodule[1] > 0) || !(cyl->mv_crystal[0] > 0) ||
            !(cyl->mv_crystal[1] > 0))
            return SRF_ERR_INVALID_ARG;
        C = *cyl;
    } else {
        return SRF_ERR_INVALID_ARG;
    }
    const dim3 grid((unsigned)((2 * n + 255) / 256));
    if (kind == 0)
        hipLaunchKernelGGL(k_crystal_index<0>, grid, dim3(256), 0, (hipStream_t)stream, lors, n, lor_stride, E, C, ind);
    else
        hipLaunchKernelGGL(k_crystal_index<1>, grid, dim3(256), 0, (hipStream_t)stream, lors, n, lor_stride, E, C, ind);
    return launch_status();
}

extern "C" int srf_snap_lors(const float *lors, int64_t n, int64_t lor_stride, const int64_t *ind,
                             const float *crystal_pos, int64_t n_crystals, float *out, uint64_t *outside,
                             void *stream)
{
    if (n == 0) return SRF_OK;
    if (!lors || !ind || !crystal_pos || !out || !outside || n < 0 || lor_stride < 6 || n_crystals <= 0)
        return SRF_ERR_INVALID_ARG;
    hipLaunchKernelGGL(k_snap, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, lors, n,
                       lor_stride, ind, crystal_pos, n_crystals, out, (unsigned long long *)outside);
    return launch_status();
}

extern "C" int srf_pair_sinogram(const int64_t *ind, const float *values, int64_t n, int64_t n_crystals,
                                 double *sino, uint64_t *outside, void *stream)
{
    if (n == 0) return SRF_OK;
    if (!ind || !sino || !outside || n < 0 || n_crystals < 2) return SRF_ERR_INVALID_ARG;
    hipLaunchKernelGGL(k_pair_sinogram, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, ind,
                       values, n, n_crystals, sino, (unsigned long long *)outside);
    return launch_status();
}

extern "C" int srf_sinogram_lors(const double *sino, int64_t n_crystals, const float *crystal_pos, float *lors,
                                 double *values, int64_t capacity, int64_t *count, void *stream)
{
    if (!sino || !crystal_pos || !count || n_crystals < 2) return SRF_ERR_INVALID_ARG;
    if (lors && (!values || capacity < 0)) return SRF_ERR_INVALID_ARG;
    const int64_t nbins = n_crystals * (n_crystals - 1) / 2;
    if (nbins >= ((int64_t)1 << 31)) return SRF_ERR_UNSUPPORTED;
    hipStream_t s = (hipStream_t)stream;
    int64_t *bins = nullptr, *dcount = nullptr;
    void *tmp = nullptr;
    size_t tmp_bytes = 0;
    int rc = SRF_OK;
    hipcub::CountingInputIterator<int64_t> it(0);
    SinoKeep keep{sino};
    *count = 0;
    if (hipMalloc(&bins, (size_t)nbins * sizeof(int64_t)) != hipSuccess ||
        hipMalloc(&dcount, sizeof(int64_t)) != hipSuccess ||
        hipcub::DeviceSelect::If(nullptr, tmp_bytes, it, bins, dcount, (int)nbins, keep, s) != hipSuccess ||
        hipMalloc(&tmp, tmp_bytes ? tmp_bytes : 1) != hipSuccess ||
        hipcub::DeviceSelect::If(tmp, tmp_bytes, it, bins, dcount, (int)nbins, keep, s) != hipSuccess ||
        hipMemcpyAsync(count, dcount, sizeof(int64_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        (void)hipGetLastError();
        rc = SRF_ERR_LAUNCH;
    }
    if (rc == SRF_OK && lors) {
        if (*count > capacity) {
            rc = SRF_ERR_INVALID_ARG;
        } else if (*count > 0) {
            hipLaunchKernelGGL(k_sino_lors, dim3((unsigned)((*count + 255) / 256)), dim3(256), 0, s, bins, *count,
                               sino, crystal_pos, lors, values);
            if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
                (void)hipGetLastError();
                rc = SRF_ERR_LAUNCH;
            }
        }
    }
    (void)hipFree(bins);
    (void)hipFree(dcount);
    (void)hipFree(tmp);
    return rc;
}
