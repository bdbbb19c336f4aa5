// srf_elementwise.hip -- the EM loop's elementwise kernels, image layout
// conversions and the ABI's small utilities.
//
// All of these are HBM-bound streaming kernels: 16-byte vector accesses,
// grid sized to the data (no grid-stride), so a 32 MiB image is one pass.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <mutex>

#include "srf_hip.h"
#include "srf_internal.hpp"

namespace srf {

// Scratch of srf_smooth_noalloc: the intermediate image of the separable
// passes, 2^23 floats (32 MiB, a 256 x 256 x 128 image) in a file-scope device
// array, so that entry point allocates nothing.  Each device holds its own copy
// for the life of the process.  In host code `scratch` only names the symbol:
// its device address, per device, comes from hipGetSymbolAddress.
constexpr int64_t SMOOTH_SCRATCH_FLOATS = (int64_t)1 << 23;
static __device__ __align__(256) float scratch[SMOOTH_SCRATCH_FLOATS];

__global__ __launch_bounds__(256) void k_fill(float *__restrict__ dst, float v, int64_t n)
{
    const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
    if (i + 3 < n) {
        *reinterpret_cast<float4 *>(dst + i) = make_float4(v, v, v, v);
    } else {
        for (int64_t j = i; j < n; ++j) dst[j] = v;
    }
}

__global__ __launch_bounds__(256) void k_fill_scalar(float *__restrict__ dst, float v, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = v;
}

int fill_f32(float *dst, float v, int64_t n, hipStream_t s)
{
    if (n <= 0) return SRF_OK;
    if (!dst) return SRF_ERR_INVALID_ARG;
    if ((reinterpret_cast<uintptr_t>(dst) & 15) != 0)
        hipLaunchKernelGGL(k_fill_scalar, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, dst, v, n);
    else
        hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 1023) / 1024)), dim3(256), 0, s, dst, v, n);
    return launch_status();
}

// q = ratio(y, p)  (srf_ratio_kind_t)
__global__ __launch_bounds__(256) void k_ratio(const float *__restrict__ y, const float *__restrict__ p,
                                               float *__restrict__ q, int64_t n, int kind, float eps)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float d = p[i] + eps;
    float r = y[i] / d;
    if (kind == SRF_RATIO_DIV_NO_NAN && d == 0.0f) r = 0.0f;
    q[i] = r;
}

// x <- x * bp / e   (ArithmeticMixin (x*bp)/emap, arithmetic_mixins.py:38-51)
__global__ __launch_bounds__(256) void k_em_update(float *__restrict__ x, const float *__restrict__ bp,
                                                   const float *__restrict__ emap, int64_t n, float guard)
{
    const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
    if (i + 3 < n) {
        float4 xv = *reinterpret_cast<const float4 *>(x + i);
        const float4 bv = *reinterpret_cast<const float4 *>(bp + i);
        float4 ev = *reinterpret_cast<const float4 *>(emap + i);
        if (guard != 0.0f) {
            ev.x = ev.x == 0.0f ? guard : ev.x;
            ev.y = ev.y == 0.0f ? guard : ev.y;
            ev.z = ev.z == 0.0f ? guard : ev.z;
            ev.w = ev.w == 0.0f ? guard : ev.w;
        }
        xv.x = xv.x * bv.x / ev.x;
        xv.y = xv.y * bv.y / ev.y;
        xv.z = xv.z * bv.z / ev.z;
        xv.w = xv.w * bv.w / ev.w;
        *reinterpret_cast<float4 *>(x + i) = xv;
    } else {
        for (int64_t j = i; j < n; ++j) {
            float e = emap[j];
            if (guard != 0.0f && e == 0.0f) e = guard;
            x[j] = x[j] * bp[j] / e;
        }
    }
}

// 2-D tiled transpose of each y-plane: dst[c][y][r] = src[r][y][c] for a
// (R x Y x C) C-order source.  32x32 tiles staged through LDS (padded row) so
// both the read and the write are coalesced.
__global__ __launch_bounds__(256) void k_swap_outer(const float *__restrict__ src, float *__restrict__ dst, int R,
                                                    int Y, int C)
{
    __shared__ float tile[32][33];
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32, yy = blockIdx.z;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    for (int k = ty; k < 32; k += 8) {
        const int r = r0 + k, c = c0 + tx;
        if (r < R && c < C) tile[k][tx] = src[((int64_t)r * Y + yy) * C + c];
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const int c = c0 + k, r = r0 + tx;
        if (r < R && c < C) dst[((int64_t)c * Y + yy) * R + r] = tile[tx][k];
    }
}

static int swap_outer(const float *src, float *dst, int R, int Y, int C, hipStream_t s)
{
    if (!src || !dst || R <= 0 || Y <= 0 || C <= 0) return SRF_ERR_INVALID_ARG;
    if (Y > 65535) return SRF_ERR_UNSUPPORTED;
    dim3 grid((unsigned)((C + 31) / 32), (unsigned)((R + 31) / 32), (unsigned)Y);
    hipLaunchKernelGGL(k_swap_outer, grid, dim3(256), 0, s, src, dst, R, Y, C);
    return launch_status();
}

// One pass of a separable Gaussian along an axis of a C-order (n0, n1, n2)
// image: out[i] = sum_k w_k x[i + k*stride] / sum_k w_k over the taps inside
// the image (weights renormalised at the edges, so a constant stays constant),
// w_k = exp(-k^2 / (2 sigma^2)), |k| <= radius.  Optional mask: out *= mask.
__global__ __launch_bounds__(256) void k_gauss_pass(const float *__restrict__ x, int64_t n, int len, int64_t stride,
                                                    float sigma, int radius, const float *__restrict__ mask,
                                                    float *__restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int pos = (int)((i / stride) % len);
    const float c = -0.5f / (sigma * sigma);
    float s = 0.0f, ws = 0.0f;
    const int k0 = max(-radius, -pos), k1 = min(radius, len - 1 - pos);
    for (int k = k0; k <= k1; ++k) {
        const float w = expf(c * (float)(k * k));
        s += w * x[i + (int64_t)k * stride];
        ws += w;
    }
    float v = s / ws;
    if (mask) v *= mask[i];
    out[i] = v;
}

// PSF blur in the plane (corrections/psf/deconvolute.py:56-62): the sparse
// K_xy kept as each source pixel's (2 half + 1)^2 patch weights w[p][(a + half)
// (2 half + 1) + b + half] on target pixel p + (a, b); x is (nx ny, nz), slices
// fastest.  adjoint 0: out[q] = sum_o w[q - o][o] x[q - o] (K_xy x), adjoint 1:
// out[p] = sum_o w[p][o] x[p + o] (K_xy^T x); with d, out = d / (sum + eps).
// One thread per (pixel, slice): the patch weights are shared by a pixel's
// slices, the x reads coalesce along the slices.
__global__ __launch_bounds__(256) void k_psf_xy(const float *__restrict__ x, int nx, int ny, int nz,
                                                const float *__restrict__ w, int half, int adjoint,
                                                const float *__restrict__ d, float eps, float *__restrict__ out)
{
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= (int64_t)nx * ny * nz) return;
    const int z = (int)(t % nz);
    const int64_t p = t / nz;
    const int i = (int)(p / ny), j = (int)(p % ny);
    const int k = 2 * half + 1;
    float s = 0.0f;
    for (int a = -half; a <= half; ++a) {
        const int ii = adjoint ? i + a : i - a;
        if (ii < 0 || ii >= nx) continue;
        for (int b = -half; b <= half; ++b) {
            const int jj = adjoint ? j + b : j - b;
            if (jj < 0 || jj >= ny) continue;
            const int64_t q = (int64_t)ii * ny + jj;
            const int o = (a + half) * k + (b + half);
            const float wt = adjoint ? w[p * k * k + o] : w[q * k * k + o];
            s = fmaf(wt, x[q * nz + z], s);
        }
    }
    if (d) s = d[t] / (s + eps);
    out[t] = s;
}

}  // namespace srf

using namespace srf;

extern "C" {

const char *srf_strerror(int code)
{
    switch (code) {
    case SRF_OK: return "ok";
    case SRF_ERR_INVALID_ARG: return "invalid argument";
    case SRF_ERR_UNSUPPORTED: return "unsupported size or configuration";
    case SRF_ERR_LAUNCH: return "HIP launch or API call failed";
    case SRF_ERR_WORKSPACE: return "workspace too small";
    default: return "unknown error";
    }
}

int srf_version(void) { return 1; }

int srf_ratio(const float *y, const float *p, float *q, int64_t n, srf_ratio_kind_t kind, float eps, void *stream)
{
    if (n < 0) return SRF_ERR_INVALID_ARG;
    if (n == 0) return SRF_OK;
    if (!y || !p || !q) return SRF_ERR_INVALID_ARG;
    if (kind != SRF_RATIO_PLAIN && kind != SRF_RATIO_DIV_NO_NAN) return SRF_ERR_INVALID_ARG;
    hipLaunchKernelGGL(k_ratio, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, y, p, q, n,
                       (int)kind, eps);
    return launch_status();
}

int srf_em_update(float *x, const float *bp, const float *emap, int64_t nvox, float zero_guard, void *stream)
{
    if (nvox < 0) return SRF_ERR_INVALID_ARG;
    if (nvox == 0) return SRF_OK;
    if (!x || !bp || !emap) return SRF_ERR_INVALID_ARG;
    if (((reinterpret_cast<uintptr_t>(x) | reinterpret_cast<uintptr_t>(bp) | reinterpret_cast<uintptr_t>(emap)) &
         15) != 0)
        return SRF_ERR_INVALID_ARG;
    hipLaunchKernelGGL(k_em_update, dim3((unsigned)((nvox + 1023) / 1024)), dim3(256), 0, (hipStream_t)stream, x, bp,
                       emap, nvox, zero_guard);
    return launch_status();
}

int srf_smooth(const float *x, int32_t n0, int32_t n1, int32_t n2, float sigma0, float sigma1, float sigma2,
               const float *mask, float *tmp, float *out, void *stream)
{
    if (!x || !tmp || !out || n0 <= 0 || n1 <= 0 || n2 <= 0) return SRF_ERR_INVALID_ARG;
    if (!(sigma0 > 0.0f) || !(sigma1 > 0.0f) || !(sigma2 > 0.0f)) return SRF_ERR_INVALID_ARG;
    if (sigma0 > 64.0f || sigma1 > 64.0f || sigma2 > 64.0f) return SRF_ERR_UNSUPPORTED;
    if (tmp == x || tmp == out || out == x) return SRF_ERR_INVALID_ARG;
    const int64_t n = (int64_t)n0 * n1 * n2;
    const unsigned nb = (unsigned)((n + 255) / 256);
    hipStream_t s = (hipStream_t)stream;
    auto rad = [](float sg) { return (int)ceilf(3.0f * sg); };
    // axis 2 (contiguous) x -> out, axis 1 out -> tmp, axis 0 tmp -> out (masked)
    hipLaunchKernelGGL(k_gauss_pass, dim3(nb), dim3(256), 0, s, x, n, n2, (int64_t)1, sigma2, rad(sigma2),
                       (const float *)nullptr, out);
    hipLaunchKernelGGL(k_gauss_pass, dim3(nb), dim3(256), 0, s, (const float *)out, n, n1, (int64_t)n2, sigma1,
                       rad(sigma1), (const float *)nullptr, tmp);
    hipLaunchKernelGGL(k_gauss_pass, dim3(nb), dim3(256), 0, s, (const float *)tmp, n, n0, (int64_t)n1 * n2,
                       sigma0, rad(sigma0), mask, out);
    return launch_status();
}

namespace {
constexpr int SCRATCH_DEVICES = 64;
std::mutex scratch_mu;
float *scratch_dev[SCRATCH_DEVICES];   // device address of `scratch` on each device, once resolved
}  // namespace

// The current device's address of `scratch`, resolved by hipGetSymbolAddress on
// the first call on that device (with it current) and kept.
static int scratch_address(float **ptr)
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= SCRATCH_DEVICES) {
        (void)hipGetLastError();
        return SRF_ERR_LAUNCH;
    }
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_dev[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
            (void)hipGetLastError();
            return SRF_ERR_LAUNCH;
        }
        scratch_dev[dev] = static_cast<float *>(p);
    }
    *ptr = scratch_dev[dev];
    return SRF_OK;
}

int srf_smooth_noalloc(const float *x, int32_t n0, int32_t n1, int32_t n2, float sigma0, float sigma1, float sigma2,
                       const float *mask, float *out, void *stream)
{
    if (!x || !out || n0 <= 0 || n1 <= 0 || n2 <= 0) return SRF_ERR_INVALID_ARG;
    if ((int64_t)n0 * n1 * n2 > SMOOTH_SCRATCH_FLOATS) return SRF_ERR_WORKSPACE;
    float *tmp = nullptr;
    const int rc = scratch_address(&tmp);
    if (rc) return rc;
    // every scratch word a call reads is written by the same call first (the
    // second pass fills it whole before the third reads it): nothing to clear
    return srf_smooth(x, n0, n1, n2, sigma0, sigma1, sigma2, mask, tmp, out, stream);
}

int srf_image_to_kernel_layout(const float *src_xyz, float *dst_kernel, int32_t nx, int32_t ny, int32_t nz,
                               void *stream)
{
    // (nx, ny, nz) z-fastest  ->  (nz, ny, nx) x-fastest
    return swap_outer(src_xyz, dst_kernel, nx, ny, nz, (hipStream_t)stream);
}

int srf_image_from_kernel_layout(const float *src_kernel, float *dst_xyz, int32_t nx, int32_t ny, int32_t nz,
                                 void *stream)
{
    return swap_outer(src_kernel, dst_xyz, nz, ny, nx, (hipStream_t)stream);
}

int srf_psf_xy(const float *x, int32_t nx, int32_t ny, int32_t nz, const float *weights, int32_t half,
               int32_t adjoint, const float *d, float eps, float *out, void *stream)
{
    if (nx <= 0 || ny <= 0 || nz <= 0) return SRF_ERR_INVALID_ARG;
    if (!x || !weights || !out || half < 0 || half > 32 || (adjoint != 0 && adjoint != 1)) return SRF_ERR_INVALID_ARG;
    if (x == out || (d && d == out)) return SRF_ERR_INVALID_ARG;
    const int64_t n = (int64_t)nx * ny * nz;
    hipLaunchKernelGGL(k_psf_xy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, x, nx, ny, nz,
                       weights, half, adjoint, d, eps, out);
    return launch_status();
}

}  // extern "C"
