// srf_project.hip -- projector / backprojector kernels and their C-ABI entry
// points (declared in include/srf_hip.h).
//
// Kernels (one event per lane, 256-lane workgroups, N/256 workgroups):
//   k_gather   forward projection            (Projection[Tof] op)
//   k_scatter  backprojection                (Backprojection[Tof] op / adjoint)
//   k_digest   per-event term digest         (verification of voxel indexing)
// All three instantiate the same srf::march_event, so the digest proves the
// indexing of the projector and backprojector themselves.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "marcher.hpp"
#include "srf_hip.h"
#include "srf_internal.hpp"

namespace srf {

struct GatherAct {
    const float *__restrict__ img;
    float acc;
    __device__ __forceinline__ void term(int v, int, float w) { acc += img[v] * w; }
    __device__ __forceinline__ void nan_event() { acc = __builtin_nanf(""); }
};

struct ScatterAct {
    float *__restrict__ img;
    float value;
    __device__ __forceinline__ void term(int v, int, float w) { unsafeAtomicAdd(img + v, value * w); }
    __device__ __forceinline__ void nan_event() {}
};

struct DigestAct {
    uint64_t h;
    int cnt;
    __device__ __forceinline__ void term(int v, int br, float) {
        h += mix64(((uint64_t)(uint32_t)v << 4) | (uint64_t)br);
        cnt += 1;
    }
    __device__ __forceinline__ void nan_event() {}
};

__device__ __forceinline__ float event_ratio(const float *__restrict__ y, const float *__restrict__ p,
                                             int64_t e, int kind, float eps)
{
    const float yy = y[e];
    const float d = p[e] + eps;
    if (kind == SRF_RATIO_DIV_NO_NAN) return d == 0.0f ? 0.0f : yy / d;
    return yy / d;
}

template <bool TOF, int VAR, bool DS, bool DZ>
__global__ __launch_bounds__(256) void k_gather(const float *__restrict__ lors, int64_t n, int stride,
                                                const float *__restrict__ tofv, int tof_stride,
                                                const float *__restrict__ img, LaunchGeom G,
                                                float *__restrict__ out)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    GatherAct act{img, 0.0f};
    march_event<TOF, VAR, DS, DZ>(lors + e * stride, TOF ? tofv[e * tof_stride] : 0.0f, G, act);
    out[e] = act.acc;
}

// values == nullptr: all ones.  y/p != nullptr: value = ratio(y, p).
template <bool TOF, int VAR, bool DS, bool DZ>
__global__ __launch_bounds__(256) void k_scatter(const float *__restrict__ lors, int64_t n, int stride,
                                                 const float *__restrict__ tofv, int tof_stride,
                                                 const float *__restrict__ values,
                                                 const float *__restrict__ y, const float *__restrict__ p,
                                                 int kind, float eps, LaunchGeom G, float *__restrict__ img)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    float v = 1.0f;
    if (y) v = event_ratio(y, p, e, kind, eps);
    else if (values) v = values[e];
    ScatterAct act{img, v};
    march_event<TOF, VAR, DS, DZ>(lors + e * stride, TOF ? tofv[e * tof_stride] : 0.0f, G, act);
}

template <bool TOF, int VAR, bool DS, bool DZ>
__global__ __launch_bounds__(256) void k_digest(const float *__restrict__ lors, int64_t n, int stride,
                                                const float *__restrict__ tofv, int tof_stride, LaunchGeom G,
                                                uint64_t *__restrict__ hash, int32_t *__restrict__ count)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    DigestAct act{0ull, 0};
    march_event<TOF, VAR, DS, DZ>(lors + e * stride, TOF ? tofv[e * tof_stride] : 0.0f, G, act);
    hash[e] = act.h;
    count[e] = act.cnt;
}

// ----------------------------------------------------------------- host side --

int make_geom(const srf_grid_t &g, const srf_tof_t *tof, LaunchGeom &G)
{
    if (g.nx <= 0 || g.ny <= 0 || g.nz <= 0) return SRF_ERR_INVALID_ARG;
    if ((int64_t)g.nx * g.ny * g.nz >= ((int64_t)1 << 31)) return SRF_ERR_UNSUPPORTED;
    if (!(g.sx > 0.0f && g.sy > 0.0f && g.sz > 0.0f)) return SRF_ERR_INVALID_ARG;
    if (!(isfinite(g.cx) && isfinite(g.cy) && isfinite(g.cz))) return SRF_ERR_INVALID_ARG;
    G.nx = g.nx; G.ny = g.ny; G.nz = g.nz;
    G.cx = g.cx; G.cy = g.cy; G.cz = g.cz;
    // siddon.cu.cc:19-22, 39-41 -- same float expressions, host IEEE
    G.dx_ = g.sx / (float)g.nx;
    G.dy = g.sy / (float)g.ny / G.dx_;
    G.dz = g.sz / (float)g.nz / G.dx_;
    G.nx2 = g.nx / 2.0f;
    G.ny2 = g.ny / 2.0f;
    G.nz2 = g.nz / 2.0f;
    G.tof_bin_ = G.tof_sigma2_ = G.tse = G.tof_c0 = G.tof_c1 = 0.0f;
    if (tof) {
        if (!(tof->tof_sigma2 > 0.0f) || !(tof->tof_bin >= 0.0f)) return SRF_ERR_INVALID_ARG;
        // siddon_tof.cu.cc:46-47, 57, 59
        G.tof_bin_ = tof->tof_bin / G.dx_;
        G.tof_sigma2_ = tof->tof_sigma2 / G.dx_ / G.dx_;
        G.tse = G.tof_sigma2_ + (G.tof_bin_ * G.tof_bin_) / 12;
        const float PI_REF = 3.1415926f;  // siddon_tof.cu.cc:10
        G.tof_c0 = (float)((double)G.tof_bin_ / sqrt(2.0 * (double)PI_REF * (double)G.tse));
        G.tof_c1 = (float)(-0.5 * 1.4426950408889634 / (double)G.tse);
    }
    return SRF_OK;
}

static int check_events(const float *lors, int64_t n, int64_t stride, const float *tofv, int64_t tof_stride,
                        const srf_tof_t *tof)
{
    if (n < 0) return SRF_ERR_INVALID_ARG;
    if (n == 0) return SRF_OK;
    if (!lors || stride < 6 || stride > (1 << 20)) return SRF_ERR_INVALID_ARG;
    if (tof && (!tofv || tof_stride < 1 || tof_stride > (1 << 20))) return SRF_ERR_INVALID_ARG;
    if (n > ((int64_t)1 << 40)) return SRF_ERR_UNSUPPORTED;
    return SRF_OK;
}

#define SRF_DISPATCH(KERNEL, TOFV, VARV, DSV, DZV, ...)                                                     \
    do {                                                                                                   \
        if (TOFV) {                                                                                        \
            if (VARV == V_FP) {                                                                            \
                if (DSV) { if (DZV) KERNEL<true, V_FP, true, true> __VA_ARGS__; else KERNEL<true, V_FP, true, false> __VA_ARGS__; } \
                else { if (DZV) KERNEL<true, V_FP, false, true> __VA_ARGS__; else KERNEL<true, V_FP, false, false> __VA_ARGS__; } \
            } else {                                                                                       \
                if (DSV) { if (DZV) KERNEL<true, V_BP, true, true> __VA_ARGS__; else KERNEL<true, V_BP, true, false> __VA_ARGS__; } \
                else { if (DZV) KERNEL<true, V_BP, false, true> __VA_ARGS__; else KERNEL<true, V_BP, false, false> __VA_ARGS__; } \
            }                                                                                              \
        } else {                                                                                           \
            if (VARV == V_FP) {                                                                            \
                if (DSV) { if (DZV) KERNEL<false, V_FP, true, true> __VA_ARGS__; else KERNEL<false, V_FP, true, false> __VA_ARGS__; } \
                else { if (DZV) KERNEL<false, V_FP, false, true> __VA_ARGS__; else KERNEL<false, V_FP, false, false> __VA_ARGS__; } \
            } else {                                                                                       \
                if (DSV) { if (DZV) KERNEL<false, V_BP, true, true> __VA_ARGS__; else KERNEL<false, V_BP, true, false> __VA_ARGS__; } \
                else { if (DZV) KERNEL<false, V_BP, false, true> __VA_ARGS__; else KERNEL<false, V_BP, false, false> __VA_ARGS__; } \
            }                                                                                              \
        }                                                                                                  \
    } while (0)

static inline unsigned blocks_for(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace srf

using namespace srf;

extern "C" {

int srf_project(const float *lors, int64_t n, int64_t lor_stride, const float *tof_values, int64_t tof_stride,
                const float *image, srf_grid_t grid, const srf_tof_t *tof, float *out, void *stream)
{
    int rc = check_events(lors, n, lor_stride, tof_values, tof_stride, tof);
    if (rc) return rc;
    if (n > 0 && (!out || !image)) return SRF_ERR_INVALID_ARG;
    LaunchGeom G;
    if ((rc = make_geom(grid, tof, G))) return rc;
    if (n == 0) return SRF_OK;
    hipStream_t s = (hipStream_t)stream;
    const bool ds = G.dy != 1.0f, dz = G.dz != 1.0f;
    SRF_DISPATCH(k_gather, tof != nullptr, V_FP, ds, dz,
                 <<<blocks_for(n), 256, 0, s>>>(lors, n, (int)lor_stride, tof_values, (int)tof_stride, image, G,
                                                out));
    return launch_status();
}

static int backproject_impl(const float *lors, int64_t n, int64_t lor_stride, const float *tof_values,
                            int64_t tof_stride, const float *values, const float *y, const float *p, int kind,
                            float eps, srf_grid_t grid, const srf_tof_t *tof, srf_bp_mode_t mode, int accumulate,
                            float *image, void *stream)
{
    int rc = check_events(lors, n, lor_stride, tof_values, tof_stride, tof);
    if (rc) return rc;
    if (!image) return SRF_ERR_INVALID_ARG;
    if (mode != SRF_BP_REFERENCE && mode != SRF_BP_ADJOINT) return SRF_ERR_INVALID_ARG;
    LaunchGeom G;
    if ((rc = make_geom(grid, tof, G))) return rc;
    hipStream_t s = (hipStream_t)stream;
    const int64_t nvox = (int64_t)grid.nx * grid.ny * grid.nz;
    if (!accumulate) {
        if ((rc = fill_f32(image, 0.0f, nvox, s))) return rc;
    }
    if (n == 0) return SRF_OK;
    const int var = mode == SRF_BP_ADJOINT ? V_FP : V_BP;
    const bool ds = G.dy != 1.0f, dz = G.dz != 1.0f;
    SRF_DISPATCH(k_scatter, tof != nullptr, var, ds, dz,
                 <<<blocks_for(n), 256, 0, s>>>(lors, n, (int)lor_stride, tof_values, (int)tof_stride, values, y,
                                                p, kind, eps, G, image));
    return launch_status();
}

int srf_backproject(const float *lors, int64_t n, int64_t lor_stride, const float *tof_values, int64_t tof_stride,
                    const float *values, srf_grid_t grid, const srf_tof_t *tof, srf_bp_mode_t mode, int accumulate,
                    float *image, void *stream)
{
    return backproject_impl(lors, n, lor_stride, tof_values, tof_stride, values, nullptr, nullptr, 0, 0.0f, grid,
                            tof, mode, accumulate, image, stream);
}

int srf_backproject_ratio(const float *lors, int64_t n, int64_t lor_stride, const float *tof_values,
                          int64_t tof_stride, const float *y, const float *p, srf_ratio_kind_t kind, float eps,
                          srf_grid_t grid, const srf_tof_t *tof, srf_bp_mode_t mode, int accumulate, float *image,
                          void *stream)
{
    if (n > 0 && (!y || !p)) return SRF_ERR_INVALID_ARG;
    if (kind != SRF_RATIO_PLAIN && kind != SRF_RATIO_DIV_NO_NAN) return SRF_ERR_INVALID_ARG;
    return backproject_impl(lors, n, lor_stride, tof_values, tof_stride, nullptr, y, p, (int)kind, eps, grid, tof,
                            mode, accumulate, image, stream);
}

int srf_lor_digest(const float *lors, int64_t n, int64_t lor_stride, const float *tof_values, int64_t tof_stride,
                   srf_grid_t grid, const srf_tof_t *tof, srf_bp_mode_t mode, uint64_t *hash, int32_t *count,
                   void *stream)
{
    int rc = check_events(lors, n, lor_stride, tof_values, tof_stride, tof);
    if (rc) return rc;
    if (n > 0 && (!hash || !count)) return SRF_ERR_INVALID_ARG;
    if (mode != SRF_BP_REFERENCE && mode != SRF_BP_ADJOINT) return SRF_ERR_INVALID_ARG;
    LaunchGeom G;
    if ((rc = make_geom(grid, tof, G))) return rc;
    if (n == 0) return SRF_OK;
    hipStream_t s = (hipStream_t)stream;
    const int var = mode == SRF_BP_ADJOINT ? V_FP : V_BP;
    const bool ds = G.dy != 1.0f, dz = G.dz != 1.0f;
    SRF_DISPATCH(k_digest, tof != nullptr, var, ds, dz,
                 <<<blocks_for(n), 256, 0, s>>>(lors, n, (int)lor_stride, tof_values, (int)tof_stride, G, hash,
                                                count));
    return launch_status();
}

}  // extern "C"
