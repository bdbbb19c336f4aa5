// srf_cnn.hip -- BF16 MFMA kernels of the scatter surrogate CNN (training and
// inference), and their C-ABI entry points (include/srf_hip.h, srf_cnn_*).
//
// Tensors are channels-last: act[v][c], v = (z*H + y)*W + x, with the channel
// count padded to a multiple of 8 (one MFMA k-step covers 8 channels) and
// BF16 storage; weights are [tap][cout][cin] (cin fastest, padded like the
// activations) in BF16 with FP32 master copies kept by the optimiser.
//
// Convolutions (3x3x3 or 1x1x1, stride 1, zero padding) are implicit GEMMs on
// v_mfma_f32_32x32x8bf16_1k (A: 32 voxels x 8 channels, B: 8 channels x 32
// output channels, FP32 accumulation).  Each wavefront owns a 32-voxel x
// 32-channel block of the output; a workgroup stacks four of them along the
// voxels.  Fragments are loaded straight from global memory (8 bytes per lane
// per operand); these networks are small enough that L1/L2 reuse suffices.
//   k_conv        forward (bias + optional leaky ReLU epilogue), and the
//                 input gradient as a forward pass of flipped, transposed weights
//   k_conv_wgrad  weight gradient: per tap, dW[co][ci] += sum_v dY[v][co] X[v+o][ci]
//                 (split over voxel chunks, FP32 atomics into dW)
// plus the elementwise kernels of the network (pooling, upsampling, channel
// copies, activation backward, bias gradient, MSE, Adam, casts).
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <math.h>
#include <stdint.h>

#include "srf_hip.h"
#include "srf_internal.hpp"

namespace srf {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

__device__ __forceinline__ unsigned short f2bf(float f)
{
    // round to nearest even (NaN kept as a quiet NaN)
    unsigned u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (unsigned short)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

struct ConvDims {
    int D, H, W;      // spatial size (same in and out)
    int cin_p;        // padded input channels (multiple of 8), also the input row stride
    int cout;         // real output channels
    int cout_p;       // output channels padded to a multiple of 32 (weight rows)
    int out_stride;   // output row stride (channels)
    int out_off;      // first output channel
    int taps;         // 27 or 1
    float slope;      // leaky ReLU slope; < 0: no activation
};

__device__ __forceinline__ void tap_offset(int t, int taps, int &dz, int &dy, int &dx)
{
    if (taps == 1) { dz = dy = dx = 0; return; }
    dz = t / 9 - 1;
    dy = (t / 3) % 3 - 1;
    dx = t % 3 - 1;
}

// One 32 x 32 output block per wavefront, four wavefronts per workgroup along v.
__global__ __launch_bounds__(256) void k_conv(const unsigned short *__restrict__ X,
                                              const unsigned short *__restrict__ Wt,
                                              const float *__restrict__ bias, ConvDims d,
                                              unsigned short *__restrict__ Ybf, float *__restrict__ Yf)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int V = d.D * d.H * d.W;
    const int v0 = (blockIdx.x * 4 + wave) * 32;
    const int n0 = blockIdx.y * 32;
    if (v0 >= V) return;
    const int m = lane & 31;              // A row / B column within the block
    const int kq = (lane >> 5) * 4;       // first of the lane's 4 k values
    const int v = v0 + m;
    const bool vin = v < V;
    const int x = vin ? v % d.W : 0, y = vin ? (v / d.W) % d.H : 0, z = vin ? v / (d.W * d.H) : 0;
    f32x16 acc = {0};
    for (int t = 0; t < d.taps; ++t) {
        int dz, dy, dx;
        tap_offset(t, d.taps, dz, dy, dx);
        const int zz = z + dz, yy = y + dy, xx = x + dx;
        const bool ok = vin && zz >= 0 && zz < d.D && yy >= 0 && yy < d.H && xx >= 0 && xx < d.W;
        const int64_t src = ((int64_t)(zz * d.H + yy) * d.W + xx) * d.cin_p;
        const unsigned short *wrow = Wt + ((int64_t)t * d.cout_p + n0 + m) * d.cin_p;
        for (int c = 0; c < d.cin_p; c += 8) {
            bf16x4 a = {0, 0, 0, 0};
            if (ok) a = *reinterpret_cast<const bf16x4 *>(X + src + c + kq);
            const bf16x4 b = *reinterpret_cast<const bf16x4 *>(wrow + c + kq);
            acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
        }
    }
    // C layout: lane l holds column n = l % 32, rows 8*(i/4) + 4*(l/32) + i%4
    const int n = n0 + m;
    if (n >= d.cout) return;
    const float bn = bias ? bias[n] : 0.0f;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int row = v0 + 8 * (i / 4) + kq + (i % 4);
        if (row >= V) continue;
        float r = acc[i] + bn;
        if (d.slope >= 0.0f && r < 0.0f) r *= d.slope;
        const int64_t o = (int64_t)row * d.out_stride + d.out_off + n;
        if (Ybf) Ybf[o] = f2bf(r);
        if (Yf) Yf[o] = r;
    }
}

// Weight gradient of one tap: dW[t][co][ci] += sum_v dY[v][co] * X[v + o_t][ci]
// (dY rows of stride dy_stride, channels [0, cout); X rows of stride cin_p).
// Grid: (voxel chunks, taps, cout blocks of 32 * cin blocks of 32).
__global__ __launch_bounds__(256) void k_conv_wgrad(const unsigned short *__restrict__ X,
                                                    const unsigned short *__restrict__ dY, int dy_stride,
                                                    ConvDims d, int chunk, float *__restrict__ dW)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int V = d.D * d.H * d.W;
    const int t = blockIdx.y;
    const int ncb = (d.cin_p + 31) / 32;
    const int co0 = (blockIdx.z / ncb) * 32, ci0 = (blockIdx.z % ncb) * 32;
    int dz, dy, dx;
    tap_offset(t, d.taps, dz, dy, dx);
    const int m = lane & 31, kq = (lane >> 5) * 4;
    const int co = co0 + m, ci = ci0 + m;
    const int vbeg = blockIdx.x * chunk, vend = min(V, vbeg + chunk);
    f32x16 acc = {0};
    // each wavefront walks every fourth group of 8 voxels of the chunk
    for (int vb = vbeg + wave * 8; vb < vend; vb += 32) {
        bf16x4 a, b;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int v = vb + kq + j;
            unsigned short av = 0, bv = 0;
            if (v < vend) {
                if (co < d.cout) av = dY[(int64_t)v * dy_stride + co];
                const int x = v % d.W, y = (v / d.W) % d.H, z = v / (d.W * d.H);
                const int zz = z + dz, yy = y + dy, xx = x + dx;
                if (ci < d.cin_p && zz >= 0 && zz < d.D && yy >= 0 && yy < d.H && xx >= 0 && xx < d.W)
                    bv = X[((int64_t)(zz * d.H + yy) * d.W + xx) * d.cin_p + ci];
            }
            a[j] = (short)av;
            b[j] = (short)bv;
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int r = co0 + 8 * (i / 4) + kq + (i % 4);   // row = output channel
        const int c = ci;                                   // column = input channel
        if (r < d.cout && c < d.cin_p && acc[i] != 0.0f)
            atomicAdd(dW + ((int64_t)t * d.cout_p + r) * d.cin_p + c, acc[i]);
    }
}

// W'[t'][ci][co] = W[26 - t'][co][ci]: the input gradient of a stride-1,
// zero-padded convolution is a convolution of dY with this kernel.
__global__ void k_flip_weights(const unsigned short *__restrict__ W, int taps, int cout, int cout_p, int cin_p,
                               int cin_real, unsigned short *__restrict__ Wf, int fcout_p, int fcin_p)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t n = (int64_t)taps * fcout_p * fcin_p;
    if (i >= n) return;
    const int cc = (int)(i % fcin_p);            // new input channel = old output channel
    const int rr = (int)((i / fcin_p) % fcout_p);// new output channel = old input channel
    const int t = (int)(i / ((int64_t)fcin_p * fcout_p));
    unsigned short val = 0;
    if (cc < cout && rr < cin_real) val = W[((int64_t)(taps - 1 - t) * cout_p + cc) * cin_p + rr];
    Wf[i] = val;
}

// ------------------------------------------------------------ elementwise --

// 2x2x2 average pooling, channels-last bf16 (C real channels, stride cs)
__global__ void k_pool(const unsigned short *__restrict__ X, int D, int H, int W, int C, int xs,
                       unsigned short *__restrict__ Y, int ys)
{
    const int D2 = D / 2, H2 = H / 2, W2 = W / 2;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (int64_t)D2 * H2 * W2 * C) return;
    const int c = (int)(i % C);
    const int64_t v = i / C;
    const int x = (int)(v % W2), y = (int)((v / W2) % H2), z = (int)(v / (W2 * H2));
    float s = 0.0f;
    for (int k = 0; k < 8; ++k) {
        const int zz = 2 * z + (k >> 2), yy = 2 * y + ((k >> 1) & 1), xx = 2 * x + (k & 1);
        s += bf2f(X[((int64_t)(zz * H + yy) * W + xx) * xs + c]);
    }
    Y[v * ys + c] = f2bf(0.125f * s);
}

// gradient of the pooling: dX[child] = dY[parent] / 8 (written, not added)
__global__ void k_pool_bwd(const float *__restrict__ dY, int D, int H, int W, int C, int dys,
                           float *__restrict__ dX, int dxs)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (int64_t)D * H * W * C) return;
    const int c = (int)(i % C);
    const int64_t v = i / C;
    const int x = (int)(v % W), y = (int)((v / W) % H), z = (int)(v / (W * H));
    const int64_t p = ((int64_t)((z / 2) * (H / 2) + y / 2) * (W / 2) + x / 2);
    dX[v * dxs + c] = 0.125f * dY[p * dys + c];
}

// nearest 2x upsampling into channels [off, off + C) of a wider tensor
__global__ void k_upsample(const unsigned short *__restrict__ X, int D, int H, int W, int C, int xs,
                           unsigned short *__restrict__ Y, int ys, int off)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (int64_t)8 * D * H * W * C) return;
    const int c = (int)(i % C);
    const int64_t v = i / C;
    const int W2 = 2 * W, H2 = 2 * H;
    const int x = (int)(v % W2), y = (int)((v / W2) % H2), z = (int)(v / (W2 * H2));
    Y[v * ys + off + c] = X[((int64_t)((z / 2) * H + y / 2) * W + x / 2) * xs + c];
}

// gradient of the upsampling: dX[parent] = sum of the 8 children's dY (f32)
__global__ void k_upsample_bwd(const float *__restrict__ dY, int D, int H, int W, int C, int dys, int off,
                               float *__restrict__ dX, int dxs)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (int64_t)D * H * W * C) return;
    const int c = (int)(i % C);
    const int64_t v = i / C;
    const int x = (int)(v % W), y = (int)((v / W) % H), z = (int)(v / (W * H));
    const int W2 = 2 * W, H2 = 2 * H;
    float s = 0.0f;
    for (int k = 0; k < 8; ++k) {
        const int zz = 2 * z + (k >> 2), yy = 2 * y + ((k >> 1) & 1), xx = 2 * x + (k & 1);
        s += dY[((int64_t)(zz * H2 + yy) * W2 + xx) * dys + off + c];
    }
    dX[v * dxs + c] = s;
}

// copy channels [0,C) of X (stride xs) into channels [off, off+C) of Y (stride ys); bf16 or f32
template <class T>
__global__ void k_copy_channels(const T *__restrict__ X, int64_t V, int C, int xs, int xoff, T *__restrict__ Y,
                                int ys, int yoff, int accumulate)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= V * C) return;
    const int c = (int)(i % C);
    const int64_t v = i / C;
    if (accumulate) Y[v * ys + yoff + c] += X[v * xs + xoff + c];
    else Y[v * ys + yoff + c] = X[v * xs + xoff + c];
}

// dPre = dAct * (act > 0 ? 1 : slope) -> bf16 (stride ps, zero-filled pad channels);
// act > 0 exactly when the pre-activation is > 0 (slope > 0)
__global__ void k_act_bwd(const float *__restrict__ dA, int das, const unsigned short *__restrict__ A, int as,
                          int64_t V, int C, int Cp, float slope, unsigned short *__restrict__ dP)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= V * Cp) return;
    const int c = (int)(i % Cp);
    const int64_t v = i / Cp;
    float g = 0.0f;
    if (c < C) {
        g = dA[v * das + c];
        if (slope >= 0.0f && !(bf2f(A[v * as + c]) > 0.0f)) g *= slope;
    }
    dP[v * Cp + c] = f2bf(g);
}

// db[c] = sum_v dP[v][c] (bf16 rows of stride Cp), one block per channel
__global__ __launch_bounds__(256) void k_bias_grad(const unsigned short *__restrict__ dP, int64_t V, int Cp,
                                                   float *__restrict__ db)
{
    __shared__ float red[256];
    const int c = blockIdx.x;
    float s = 0.0f;
    for (int64_t v = threadIdx.x; v < V; v += 256) s += bf2f(dP[v * Cp + c]);
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[c] += red[0];
}

// loss = w mean((p - t)^2) accumulated into loss[0]; grad = 2 w (p - t) / n
__global__ __launch_bounds__(256) void k_mse(const float *__restrict__ p, const float *__restrict__ t, int64_t n,
                                             float w, float *__restrict__ grad, float *__restrict__ loss)
{
    __shared__ float red[256];
    float s = 0.0f;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const float e = p[i] - t[i];
        grad[i] = 2.0f * w * e / (float)n;
        s += e * e;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, w * red[0] / (float)n);
}

// Adam (Kingma & Ba 2015) on FP32 master weights; writes the bf16 copy.
__global__ void k_adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, int64_t n, float lr, float b1, float b2, float eps, float bc1,
                       float bc2, float gscale, unsigned short *__restrict__ wbf)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i] * gscale;
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
    w[i] = wi;
    if (wbf) wbf[i] = f2bf(wi);
}

__global__ void k_f32_to_bf16(const float *__restrict__ x, int64_t n, unsigned short *__restrict__ y)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) y[i] = f2bf(x[i]);
}

__global__ void k_bf16_to_f32(const unsigned short *__restrict__ x, int64_t n, float *__restrict__ y)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) y[i] = bf2f(x[i]);
}

// Surrogate input features per sinogram bin from the emission and attenuation
// line integrals along the bin's LOR (scale[0] = 1 / mean emission integral)
// and an optional extra sinogram (scale[1]):
//   [e, exp(-mu), e * exp(-mu), mu / 4, extra * scale[1], 0, 0, 0], e = em * scale[0]
__global__ void k_features(const float *__restrict__ em, const float *__restrict__ mu,
                           const float *__restrict__ extra, int64_t n, const float *__restrict__ scale,
                           unsigned short *__restrict__ x)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float e = em[i] * scale[0];
    const float a = expf(-mu[i]);
    ushort4 lo, hi;
    lo.x = f2bf(e);
    lo.y = f2bf(a);
    lo.z = f2bf(e * a);
    lo.w = f2bf(0.25f * mu[i]);
    hi.x = extra ? f2bf(extra[i] * scale[1]) : (unsigned short)0;
    hi.y = hi.z = hi.w = 0;
    reinterpret_cast<ushort4 *>(x)[2 * i] = lo;
    reinterpret_cast<ushort4 *>(x)[2 * i + 1] = hi;
}

static inline unsigned nblk(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace srf

using namespace srf;

extern "C" {

int srf_cnn_conv(const void *x, const void *w, const float *bias, int32_t D, int32_t H, int32_t W, int32_t cin_p,
                 int32_t cout, int32_t cout_p, int32_t taps, int32_t out_stride, int32_t out_off, float slope,
                 void *y_bf16, float *y_f32, void *stream)
{
    if (!x || !w || D <= 0 || H <= 0 || W <= 0) return SRF_ERR_INVALID_ARG;
    if (cin_p <= 0 || cin_p % 8 || cout <= 0 || cout_p % 32 || cout > cout_p) return SRF_ERR_INVALID_ARG;
    if (taps != 1 && taps != 27) return SRF_ERR_INVALID_ARG;
    if (out_off < 0 || out_off + cout > out_stride || (!y_bf16 && !y_f32)) return SRF_ERR_INVALID_ARG;
    if ((reinterpret_cast<uintptr_t>(x) | reinterpret_cast<uintptr_t>(w)) & 7) return SRF_ERR_INVALID_ARG;
    ConvDims d{D, H, W, cin_p, cout, cout_p, out_stride, out_off, taps, slope};
    const int64_t V = (int64_t)D * H * W;
    dim3 grid((unsigned)((V + 127) / 128), (unsigned)(cout_p / 32));
    hipLaunchKernelGGL(k_conv, grid, dim3(256), 0, (hipStream_t)stream, (const unsigned short *)x,
                       (const unsigned short *)w, bias, d, (unsigned short *)y_bf16, y_f32);
    return launch_status();
}

int srf_cnn_conv_wgrad(const void *x, const void *dy, int32_t dy_stride, int32_t D, int32_t H, int32_t W,
                       int32_t cin_p, int32_t cout, int32_t cout_p, int32_t taps, float *dw, void *stream)
{
    if (!x || !dy || !dw || D <= 0 || H <= 0 || W <= 0) return SRF_ERR_INVALID_ARG;
    if (cin_p <= 0 || cin_p % 8 || cout <= 0 || cout_p % 32 || cout > cout_p || dy_stride < cout)
        return SRF_ERR_INVALID_ARG;
    if (taps != 1 && taps != 27) return SRF_ERR_INVALID_ARG;
    ConvDims d{D, H, W, cin_p, cout, cout_p, 0, 0, taps, -1.0f};
    const int64_t V = (int64_t)D * H * W;
    const int chunk = 4096;
    dim3 grid((unsigned)((V + chunk - 1) / chunk), (unsigned)taps,
              (unsigned)((cout_p / 32) * ((cin_p + 31) / 32)));
    hipLaunchKernelGGL(k_conv_wgrad, grid, dim3(256), 0, (hipStream_t)stream, (const unsigned short *)x,
                       (const unsigned short *)dy, dy_stride, d, chunk, dw);
    return launch_status();
}

int srf_cnn_flip_weights(const void *w, int32_t taps, int32_t cout, int32_t cout_p, int32_t cin_p, int32_t cin,
                         void *wf, int32_t fcout_p, int32_t fcin_p, void *stream)
{
    if (!w || !wf || (taps != 1 && taps != 27) || fcin_p % 8 || fcout_p % 32 || fcin_p < cout || fcout_p < cin)
        return SRF_ERR_INVALID_ARG;
    const int64_t n = (int64_t)taps * fcout_p * fcin_p;
    hipLaunchKernelGGL(k_flip_weights, dim3(nblk(n)), dim3(256), 0, (hipStream_t)stream,
                       (const unsigned short *)w, taps, cout, cout_p, cin_p, cin, (unsigned short *)wf, fcout_p,
                       fcin_p);
    return launch_status();
}

int srf_cnn_pool(const void *x, int32_t D, int32_t H, int32_t W, int32_t C, int32_t xs, void *y, int32_t ys,
                 void *stream)
{
    if (!x || !y || D % 2 || H % 2 || W % 2 || C <= 0) return SRF_ERR_INVALID_ARG;
    hipLaunchKernelGGL(k_pool, dim3(nblk((int64_t)D * H * W / 8 * C)), dim3(256), 0, (hipStream_t)stream,
                       (const unsigned short *)x, D, H, W, C, xs, (unsigned short *)y, ys);
    return launch_status();
}

int srf_cnn_pool_bwd(const float *dy, int32_t D, int32_t H, int32_t W, int32_t C, int32_t dys, float *dx,
                     int32_t dxs, void *stream)
{
    if (!dy || !dx || D % 2 || H % 2 || W % 2 || C <= 0) return SRF_ERR_INVALID_ARG;
    hipLaunchKernelGGL(k_pool_bwd, dim3(nblk((int64_t)D * H * W * C)), dim3(256), 0, (hipStream_t)stream, dy, D, H,
                       W, C, dys, dx, dxs);
    return launch_status();
}

int srf_cnn_upsample(const void *x, int32_t D, int32_t H, int32_t W, int32_t C, int32_t xs, void *y, int32_t ys,
                     int32_t off, void *stream)
{
    if (!x || !y || C <= 0 || off + C > ys) return SRF_ERR_INVALID_ARG;
    hipLaunchKernelGGL(k_upsample, dim3(nblk((int64_t)8 * D * H * W * C)), dim3(256), 0, (hipStream_t)stream,
                       (const unsigned short *)x, D, H, W, C, xs, (unsigned short *)y, ys, off);
    return launch_status();
}

int srf_cnn_upsample_bwd(const float *dy, int32_t D, int32_t H, int32_t W, int32_t C, int32_t dys, int32_t off,
                         float *dx, int32_t dxs, void *stream)
{
    if (!dy || !dx || C <= 0) return SRF_ERR_INVALID_ARG;
    hipLaunchKernelGGL(k_upsample_bwd, dim3(nblk((int64_t)D * H * W * C)), dim3(256), 0, (hipStream_t)stream, dy,
                       D, H, W, C, dys, off, dx, dxs);
    return launch_status();
}

int srf_cnn_copy_channels(const void *x, int64_t V, int32_t C, int32_t xs, int32_t xoff, void *y, int32_t ys,
                          int32_t yoff, int32_t f32, int32_t accumulate, void *stream)
{
    if (!x || !y || C <= 0 || V < 0) return SRF_ERR_INVALID_ARG;
    if (f32)
        hipLaunchKernelGGL(k_copy_channels<float>, dim3(nblk(V * C)), dim3(256), 0, (hipStream_t)stream,
                           (const float *)x, V, C, xs, xoff, (float *)y, ys, yoff, accumulate);
    else {
        if (accumulate) return SRF_ERR_UNSUPPORTED;
        hipLaunchKernelGGL(k_copy_channels<unsigned short>, dim3(nblk(V * C)), dim3(256), 0, (hipStream_t)stream,
                           (const unsigned short *)x, V, C, xs, xoff, (unsigned short *)y, ys, yoff, 0);
    }
    return launch_status();
}

int srf_cnn_act_bwd(const float *da, int32_t das, const void *a, int32_t as, int64_t V, int32_t C, int32_t Cp,
                    float slope, void *dp, void *stream)
{
    if (!da || !a || !dp || C <= 0 || Cp < C || Cp % 8) return SRF_ERR_INVALID_ARG;
    hipLaunchKernelGGL(k_act_bwd, dim3(nblk(V * Cp)), dim3(256), 0, (hipStream_t)stream, da, das,
                       (const unsigned short *)a, as, V, C, Cp, slope, (unsigned short *)dp);
    return launch_status();
}

int srf_cnn_bias_grad(const void *dp, int64_t V, int32_t C, int32_t Cp, float *db, void *stream)
{
    if (!dp || !db || C <= 0) return SRF_ERR_INVALID_ARG;
    hipLaunchKernelGGL(k_bias_grad, dim3((unsigned)C), dim3(256), 0, (hipStream_t)stream, (const unsigned short *)dp,
                       V, Cp, db);
    return launch_status();
}

int srf_cnn_mse(const float *p, const float *t, int64_t n, float weight, float *grad, float *loss, void *stream)
{
    if (!p || !t || !grad || !loss || n <= 0 || !(weight >= 0.0f)) return SRF_ERR_INVALID_ARG;
    hipLaunchKernelGGL(k_mse, dim3((unsigned)((n + 255) / 256 < 1024 ? (n + 255) / 256 : 1024)), dim3(256), 0,
                       (hipStream_t)stream, p, t, n, weight, grad, loss);
    return launch_status();
}

int srf_cnn_adam(float *w, const float *g, float *m, float *v, int64_t n, float lr, float b1, float b2, float eps,
                 int32_t step, float gscale, void *wbf, void *stream)
{
    if (!w || !g || !m || !v || n <= 0 || step <= 0) return SRF_ERR_INVALID_ARG;
    const float bc1 = 1.0f - powf(b1, (float)step), bc2 = 1.0f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3(nblk(n)), dim3(256), 0, (hipStream_t)stream, w, g, m, v, n, lr, b1, b2, eps, bc1,
                       bc2, gscale, (unsigned short *)wbf);
    return launch_status();
}

int srf_cnn_features(const float *em_int, const float *mu_int, const float *extra, int64_t n, const float *scale,
                     void *x, void *stream)
{
    if (!em_int || !mu_int || !scale || !x || n < 0) return SRF_ERR_INVALID_ARG;
    if (reinterpret_cast<uintptr_t>(x) & 7) return SRF_ERR_INVALID_ARG;
    if (n == 0) return SRF_OK;
    hipLaunchKernelGGL(k_features, dim3(nblk(n)), dim3(256), 0, (hipStream_t)stream, em_int, mu_int, extra, n,
                       scale, (unsigned short *)x);
    return launch_status();
}

int srf_cnn_cast(const void *x, int64_t n, int32_t to_bf16, void *y, void *stream)
{
    if (!x || !y || n < 0) return SRF_ERR_INVALID_ARG;
    if (to_bf16)
        hipLaunchKernelGGL(k_f32_to_bf16, dim3(nblk(n)), dim3(256), 0, (hipStream_t)stream, (const float *)x, n,
                           (unsigned short *)y);
    else
        hipLaunchKernelGGL(k_bf16_to_f32, dim3(nblk(n)), dim3(256), 0, (hipStream_t)stream,
                           (const unsigned short *)x, n, (float *)y);
    return launch_status();
}

}  // extern "C"
