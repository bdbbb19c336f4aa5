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
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

// float -> bf16 bits, round to nearest even (gfx950's v_cvt_pk_bf16_f32)
__device__ __forceinline__ unsigned short f2bf(float f) { return __builtin_bit_cast(unsigned short, (__bf16)f); }

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

// Convolution as an implicit GEMM on gfx950's v_mfma_f32_16x16x32_bf16
// (A: 16 voxels x 32 K values, B: 32 K values x 16 output channels, FP32
// accumulation).  K runs over (tap, group of 8 input channels).
//
// A workgroup of four wavefronts owns a tile of up to 128 voxels (tx x ty x tz,
// x fastest; wavefront w rows 32w..32w+31, two 16-row blocks) and NBW blocks of
// 16 output channels.  Per chunk of 8, 16 or 32 input channels the tile's input
// halo is staged in LDS with coalesced 16-byte loads: zero outside the volume
// and outside the tile's own sample (a batch stacks samples along z, Ds slices
// each; a tile never straddles two), so taps never read a neighbouring
// sample.  A lane then reads its 8 K values of a voxel as one 16-byte LDS
// access; a halo voxel's channel groups are padded to an odd count of 16-byte
// words so that the 16 voxels of a row block start in different banks.
// Weights (a few KiB to 440 KiB per layer) are read from global memory
// through the caches.
typedef float f32x4 __attribute__((ext_vector_type(4)));

struct ConvTile {
    int tx, ty, tz;             // tile extent
    int ntx, nty;               // tiles along x and y
    int Ds;                     // slices per sample (tz divides it)
    int nblk;                   // tiles of the launch
    float inv_hx, inv_hyx;      // 1 / halo row, 1 / halo plane (voxels)
    float inv_tx, inv_txy;      // 1 / tile row, 1 / tile plane
};

// The input of a convolution: channels [0, cc) are the nearest-neighbour 2x
// upsampling of `coarse` (rows of cc channels on the half-size grid) and
// channels [cc, cin_p) come from `fine` (rows of cin_p - cc channels): the
// decoder's concatenation read in place.  cc = 0: every channel from fine.
struct ConvIn {
    const unsigned short *fine;
    const unsigned short *coarse;
    int cc;
};

// A 1x1 convolution to one channel applied to a layer's output in its
// epilogue: out[v] = b + sum_c bf16(y[v][c]) w[c] (the bf16 values the layer
// stores), so that the network's last layer needs no pass of its own.
struct ConvHead {
    const unsigned short *w;   // cout bf16 weights
    const float *b;            // the bias (one value)
    float *out;                // one value per voxel
};

// floor(a / b) for 0 <= a < 2^24 and b >= 1 given inv = 1/b in float: the
// float quotient is off by at most one, and one integer step corrects it.
__device__ __forceinline__ int idiv(int a, int b, float inv)
{
    int q = (int)((float)a * inv);
    const int r = a - q * b;
    q += r >= b ? 1 : 0;
    q -= r < 0 ? 1 : 0;
    return q;
}
constexpr int CONV_TILE = 128;
constexpr int CONV_LDS_MAX = 65536;

// Offset in halo voxels of tap t from the halo voxel of the tap (-1,-1,-1) corner.
__device__ __forceinline__ int tap_delta(int t, int taps, int hx, int hyx)
{
    if (taps == 1) return 0;
    return (t / 9) * hyx + ((t / 3) % 3) * hx + t % 3;
}

template <int TAPS, int NG, int NBW>
__device__ __forceinline__ void conv_chunk(const u16x8 *__restrict__ halo, const int (&hb)[2], int gp, int hx, int hyx,
                                           const unsigned short *__restrict__ wl, int64_t tap_stride, int nb_stride,
                                           int q, f32x4 (&acc)[2][NBW])
{
    constexpr int G = TAPS * NG;          // K groups of 8 in this chunk
    constexpr int S = (G + 3) / 4;        // MFMA steps (4 groups each)
    constexpr int K = 4 / NG;             // taps per step
    // lane l's K group q is (tap K s + q / NG, channel group q % NG)
    const int g = q & (NG - 1), tq = q / NG;
    const u16x8 *a0 = halo + hb[0] * gp + g, *a1 = halo + hb[1] * gp + g;
    const unsigned short *wq = wl + tq * tap_stride + 8 * g;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        // the tap offsets of this step: uniform candidates, one picked per lane
        int delta;
        if constexpr (K == 1) {
            delta = tap_delta(s, TAPS, hx, hyx);
        } else if constexpr (K == 2) {
            const int d0 = tap_delta(2 * s, TAPS, hx, hyx), d1 = tap_delta(2 * s + 1, TAPS, hx, hyx);
            delta = tq ? d1 : d0;
        } else {
            const int d0 = tap_delta(4 * s, TAPS, hx, hyx), d1 = tap_delta(4 * s + 1, TAPS, hx, hyx);
            const int d2 = tap_delta(4 * s + 2, TAPS, hx, hyx), d3 = tap_delta(4 * s + 3, TAPS, hx, hyx);
            delta = tq == 0 ? d0 : tq == 1 ? d1 : tq == 2 ? d2 : d3;
        }
        const bool valid = (G % 4 == 0 || s < S - 1) || (4 * s + q < G);
        u16x8 a[2], w[NBW];
#pragma unroll
        for (int r = 0; r < 2; ++r) a[r] = u16x8{0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int j = 0; j < NBW; ++j) w[j] = u16x8{0, 0, 0, 0, 0, 0, 0, 0};
        if (valid) {
            a[0] = a0[delta * gp];
            a[1] = a1[delta * gp];
            const unsigned short *ws = wq + (int64_t)(K * s) * tap_stride;
#pragma unroll
            for (int j = 0; j < NBW; ++j) w[j] = *reinterpret_cast<const u16x8 *>(ws + j * nb_stride);
        }
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int j = 0; j < NBW; ++j)
                acc[r][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a[r]),
                                                                    __builtin_bit_cast(bf16x8, w[j]), acc[r][j], 0,
                                                                    0, 0);
    }
}

template <int TAPS, int NBW>
__global__ __launch_bounds__(256) void k_conv(ConvIn in, const unsigned short *__restrict__ Wt,
                                              const float *__restrict__ bias, ConvDims d, ConvTile T,
                                              unsigned short *__restrict__ Ybf, float *__restrict__ Yf, ConvHead hd)
{
    constexpr int R = TAPS == 27 ? 1 : 0;
    constexpr int STAGE = 4;   // halo loads in flight per thread
    extern __shared__ u16x8 halo[];
    // the hardware deals workgroup i to XCD i % 8: give every XCD a contiguous
    // run of tiles, so that the halos neighbouring tiles share hit its L2
    const int per = gridDim.x >> 3;
    const int b = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (b >= T.nblk) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int col = lane & 15, q = lane >> 4;
    const int bx = b % T.ntx, by = (b / T.ntx) % T.nty, bz = b / (T.ntx * T.nty);
    const int x0 = bx * T.tx, y0 = by * T.ty, z0 = bz * T.tz;
    const int zs0 = (z0 / T.Ds) * T.Ds;   // first slice of the tile's sample
    const int hx = T.tx + 2 * R, hy = T.ty + 2 * R, hz = T.tz + 2 * R;
    const int hyx = hx * hy, HV = hyx * hz;
    const int txy = T.tx * T.ty, nvox = txy * T.tz;
    const int n0 = blockIdx.y * 16 * NBW;
    const int Hc = d.H >> 1, Wc = d.W >> 1;
    // this lane's A rows (clamped into the tile; rows past it are not stored)
    int hb[2];
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        const int tv = min(wave * 32 + r * 16 + col, nvox - 1);
        const int ltz = idiv(tv, txy, T.inv_txy), rem = tv - ltz * txy;
        const int lty = idiv(rem, T.tx, T.inv_tx), ltx = rem - lty * T.tx;
        hb[r] = (ltz * hy + lty) * hx + ltx;
    }
    f32x4 acc[2][NBW];
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int j = 0; j < NBW; ++j) acc[r][j] = f32x4{0, 0, 0, 0};
    const int64_t tap_stride = (int64_t)d.cout_p * d.cin_p;
    const int nb_stride = 16 * d.cin_p;
    for (int c0 = 0; c0 < d.cin_p;) {
        const bool up = c0 < in.cc;   // a chunk lies wholly in one of the two inputs
        const int left = ((up ? in.cc : d.cin_p) - c0) >> 3;
        const int lg = left >= 4 ? 2 : left >= 2 ? 1 : 0;   // log2 of the chunk's 8-channel groups
        const int ng = 1 << lg, gp = ng | 1;
        const unsigned short *src = up ? in.coarse + c0 : in.fine + (c0 - in.cc);
        const int stride = up ? in.cc : d.cin_p - in.cc;
        // thread i stages items i, i + 256, ... (item = 8-channel group g of
        // halo voxel hv); the halo coordinates advance by a fixed step per pass
        const int g = threadIdx.x & (ng - 1);
        const int dv = 256 >> lg;                                  // voxels per pass
        const int sa = dv / hyx, sb = (dv - sa * hyx) / hx, sc = dv - sa * hyx - sb * hx;
        int hv = threadIdx.x >> lg;
        int hzi = idiv(hv, hyx, T.inv_hyx);
        int hyi = idiv(hv - hzi * hyx, hx, T.inv_hx);
        int hxi = hv - hzi * hyx - hyi * hx;
        __syncthreads();   // the previous chunk's reads are done
        while (hv < HV) {
            u16x8 v[STAGE];
            int dst[STAGE];
#pragma unroll
            for (int u = 0; u < STAGE; ++u) {
                v[u] = u16x8{0, 0, 0, 0, 0, 0, 0, 0};
                dst[u] = hv < HV ? hv * gp + g : -1;
                const int gx = x0 + hxi - R, gy = y0 + hyi - R, gz = z0 + hzi - R;
                if (hv < HV && (unsigned)gx < (unsigned)d.W && (unsigned)gy < (unsigned)d.H && gz >= zs0 &&
                    gz < zs0 + T.Ds) {
                    const int vox = up ? ((gz >> 1) * Hc + (gy >> 1)) * Wc + (gx >> 1) : (gz * d.H + gy) * d.W + gx;
                    v[u] = *reinterpret_cast<const u16x8 *>(src + ((unsigned)vox * (unsigned)stride + 8 * g));
                }
                hv += dv;
                hxi += sc;
                hyi += sb;
                hzi += sa;
                if (hxi >= hx) {
                    hxi -= hx;
                    ++hyi;
                }
                if (hyi >= hy) {
                    hyi -= hy;
                    ++hzi;
                }
            }
#pragma unroll
            for (int u = 0; u < STAGE; ++u)
                if (dst[u] >= 0) halo[dst[u]] = v[u];
        }
        __syncthreads();
        const unsigned short *wl = Wt + (int64_t)(n0 + col) * d.cin_p + c0;
        if (lg == 2)
            conv_chunk<TAPS, 4, NBW>(halo, hb, gp, hx, hyx, wl, tap_stride, nb_stride, q, acc);
        else if (lg == 1)
            conv_chunk<TAPS, 2, NBW>(halo, hb, gp, hx, hyx, wl, tap_stride, nb_stride, q, acc);
        else
            conv_chunk<TAPS, 1, NBW>(halo, hb, gp, hx, hyx, wl, tap_stride, nb_stride, q, acc);
        c0 += 8 * ng;
    }
    // Epilogue through LDS: the wavefront's 32 x 16 NBW block (bias and
    // activation applied) goes to LDS in the MFMA layout (lane l holds column
    // l % 16, rows 4 (l / 16) + i of each 16 x 16 block) and comes back row by
    // row, so that consecutive lanes store consecutive bytes: 16-byte groups
    // of 8 bf16 / 4 float channels where the row layout allows, else single
    // values (the 1-channel head: one value per voxel, consecutive voxels).
    constexpr int RL = 16 * NBW + 1;   // floats per LDS row
    __syncthreads();                   // every wavefront is done with the halo
    float *cbuf = reinterpret_cast<float *>(halo) + wave * (32 * RL + 64);
    int64_t *orow = reinterpret_cast<int64_t *>(cbuf + 32 * RL);   // row -> global offset
#pragma unroll
    for (int j = 0; j < NBW; ++j) {
        const int n = n0 + 16 * j + col;
        const float bn = (bias && n < d.cout) ? bias[n] : 0.0f;
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                float v = acc[r][j][i] + bn;
                if (d.slope >= 0.0f && v < 0.0f) v *= d.slope;
                cbuf[(r * 16 + 4 * q + i) * RL + 16 * j + col] = v;
            }
    }
    // global offset of each row of this wavefront, -1 outside the volume
    int myvox = -1;
    if (lane < 32) {
        const int tv = wave * 32 + lane;
        const int ltz = idiv(tv, txy, T.inv_txy), rem = tv - ltz * txy;
        const int lty = idiv(rem, T.tx, T.inv_tx), ltx = rem - lty * T.tx;
        const int gx = x0 + ltx, gy = y0 + lty, gz = z0 + ltz;
        if (tv < nvox && gx < d.W && gy < d.H) myvox = (gz * d.H + gy) * d.W + gx;
        orow[lane] = myvox >= 0 ? (int64_t)myvox * d.out_stride + d.out_off + n0 : -1;
    }
    __syncthreads();   // LDS writes visible to the whole wavefront
    if (hd.out && myvox >= 0) {
        // the head on this lane's row: products of bf16 values are exact in float
        const float *src = cbuf + lane * RL;
        float sum = 0.0f;
        for (int c = 0; c < d.cout; ++c) sum += bf2f(f2bf(src[c])) * bf2f(hd.w[c]);
        hd.out[myvox] = sum + hd.b[0];
    }
    const int cvalid = min(16 * NBW, d.cout - n0);   // channels of this block to store
    const bool vec_bf = Ybf && (cvalid % 8 == 0) && ((d.out_stride | d.out_off | n0) % 8 == 0) &&
                        !(reinterpret_cast<uintptr_t>(Ybf) & 15);
    const bool vec_f = Yf && (cvalid % 4 == 0) && ((d.out_stride | d.out_off | n0) % 4 == 0) &&
                       !(reinterpret_cast<uintptr_t>(Yf) & 15);
    if (Ybf) {
        const int per = vec_bf ? cvalid >> 3 : cvalid;          // items per row
        const float inv = 1.0f / (float)per;
        const int sh = (per & (per - 1)) ? -1 : __builtin_ctz(per);
        for (int e = lane; e < 32 * per; e += 64) {
            const int row = sh >= 0 ? e >> sh : idiv(e, per, inv), k = e - row * per;
            const int64_t o = orow[row];
            if (o < 0) continue;
            const float *src = cbuf + row * RL;
            if (vec_bf) {
                u16x8 w;
#pragma unroll
                for (int c = 0; c < 8; ++c) w[c] = f2bf(src[8 * k + c]);
                *reinterpret_cast<u16x8 *>(Ybf + o + 8 * k) = w;
            } else {
                Ybf[o + k] = f2bf(src[k]);
            }
        }
    }
    if (Yf) {
        const int per = vec_f ? cvalid >> 2 : cvalid;
        const float inv = 1.0f / (float)per;
        const int sh = (per & (per - 1)) ? -1 : __builtin_ctz(per);
        for (int e = lane; e < 32 * per; e += 64) {
            const int row = sh >= 0 ? e >> sh : idiv(e, per, inv), k = e - row * per;
            const int64_t o = orow[row];
            if (o < 0) continue;
            const float *src = cbuf + row * RL;
            if (vec_f)
                *reinterpret_cast<float4 *>(Yf + o + 4 * k) =
                    float4{src[4 * k], src[4 * k + 1], src[4 * k + 2], src[4 * k + 3]};
            else
                Yf[o + k] = src[k];
        }
    }
}

// Tile of k_conv for samples of Ds x H x W: x first (up to 32), then y, then z
// (a divisor of Ds), at most 128 voxels, and a halo that fits CONV_LDS_MAX
// with 5 padded 16-byte words per voxel.
static void conv_tile(int Ds, int H, int W, int R, ConvTile &T)
{
    T.tx = W < 32 ? W : 32;
    T.ty = H < CONV_TILE / T.tx ? H : CONV_TILE / T.tx;
    T.tz = 1;
    for (int z = CONV_TILE / (T.tx * T.ty); z > 1; --z)
        if (Ds % z == 0) {
            T.tz = z;
            break;
        }
    auto bytes = [&] { return (T.tx + 2 * R) * (T.ty + 2 * R) * (T.tz + 2 * R) * 5 * 16; };
    while (bytes() > CONV_LDS_MAX && T.tz > 1) {
        int z = T.tz - 1;
        while (Ds % z) --z;
        T.tz = z;
    }
    while (bytes() > CONV_LDS_MAX && T.ty > 1) --T.ty;
    T.ntx = (W + T.tx - 1) / T.tx;
    T.nty = (H + T.ty - 1) / T.ty;
    T.Ds = Ds;
    T.nblk = 0;
    const int hx = T.tx + 2 * R, hy = T.ty + 2 * R;
    T.inv_hx = 1.0f / (float)hx;
    T.inv_hyx = 1.0f / (float)(hx * hy);
    T.inv_tx = 1.0f / (float)T.tx;
    T.inv_txy = 1.0f / (float)(T.tx * T.ty);
}

// Weight gradient of one tap: dW[t][co][ci] += sum_v dY[v][co] * X[v + o_t][ci]
// (dY rows of stride dy_stride, channels [0, cout); X rows of stride cin_p;
// samples of Ds slices stacked along z).  The general-shape path: element
// loads from global memory.  Grid: (voxel chunks, taps, cout blocks of 32 *
// cin blocks of 32).
__global__ __launch_bounds__(256) void k_conv_wgrad(const unsigned short *__restrict__ X,
                                                    const unsigned short *__restrict__ dY, int dy_stride,
                                                    ConvDims d, int Ds, int chunk, float *__restrict__ dW)
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
                const int zz = z + dz, yy = y + dy, xx = x + dx, zs = z - z % Ds;
                if (ci < d.cin_p && zz >= zs && zz < zs + Ds && yy >= 0 && yy < d.H && xx >= 0 && xx < d.W)
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

// Weight gradient on gfx950's v_mfma_f32_16x16x32_bf16: dW[t][co][ci] +=
// sum_v dY[v][co] X[v + o_t][ci], an M = cout, N = cin, K = voxels GEMM per
// tap.  A workgroup owns a 32 x 32 (co, ci) block of all taps and a run of
// voxel tiles (the forward kernel's tiles; tx a multiple of 8).  Per tile the
// dY tile and the X halo are staged in LDS channel-major, so that a lane's 8 K
// values (8 voxels along x) are one aligned 16-byte read of dY; for X the
// tap's x shift (-1, 0, +1 voxel: 0, 2, 4 bytes past an aligned word) is a
// funnel shift of an aligned 16-byte and a 4-byte read.  Wavefront w keeps
// taps w, w + 4, ... for the whole run (27 taps) or K step w (1 tap) in
// registers and adds them to dW once.  With db, the workgroups of the first
// cin block also sum the staged dY tile into the bias gradient.
constexpr int WG_DS = 136;   // dY tile row (128 voxels + 8: rows 272 bytes apart, 4 banks)

__device__ __forceinline__ unsigned funnel2(unsigned hi, unsigned lo)
{
    return __builtin_amdgcn_alignbyte(hi, lo, 2);
}

template <int TAPS>
__global__ __launch_bounds__(256) void k_conv_wgrad_lds(const unsigned short *__restrict__ X,
                                                        const unsigned short *__restrict__ dY, int dy_stride,
                                                        ConvDims d, ConvTile T, int ntiles, int run, int RW, int CS,
                                                        float *__restrict__ dW, float *__restrict__ db)
{
    constexpr int R = TAPS == 27 ? 1 : 0;
    constexpr int NT = TAPS == 27 ? 7 : 1;   // taps per wavefront
    extern __shared__ unsigned short lds16[];
    unsigned short *XT = lds16;              // [32 ci][hz][hy][RW]
    unsigned short *DT = lds16 + 32 * CS;    // [32 co][WG_DS]
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int col = lane & 15, q = lane >> 4;
    const int co0 = blockIdx.y * 32, ci0 = blockIdx.z * 32;
    const int hy = T.ty + 2 * R, hz = T.tz + 2 * R, hx = T.tx + 2 * R;
    const int HV = hx * hy * hz;
    const int nvox = T.tx * T.ty * T.tz;
    const int ngx = min(4, (d.cin_p - ci0) >> 3);
    // this lane's voxel groups (8 along x) of the four K steps
    int vb[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int tv = 32 * k + 8 * q;
        vb[k] = 0;
        if (tv < nvox) {
            const int ltx = tv % T.tx, lty = (tv / T.tx) % T.ty, ltz = tv / (T.tx * T.ty);
            vb[k] = (ltz * hy + lty) * RW + ltx;
        }
    }
    f32x4 acc[NT][2][2];
#pragma unroll
    for (int i = 0; i < NT; ++i)
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[i][r][j] = f32x4{0, 0, 0, 0};
    float bsum = 0.0f;
    const int t_end = min(ntiles, (int)(blockIdx.x + 1) * run);
    for (int tile = blockIdx.x * run; tile < t_end; ++tile) {
        const int bx = tile % T.ntx, by = (tile / T.ntx) % T.nty, bz = tile / (T.ntx * T.nty);
        const int x0 = bx * T.tx, y0 = by * T.ty, z0 = bz * T.tz;
        const int zs0 = (z0 / T.Ds) * T.Ds;
        __syncthreads();   // the previous tile's reads are done
        // X halo, transposed: 16-byte loads of 8 channels, 8 two-byte LDS stores
        for (int i = threadIdx.x; i < HV * ngx; i += 256) {
            const int g = i % ngx, hv = i / ngx;
            const int hxi = hv % hx, hyi = (hv / hx) % hy, hzi = hv / (hx * hy);
            const int gx = x0 + hxi - R, gy = y0 + hyi - R, gz = z0 + hzi - R;
            u16x8 val = {0, 0, 0, 0, 0, 0, 0, 0};
            if (gx >= 0 && gx < d.W && gy >= 0 && gy < d.H && gz >= zs0 && gz < zs0 + T.Ds)
                val = *reinterpret_cast<const u16x8 *>(X + ((int64_t)(gz * d.H + gy) * d.W + gx) * d.cin_p + ci0 +
                                                       8 * g);
            const int pos = (hzi * hy + hyi) * RW + hxi;
#pragma unroll
            for (int j = 0; j < 8; ++j) XT[(8 * g + j) * CS + pos] = val[j];
        }
        // dY tile, transposed (zero outside the volume, past the tile and for co >= cout)
        for (int i = threadIdx.x; i < 128 * 4; i += 256) {
            const int g = i & 3, tv = i >> 2;
            u16x8 val = {0, 0, 0, 0, 0, 0, 0, 0};
            const int gx = x0 + tv % T.tx, gy = y0 + (tv / T.tx) % T.ty, gz = z0 + tv / (T.tx * T.ty);
            const int c = co0 + 8 * g;
            if (tv < nvox && gx < d.W && gy < d.H && c < dy_stride) {
                val = *reinterpret_cast<const u16x8 *>(dY + ((int64_t)(gz * d.H + gy) * d.W + gx) * dy_stride + c);
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    if (c + j >= d.cout) val[j] = 0;
            }
#pragma unroll
            for (int j = 0; j < 8; ++j) DT[(8 * g + j) * WG_DS + tv] = val[j];
        }
        __syncthreads();
        if (db && blockIdx.z == 0) {
            const int c = threadIdx.x >> 3, part = threadIdx.x & 7;
#pragma unroll
            for (int v = 0; v < 16; ++v) bsum += bf2f(DT[c * WG_DS + part * 16 + v]);
        }
        // A fragments (dY) of the four K steps, shared by every tap
        u16x8 a[4][2];
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int r = 0; r < 2; ++r)
                a[k][r] = *reinterpret_cast<const u16x8 *>(DT + (r * 16 + col) * WG_DS + 32 * k + 8 * q);
#pragma unroll
        for (int i = 0; i < NT; ++i) {
            const int t = TAPS == 27 ? wave + 4 * i : 0;
            if (t >= TAPS) continue;
            int shift = 0, tdelta = 0;
            if (TAPS == 27) {
                shift = t % 3;   // dx + 1
                tdelta = (t / 9) * hy * RW + ((t / 3) % 3) * RW;
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (TAPS == 1 && k != wave) continue;
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    const unsigned short *p = XT + (j * 16 + col) * CS + vb[k] + tdelta;
                    uint4 w = *reinterpret_cast<const uint4 *>(p);
                    if (shift) {
                        const unsigned nx = *reinterpret_cast<const unsigned *>(p + 8);
                        if (shift == 2) {
                            w = uint4{w.y, w.z, w.w, nx};
                        } else {
                            w = uint4{funnel2(w.y, w.x), funnel2(w.z, w.y), funnel2(w.w, w.z), funnel2(nx, w.w)};
                        }
                    }
                    const bf16x8 b = __builtin_bit_cast(bf16x8, w);
#pragma unroll
                    for (int r = 0; r < 2; ++r)
                        acc[i][r][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a[k][r]), b,
                                                                               acc[i][r][j], 0, 0, 0);
                }
            }
        }
    }
    // C layout: lane l holds column (ci) l % 16, rows (co) 4 * (l / 16) + e
#pragma unroll
    for (int i = 0; i < NT; ++i) {
        const int t = TAPS == 27 ? wave + 4 * i : 0;
        if (t >= TAPS) continue;
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int co = co0 + r * 16 + 4 * q + e, ci = ci0 + j * 16 + col;
                    const float v = acc[i][r][j][e];
                    if (co < d.cout && ci < d.cin_p && v != 0.0f)
                        atomicAdd(dW + ((int64_t)t * d.cout_p + co) * d.cin_p + ci, v);
                }
    }
    if (db && blockIdx.z == 0) {
        bsum += __shfl_xor(bsum, 1, 8);
        bsum += __shfl_xor(bsum, 2, 8);
        bsum += __shfl_xor(bsum, 4, 8);
        const int c = co0 + (threadIdx.x >> 3);
        if ((threadIdx.x & 7) == 0 && c < d.cout && bsum != 0.0f) atomicAdd(db + c, bsum);
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

static int conv_launch(ConvIn in, const void *w, const float *bias, int32_t N, int32_t D, int32_t H, int32_t W,
                       int32_t cin_p, int32_t cout, int32_t cout_p, int32_t taps, int32_t out_stride, int32_t out_off,
                       float slope, void *y_bf16, float *y_f32, void *stream, ConvHead hd = ConvHead{nullptr, nullptr,
                                                                                                     nullptr})
{
    if (!in.fine || !w || N <= 0 || D <= 0 || H <= 0 || W <= 0) return SRF_ERR_INVALID_ARG;
    if (cin_p <= 0 || cin_p % 8 || cout <= 0 || cout_p % 32 || cout > cout_p) return SRF_ERR_INVALID_ARG;
    if (taps != 1 && taps != 27) return SRF_ERR_INVALID_ARG;
    if (out_off < 0 || out_off + cout > out_stride || (!y_bf16 && !y_f32)) return SRF_ERR_INVALID_ARG;
    if ((reinterpret_cast<uintptr_t>(in.fine) | reinterpret_cast<uintptr_t>(in.coarse) |
         reinterpret_cast<uintptr_t>(w)) & 15)
        return SRF_ERR_INVALID_ARG;
    if ((int64_t)N * D * H * W * (cin_p > out_stride ? cin_p : out_stride) >= ((int64_t)1 << 31))
        return SRF_ERR_UNSUPPORTED;
    ConvDims d{N * D, H, W, cin_p, cout, cout_p, out_stride, out_off, taps, slope};
    const int R = taps == 27 ? 1 : 0;
    ConvTile T;
    conv_tile(D, H, W, R, T);
    T.nblk = T.ntx * T.nty * (N * D / T.tz);
    const int nb = (cout + 15) / 16;
    const int nbw = (nb >= 4 && cout_p % 64 == 0) ? 4 : nb >= 2 ? 2 : 1;
    const int lg = cin_p >= 32 ? 2 : cin_p >= 16 ? 1 : 0;
    size_t lds = (size_t)(T.tx + 2 * R) * (T.ty + 2 * R) * (T.tz + 2 * R) * ((1 << lg) | 1) * 16;
    const size_t lds_epi = (size_t)4 * (32 * (16 * nbw + 1) + 64) * sizeof(float);
    if (lds < lds_epi) lds = lds_epi;
    // a multiple of 8 workgroups (the XCD-contiguous tile order; extra ones exit)
    dim3 grid((unsigned)((T.nblk + 7) / 8 * 8), (unsigned)((nb + nbw - 1) / nbw));
    if (hd.out && (grid.y != 1 || !hd.w || !hd.b)) return SRF_ERR_UNSUPPORTED;   // the head needs every channel
    auto *ws = (const unsigned short *)w;
    auto *yb = (unsigned short *)y_bf16;
    auto st = (hipStream_t)stream;
#define SRF_CONV(TP, NBW) \
    hipLaunchKernelGGL((k_conv<TP, NBW>), grid, dim3(256), lds, st, in, ws, bias, d, T, yb, y_f32, hd)
    if (taps == 27) {
        if (nbw == 4) SRF_CONV(27, 4);
        else if (nbw == 2) SRF_CONV(27, 2);
        else SRF_CONV(27, 1);
    } else {
        if (nbw == 4) SRF_CONV(1, 4);
        else if (nbw == 2) SRF_CONV(1, 2);
        else SRF_CONV(1, 1);
    }
#undef SRF_CONV
    return launch_status();
}

extern "C" {

int srf_cnn_conv_batch(const void *x, const void *w, const float *bias, int32_t N, int32_t D, int32_t H, int32_t W,
                       int32_t cin_p, int32_t cout, int32_t cout_p, int32_t taps, int32_t out_stride, int32_t out_off,
                       float slope, void *y_bf16, float *y_f32, void *stream)
{
    return conv_launch(ConvIn{(const unsigned short *)x, nullptr, 0}, w, bias, N, D, H, W, cin_p, cout, cout_p, taps,
                       out_stride, out_off, slope, y_bf16, y_f32, stream);
}

int srf_cnn_conv_head_batch(const void *x, const void *w, const float *bias, int32_t N, int32_t D, int32_t H,
                            int32_t W, int32_t cin_p, int32_t cout, int32_t cout_p, int32_t taps, float slope,
                            void *y_bf16, const void *head_w, const float *head_b, float *head_out, void *stream)
{
    if (!y_bf16 || !head_out || !head_w || !head_b) return SRF_ERR_INVALID_ARG;
    return conv_launch(ConvIn{(const unsigned short *)x, nullptr, 0}, w, bias, N, D, H, W, cin_p, cout, cout_p, taps,
                       cout, 0, slope, y_bf16, nullptr, stream,
                       ConvHead{(const unsigned short *)head_w, head_b, head_out});
}

int srf_cnn_conv_up_batch(const void *fine, const void *coarse, int32_t cc, const void *w, const float *bias,
                          int32_t N, int32_t D, int32_t H, int32_t W, int32_t cin_p, int32_t cout, int32_t cout_p,
                          int32_t taps, int32_t out_stride, int32_t out_off, float slope, void *y_bf16, float *y_f32,
                          void *stream)
{
    if (!coarse || cc <= 0 || cc % 8 || cc >= cin_p || D % 2 || H % 2 || W % 2) return SRF_ERR_INVALID_ARG;
    return conv_launch(ConvIn{(const unsigned short *)fine, (const unsigned short *)coarse, cc}, w, bias, N, D, H, W,
                       cin_p, cout, cout_p, taps, out_stride, out_off, slope, y_bf16, y_f32, stream);
}

int srf_cnn_conv(const void *x, const void *w, const float *bias, int32_t D, int32_t H, int32_t W, int32_t cin_p,
                 int32_t cout, int32_t cout_p, int32_t taps, int32_t out_stride, int32_t out_off, float slope,
                 void *y_bf16, float *y_f32, void *stream)
{
    return srf_cnn_conv_batch(x, w, bias, 1, D, H, W, cin_p, cout, cout_p, taps, out_stride, out_off, slope, y_bf16,
                              y_f32, stream);
}

int srf_cnn_conv_wgrad_batch(const void *x, const void *dy, int32_t dy_stride, int32_t N, int32_t D, int32_t H,
                             int32_t W, int32_t cin_p, int32_t cout, int32_t cout_p, int32_t taps, float *dw,
                             float *db, void *stream)
{
    if (!x || !dy || !dw || N <= 0 || D <= 0 || H <= 0 || W <= 0) return SRF_ERR_INVALID_ARG;
    if (cin_p <= 0 || cin_p % 8 || cout <= 0 || cout_p % 32 || cout > cout_p || dy_stride < cout)
        return SRF_ERR_INVALID_ARG;
    if (taps != 1 && taps != 27) return SRF_ERR_INVALID_ARG;
    if ((int64_t)N * D * H * W * (cin_p > dy_stride ? cin_p : dy_stride) >= ((int64_t)1 << 31))
        return SRF_ERR_UNSUPPORTED;
    ConvDims d{N * D, H, W, cin_p, cout, cout_p, 0, 0, taps, -1.0f};
    auto st = (hipStream_t)stream;
    const int R = taps == 27 ? 1 : 0;
    ConvTile T;
    conv_tile(D, H, W, R, T);
    // the LDS path: x tiles of whole 8-voxel groups, 16-byte aligned rows
    const int RW = (T.tx + 2 * R + 7) / 8 * 8;
    int CS = 0;
    bool lds_ok = T.tx % 8 == 0 && dy_stride % 8 == 0 &&
                  !((reinterpret_cast<uintptr_t>(x) | reinterpret_cast<uintptr_t>(dy)) & 15);
    if (lds_ok) {
        const int hv = RW * (T.ty + 2 * R) * (T.tz + 2 * R);
        CS = (hv + 119) / 128 * 128 + 8;   // >= hv; channel rows 4 banks apart
        lds_ok = (size_t)64 * (CS + WG_DS) <= CONV_LDS_MAX;
    }
    if (lds_ok) {
        const int ntiles = T.ntx * T.nty * (N * D / T.tz);
        const int nblk = (cout_p / 32) * ((cin_p + 31) / 32);
        // about 1024 workgroups, each adding its taps to dW once per run of tiles
        int run = (int)(((int64_t)ntiles * nblk + 1023) / 1024);
        run = run < 1 ? 1 : run > 64 ? 64 : run;
        dim3 grid((unsigned)((ntiles + run - 1) / run), (unsigned)(cout_p / 32), (unsigned)((cin_p + 31) / 32));
        const size_t lds = (size_t)64 * (CS + WG_DS);
        if (taps == 27)
            hipLaunchKernelGGL(k_conv_wgrad_lds<27>, grid, dim3(256), lds, st, (const unsigned short *)x,
                               (const unsigned short *)dy, dy_stride, d, T, ntiles, run, RW, CS, dw, db);
        else
            hipLaunchKernelGGL(k_conv_wgrad_lds<1>, grid, dim3(256), lds, st, (const unsigned short *)x,
                               (const unsigned short *)dy, dy_stride, d, T, ntiles, run, RW, CS, dw, db);
        return launch_status();
    }
    const int64_t V = (int64_t)N * D * H * W;
    const int chunk = 4096;
    dim3 grid((unsigned)((V + chunk - 1) / chunk), (unsigned)taps, (unsigned)((cout_p / 32) * ((cin_p + 31) / 32)));
    hipLaunchKernelGGL(k_conv_wgrad, grid, dim3(256), 0, st, (const unsigned short *)x, (const unsigned short *)dy,
                       dy_stride, d, D, chunk, dw);
    int rc = launch_status();
    if (rc == SRF_OK && db)
        hipLaunchKernelGGL(k_bias_grad, dim3((unsigned)cout), dim3(256), 0, st, (const unsigned short *)dy, V,
                           dy_stride, db);
    return rc == SRF_OK ? launch_status() : rc;
}

int srf_cnn_conv_wgrad(const void *x, const void *dy, int32_t dy_stride, int32_t D, int32_t H, int32_t W,
                       int32_t cin_p, int32_t cout, int32_t cout_p, int32_t taps, float *dw, void *stream)
{
    return srf_cnn_conv_wgrad_batch(x, dy, dy_stride, 1, D, H, W, cin_p, cout, cout_p, taps, dw, nullptr, stream);
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
