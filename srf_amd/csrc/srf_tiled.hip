// srf_tiled.hip -- LDS-tiled projector / backprojector over a binned event list
// ("plan"), and its C-ABI entry points (include/srf_hip.h, srf_plan_*).
//
// Why: with one event per lane every projector term is a gather that misses
// the per-XCD L2 (a 256x256x128 image is 32 MiB, an XCD's L2 4 MiB) and every
// backprojector term is a device-scope atomic.  Here the image is cut into
// TX x TY x TZ = 32 x 32 x 8 voxel tiles and each event's slab range is
// cut into segments, one per tile it can touch.  A work item is a run of
// segments of one tile: the projector stages the tile in LDS once and every
// term is an LDS read; the backprojector accumulates into an LDS tile (ds_add)
// and writes it out once, with no global atomics.
//
// Binning (once per event list, reused by every iteration and both directions):
//   k_bin_count  segments per event and per (tile, length class) key
//   scan         event offsets (hipcub) and key offsets (host, <= 16384 keys)
//   k_bin_emit   48-byte segment records, grouped by key: records of one tile
//                are contiguous and sorted into slab-count classes, then (radix
//                sort) by branch and slab count, so a wavefront runs one
//                branch's marcher over similar lengths.
// A record carries everything the marcher needs (the event's exact branch
// parameters), so the passes stream records instead of gathering event data.
//
// Exactness: segments are a superset (value-space margin 1e-2 voxel, far above
// float round-off) of the slabs whose terms fall in the tile; inside, march_tile runs the
// reference's exact per-slab arithmetic and the sink keeps only the terms whose
// cell lies in the tile, so every term of every event is applied exactly once.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <limits.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "marcher.hpp"
#include "srf_hip.h"
#include "srf_internal.hpp"

namespace srf {

// 32 x 32 x 8 tiles: the backprojector's 64-bit fixed-point tile is 64 KiB, so
// two 1024-thread workgroups fit a CU's LDS (a 16-deep tile, 128 KiB, allowed
// one: 1.45x slower backprojection at the OSEM workload for 15% more segments).
#ifndef SRF_TILE_Z
#define SRF_TILE_Z 8
#endif
constexpr int TX = 32, TY = 32, TZ = SRF_TILE_Z;
constexpr int TVOX = TX * TY * TZ;      // 8192 voxels
// LDS layout of a tile: cell (lx, ly, lz) at word lx + LSY*ly + LSZ*lz.  The
// padded strides put a cell in bank (lx + ly + lz) mod 32, so the cells one
// primary step reaches (a plane of fixed lx for x-dominant rays, of fixed ly
// for y-dominant ones) spread over the banks instead of sharing one.
constexpr int LSY = TX + 1, LSZ = LSY * TY + 1;
constexpr int TPAD = LSZ * (TZ - 1) + LSY * (TY - 1) + TX;   // 8454 words: 33 KiB (float) / 66 KiB (fixed point)
__device__ __forceinline__ int tile_word(int i) { return (i & (TX - 1)) + LSY * ((i / TX) & (TY - 1)) + LSZ * (i / (TX * TY)); }
// The projector's tile carries a one-cell zero halo on every side: cell
// (lx, ly, lz), -1 <= l <= T, at word (lx + 1) + HSY (ly + 1) + HSZ (lz + 1).
// A gathered cell is clamped into [-1, in-image extent] per axis instead of
// being tested against the tile: every cell outside the tile's in-image part
// then reads a zero (its term belongs to another tile's segment or to none).
// Odd strides keep the planes of one primary step spread over the banks.
constexpr int HX = TX + 2, HY = TY + 2, HZ = TZ + 2;
constexpr int HSY = HX + 1, HSZ = HSY * HY + 1;
constexpr int HPAD = HSZ * (HZ - 1) + HSY * (HY - 1) + HX;   // 11908 words: 47 KiB
constexpr int NCLASS = 4;               // slab-count classes per tile
constexpr int ITEM = 8192;              // records per work item (at most)
constexpr int MAX_KEYS = 16384;         // LDS histogram of the binning kernels (64 KiB): up to 4096 tiles
constexpr int BIN_THREADS = 256;
constexpr int TILE_THREADS = 1024;

struct TileCfg {
    int ntx, nty, ntz, ntiles, nkeys;
};

// One segment: 48 bytes, read as three 16-byte loads.  Its event id is kept
// beside the records (rec_ev), where only the backprojector's value gathers
// and the digest read it.
struct __align__(16) SegRec {
    float p1, sec1, ks, z1;
    float kz, Wbp, Wfp;
    float tA, tB;            // TOF: exp(-d^2 / (2 tse)) == exp2(-(pp1 * tA + tB)^2), pp1 = ip - np2
    int slot;
    uint32_t ab_bp, ab_fp;   // slab range a | x-dominant branch << 15 | b << 16 (backprojector / projector)
};
static_assert(sizeof(SegRec) == 48, "record is three 16-byte words");

struct Item {
    int start, count, tile, cls;   // cls: slab-count class of the item's key
    float wmax_bp, wmax_fp;   // largest weight constant of the item's records, per mode
    float pad2[2];
};

// ------------------------------------------------------------ enumeration --

struct SegCtx {
    bool fp_use, bp_use;   // this branch serves the projector / backprojector
    int fa, fb;            // projector's exact cut range on this branch
};

// Calls f(tile, cls, a, b) for every segment of slabs [lo, hi) of branch B.
template <class F>
__device__ __forceinline__ void for_each_segment(const Branch &B, const RaySetup &r, const LaunchGeom &G,
                                                 const TileCfg &T, int lo, int hi, F &&f)
{
    const int TP = B.xdom ? TX : TY;
    const int TS = B.xdom ? TY : TX;
    const int nts = B.xdom ? T.nty : T.ntx;
    const double eps = 1e-2;
    const double p1 = B.p1, np2 = B.np2;
    for (int c = lo / TP; c * TP < hi; ++c) {
        const int ca = max(lo, c * TP), cb = min(hi, (c + 1) * TP);
        if (ca >= cb) continue;
        const double xa = (double)ca - np2, xb = (double)cb - np2;
        const double Sa = ((double)B.sec1 + (double)B.ks * (xa - p1)) / (double)B.ds + (double)B.ns2;
        const double Sb = ((double)B.sec1 + (double)B.ks * (xb - p1)) / (double)B.ds + (double)B.ns2;
        const double Za = ((double)r.z1 + (double)B.kz * (xa - p1)) / (double)G.dz + (double)G.nz2;
        const double Zb = ((double)r.z1 + (double)B.kz * (xb - p1)) / (double)G.dz + (double)G.nz2;
        const int ts0 = max(0, dbl_to_int(floor((fmin(Sa, Sb) - eps) / TS)));
        const int ts1 = min(nts - 1, dbl_to_int(floor((fmax(Sa, Sb) + eps) / TS)));
        const int tz0 = max(0, dbl_to_int(floor((fmin(Za, Zb) - eps) / TZ)));
        const int tz1 = min(T.ntz - 1, dbl_to_int(floor((fmax(Za, Zb) + eps) / TZ)));
        for (int ts = ts0; ts <= ts1; ++ts) {
            double wa, wb;
            axis_window((double)B.sec1, (double)B.ks, p1, (double)B.ds, (double)B.ns2, (double)(ts * TS) - eps,
                        (double)((ts + 1) * TS) + eps, wa, wb);
            if (wa > wb) continue;
            // slabs [ip - np2, ip - np2 + 1] meeting [wa, wb]
            const int sa = max(ca, dbl_to_int(ceil(wa + np2)) - 1);
            const int sb = min(cb, dbl_to_int(floor(wb + np2)) + 1);
            if (sa >= sb) continue;
            for (int tz = tz0; tz <= tz1; ++tz) {
                double za, zb;
                axis_window((double)r.z1, (double)B.kz, p1, (double)G.dz, (double)G.nz2, (double)(tz * TZ) - eps,
                            (double)((tz + 1) * TZ) + eps, za, zb);
                if (za > zb) continue;
                const int a = max(sa, dbl_to_int(ceil(za + np2)) - 1);
                const int b = min(sb, dbl_to_int(floor(zb + np2)) + 1);
                if (a >= b) continue;
                const int tile = B.xdom ? c + T.ntx * (ts + T.nty * tz) : ts + T.ntx * (c + T.nty * tz);
                const int cls = min(NCLASS - 1, ((b - a - 1) * NCLASS) / TP);
                f(tile, cls, a, b);
            }
        }
    }
}

// Every segment of one event: f(tile, cls, a, b, Branch, SegCtx).  Non-TOF
// ties (|xd| == |yd|) march different branches in the projector (>=,
// siddon.cu.cc:46) and the backprojector (>, :372): both are binned, each
// marked for its own direction.
template <bool TOF>
__device__ __forceinline__ void tof_coeffs_d(const Branch &B, const RaySetup &r, const LaunchGeom &G, double &tA,
                                             double &tB);

// TOF window (twin > 0): of the slabs [lo, hi), keep those whose Gaussian
// factor exp2(-(pp1 tA + tB)^2) is at least 2^-twin of the largest factor in
// [lo, hi) -- the peak inside the image, not the Gaussian's own peak, which
// lies outside for events whose TOF centre is outside the image: the EM ratio
// y / sum(terms) spreads such an event over its tail terms, so they must stay
// -- plus one slab of margin on each side for the rounding of the float
// coefficients.  u = pp1 tA + tB is linear in the slab, so its smallest
// magnitude over the range is 0 if it changes sign there, else the smaller
// magnitude at the two ends (over the continuous range: a lower bound, so the
// kept set is a superset).
template <bool TOF>
__device__ __forceinline__ void tof_window(const Branch &B, const RaySetup &r, const LaunchGeom &G, float twin,
                                           int &lo, int &hi)
{
    if (!TOF || !(twin > 0.0f) || lo >= hi) return;
    double tA, tB;
    tof_coeffs_d<TOF>(B, r, G, tA, tB);
    if (!(fabs(tA) > 0.0) || !isfinite(tA) || !isfinite(tB)) return;
    const double ua = ((double)lo - (double)B.np2) * tA + tB;
    const double ub = ((double)(hi - 1) - (double)B.np2) * tA + tB;
    const double umin = (ua <= 0.0) != (ub <= 0.0) ? 0.0 : fmin(fabs(ua), fabs(ub));
    const double D = sqrt(umin * umin + (double)twin);
    const double u1 = (-D - tB) / tA, u2 = (D - tB) / tA;
    const double np2 = B.np2;
    lo = max(lo, dbl_to_int(ceil(fmin(u1, u2) + np2)) - 1);
    hi = min(hi, dbl_to_int(floor(fmax(u1, u2) + np2)) + 2);
}

template <bool TOF, class F>
__device__ __forceinline__ void for_each_event_segment(const RaySetup &r, const LaunchGeom &G, const TileCfg &T,
                                                       float twin, F &&f)
{
    if (!r.live) return;
    const bool fp_live = TOF || !r.q2;
    const bool xfp = TOF ? r.xdom_gt : r.xdom_ge;
    const bool xbp = r.xdom_gt;
    for (int pass = 0; pass < 2; ++pass) {
        bool xdom;
        SegCtx ctx;
        if (pass == 0) {
            xdom = xbp;
            ctx.bp_use = true;
            ctx.fp_use = fp_live && (xfp == xbp);
        } else {
            if (xfp == xbp || !fp_live) break;
            xdom = xfp;
            ctx.bp_use = false;
            ctx.fp_use = true;
        }
        Branch B;
        make_branch(r, xdom, G, B);
        int lo, hi;
        clip_range(B, r, G, lo, hi);
        ctx.fa = lo;
        ctx.fb = hi;
        if (ctx.fp_use) fp_cut_range<TOF>(B, lo, hi, ctx.fa, ctx.fb);
        tof_window<TOF>(B, r, G, twin, lo, hi);
        tof_window<TOF>(B, r, G, twin, ctx.fa, ctx.fb);
        int elo = ctx.bp_use ? lo : ctx.fa, ehi = ctx.bp_use ? hi : ctx.fb;
        for_each_segment(B, r, G, T, elo, ehi, [&](int tile, int cls, int a, int b) { f(tile, cls, a, b, B, ctx); });
    }
}

// Slab range [a, b) of a record and its branch (slab indices below 2^15).
__device__ __forceinline__ uint32_t pack_ab(int a, int b, bool xd)
{
    return (uint32_t)a | ((uint32_t)xd << 15) | ((uint32_t)b << 16);
}
__device__ __forceinline__ int ab_a(uint32_t ab) { return (int)(ab & 0x7fffu); }
__device__ __forceinline__ int ab_b(uint32_t ab) { return (int)(ab >> 16); }
__device__ __forceinline__ bool ab_xd(uint32_t ab) { return (ab >> 15) & 1u; }

// TOF Gaussian of siddon_tof.cu.cc:55-59 as a linear function of the slab's
// lower boundary pp1 = ip - np2 (exact in float):
//   pc = ((ip - np/2) + 0.5) * pd_tof = (pp1 + off) * pd_tof, off = np2 - np/2 + 0.5
//   d  = ((pc - p1) / (p2 - p1) - ratio1) * L
//   exp(-d^2 / (2 tse)) = exp2(-(d * sqrt(log2(e) / (2 tse)))^2)
// Coefficients in double, rounded once.
template <bool TOF>
__device__ __forceinline__ void tof_coeffs_d(const Branch &B, const RaySetup &r, const LaunchGeom &G, double &tA,
                                             double &tB)
{
    tA = tB = 0.0;
    if (!TOF) return;
    const double off = (double)B.np2 - (double)(B.np / 2) + 0.5;
    const double s = sqrt(0.5 * 1.4426950408889634 / (double)G.tse);
    const double k = (double)r.L / ((double)B.p2 - (double)B.p1);
    tA = (double)B.pd_tof * k * s;
    tB = ((off * (double)B.pd_tof - (double)B.p1) * k - (double)r.ratio1 * (double)r.L) * s;
}

template <bool TOF>
__device__ __forceinline__ void tof_coeffs(const Branch &B, const RaySetup &r, const LaunchGeom &G, float &tA,
                                           float &tB)
{
    double a, b;
    tof_coeffs_d<TOF>(B, r, G, a, b);
    tA = (float)a;
    tB = (float)b;
}

template <bool TOF>
__global__ __launch_bounds__(BIN_THREADS) void k_bin_count(const float *__restrict__ lors, int64_t n, int stride,
                                                           const float *__restrict__ tofv, int tof_stride,
                                                           LaunchGeom G, TileCfg T, float twin,
                                                           int *__restrict__ ev_cnt,
                                                           uint8_t *__restrict__ ev_flag,
                                                           unsigned *__restrict__ key_count)
{
    extern __shared__ unsigned hist[];
    for (int i = threadIdx.x; i < T.nkeys; i += BIN_THREADS) hist[i] = 0;
    __syncthreads();
    const int64_t e = (int64_t)blockIdx.x * BIN_THREADS + threadIdx.x;
    if (e < n) {
        RaySetup r;
        ray_setup<TOF>(lors + e * stride, TOF ? tofv[e * tof_stride] : 0.0f, G, r);
        int cnt = 0;
        for_each_event_segment<TOF>(r, G, T, twin, [&](int tile, int cls, int, int, const Branch &, const SegCtx &) {
            atomicAdd(&hist[tile * NCLASS + cls], 1u);
            ++cnt;
        });
        ev_cnt[e] = cnt;
        ev_flag[e] = r.nan ? 1 : 0;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < T.nkeys; i += BIN_THREADS)
        if (hist[i]) atomicAdd(&key_count[i], hist[i]);
}

// Each record is written at its segment slot (ev_off[e] + k: event order, a
// deterministic position) with its sort key; the sort that follows groups the
// records by key.
template <bool TOF>
__global__ __launch_bounds__(BIN_THREADS) void k_bin_emit(const float *__restrict__ lors, int64_t n, int stride,
                                                          const float *__restrict__ tofv, int tof_stride,
                                                          LaunchGeom G, TileCfg T, float twin, int prescale,
                                                          const int *__restrict__ ev_off,
                                                          SegRec *__restrict__ recs, int *__restrict__ slot_ev,
                                                          unsigned *__restrict__ rec_sk)
{
    const int64_t e = (int64_t)blockIdx.x * BIN_THREADS + threadIdx.x;
    if (e >= n) return;
    RaySetup r;
    ray_setup<TOF>(lors + e * stride, TOF ? tofv[e * tof_stride] : 0.0f, G, r);
    int j = ev_off[e];
    for_each_event_segment<TOF>(r, G, T, twin, [&](int tile, int cls, int a, int b, const Branch &B,
                                                   const SegCtx &c) {
        const int pos = j++;
        SegRec R;
        R.p1 = B.p1;
        // prescale bit 0 / 1: the secondary / axial voxel edge is a power of
        // two, and the record holds the line's constants divided by it (exact):
        // the marcher's (c + k (p - p1)) / d then equals c / d + (k / d) (p - p1)
        // bit for bit and runs without the division
        const float ss = (prescale & 1) ? 1.0f / B.ds : 1.0f, zs = (prescale & 2) ? 1.0f / G.dz : 1.0f;
        R.sec1 = B.sec1 * ss;
        R.ks = B.ks * ss;
        R.z1 = r.z1 * zs;
        R.kz = B.kz * zs;
        R.Wbp = branch_weight<TOF>(B, r, G, false);
        R.Wfp = branch_weight<TOF>(B, r, G, true);
        tof_coeffs<TOF>(B, r, G, R.tA, R.tB);
        R.slot = pos;
        R.ab_bp = c.bp_use ? pack_ab(a, b, B.xdom) : 0u;
        const int fa = max(a, c.fa), fb = min(b, c.fb);
        R.ab_fp = (c.fp_use && fa < fb) ? pack_ab(fa, fb, B.xdom) : 0u;
        recs[pos] = R;
        slot_ev[pos] = (int)e;
        // sort key: (tile, class), then the branch (a wavefront mostly runs
        // one branch's marcher), then slab count (longest first: its lanes
        // march similar lengths); a stable sort keeps slot order within a key
        const unsigned len = (unsigned)min(b - a, 63);
        rec_sk[pos] = ((unsigned)(tile * NCLASS + cls) << 7) | ((unsigned)B.xdom << 6) | (63u - len);
    });
}

// ------------------------------------------------------------ tile kernels --

// v_cvt_flr_i32_f32: floor and convert in one VALU op (NaN -> 0, saturating,
// as v_floor_f32 + v_cvt_i32_f32).
__device__ __forceinline__ int cvt_flr(float v)
{
    int r;
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(v));
    return r;
}

enum : int { M_GATHER = 0, M_SCATTER = 1, M_DIGEST = 2, M_SCATTER_FX = 3 };

// A term v * r (v already scaled by the item's 2^shift, |v * r| < 2^49 by the
// choice of shift) as a 64-bit two's complement integer, rounded to nearest.
// The float product is exact in double; adding 1.5 * 2^52 in the same fma
// rounds it to an integer (the sum's unit spacing is 1), and the sum's bit
// pattern is then 0x4338000000000000 + round(v * r) for |v * r| < 2^51.  The
// constant's low word is zero, so the subtraction touches the high word only.
__device__ __forceinline__ unsigned long long fixed_term(double v, float r)
{
    const double s = fma(v, (double)r, 6755399441055744.0);
    return (unsigned long long)__double_as_longlong(s) - 0x4338000000000000ull;
}

// med3(x, lo, hi) = min(max(x, lo), hi) for lo <= hi, one v_med3_i32 (the
// compiler emits it only for constant bounds), and a * b + c with a and b taken
// as signed 24-bit integers (v_mad_i32_i24: the clamped cell index is -1 on
// the low-edge halo of a tile at the image origin, which the unsigned form
// would read as 2^24 - 1).  Plain ALU instructions, no memory access.
__device__ __forceinline__ int med3_i32(int x, int lo, int hi)
{
    int r;
    asm("v_med3_i32 %0, %1, %2, %3" : "=v"(r) : "v"(x), "v"(lo), "v"(hi));
    return r;
}
__device__ __forceinline__ int mad_i24(int a, int b, int c)
{
    int r;
    asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

// Minimum / maximum of two non-negative finite floats as integer operations on
// their bit patterns (which order the same way): no NaN canonicalisation.
__device__ __forceinline__ float min_nonneg(float a, float b)
{
    return __uint_as_float(min(__float_as_uint(a), __float_as_uint(b)));
}
__device__ __forceinline__ float max_nonneg(float a, float b)
{
    return __uint_as_float(max(__float_as_uint(a), __float_as_uint(b)));
}

// State of one segment walk inside one LDS tile (layout: tile_word).  Terms
// whose cell lies outside the tile, or that the reference's predicates reject,
// are sent to a dummy address (see march_tile), so the slab body has no
// branches and its three LDS accesses issue back to back.
struct TileWalk {
    float *lds;
    unsigned long long *fx;  // fixed-point accumulator tile
    int dummy;
    int x0, y0, z0;    // tile origin
    float acc;         // gather result (times the segment's weight constant)
    float value;       // scatter value (for fixed point: times 2^shift)
    uint64_t h;        // digest
    int cnt;
};

// One segment: slabs [a, b) of record R inside the walk's tile.
//
// Per slab this evaluates the reference's cell indices and predicates exactly
// (siddon.cu.cc:86-189 for x-dominant rays, 231-333 for y-dominant; TOF
// siddon_tof.cu.cc:102-210, 244-366).  With cs1 <= cs2 the two secondary cells
// a slab spans and cz1 <= cz2 the axial ones, the reference's top-level test
// is cs1 >= (cs1 != cs2 ? -1 : 0) && cs1 < ns (and the same along z), term 2
// also needs its own cell below ns / nz, and term 3 both.  Restricted to the
// cells of the tile that lie inside the image (0 <= c - origin < extent, the
// extent cut at the image edge), every upper bound and every `>= 0` part is
// implied by the tile test of the term's own cell, which leaves:
//   term 1 (cs1, cz1): its tile test;
//   term 2 (cs1, cz2) if rs > rz, else (cs2, cz1): both crossings, its tile
//          test, and cz1 >= -1 (resp. cs1 >= -1);
//   term 3 (cs2, cz2): a crossing, its tile test, cs1 >= -1 and cz1 >= -1.
//
// XD selects the branch (x- or y-dominant ray), so the tile origin, strides and
// extents are wave-uniform; a work item's records are sorted by branch.
template <bool TOF, bool DS, bool DZ, int MODE, bool XD>
__device__ __forceinline__ void march_tile(const SegRec &R, const LaunchGeom &G, int a, int b, float W,
                                           TileWalk &t)
{
    const float p1 = R.p1, sec1 = R.sec1, ks = R.ks, kz = R.kz, z1 = R.z1;
    const float ds = XD ? G.dy : 1.0f;
    const float ns2 = XD ? G.ny2 : G.nx2;
    const float np2 = XD ? G.nx2 : G.ny2;
    const int ns = XD ? G.ny : G.nx, nz = G.nz;
    const int p0 = XD ? t.x0 : t.y0, s0 = XD ? t.y0 : t.x0, z0 = t.z0;
    constexpr bool HALO = MODE == M_GATHER;                // the projector's halo tile
    constexpr int RS = HALO ? HSY : LSY, ZS = HALO ? HSZ : LSZ;
    constexpr int SP = XD ? 1 : RS, SS = XD ? RS : 1;      // LDS strides of primary / secondary
    const bool ks_pos = ks >= 0.0f, kz_pos = kz >= 0.0f;
    // in-image extent of the tile along secondary / axial
    const unsigned es = (unsigned)min(TX, ns - s0), ez = (unsigned)min(TZ, nz - z0);
    float pp1 = (float)a - np2;
    float S0, Z0;
    {
        const float sv = sec1 + ks * (pp1 - p1);
        S0 = (DS ? sv / ds : sv) + ns2;
        const float zv = z1 + kz * (pp1 - p1);
        Z0 = (DZ ? zv / G.dz : zv) + G.nz2;
    }
    // LDS byte address of cell (c_s, c_z) of slab ip: base + c_s * BS + c_z *
    // BZ, the tile origin folded into base
    constexpr int ES = (MODE == M_SCATTER_FX) ? 8 : 4;
    constexpr unsigned BS = SS * ES, BZ = ZS * ES;    // byte strides of secondary / axial
    int base = HALO ? (a - p0 + 1) * SP * ES - (s0 - 1) * (int)BS - (z0 - 1) * (int)BZ
                    : (a - p0) * SP * ES - s0 * (int)BS - z0 * (int)BZ;
    constexpr int bstep = SP * ES;
    // halo clamp bounds and byte strides as registers (a VOP3 instruction reads
    // at most one scalar register on this target)
    const int slo = s0 - 1, shi = s0 + (int)es, zlo = z0 - 1, zhi = z0 + (int)ez;
    const int bs_v = (int)BS, bz_v = (int)BZ;
    // rejected terms: a zero word (gathers, never written), the lane's own
    // slot after the fixed-point tile (fixed-point scatters), or an address
    // no tile cell has (float scatters and the digest test for it)
    const int dummy = (MODE == M_SCATTER_FX) ? (t.dummy + (int)threadIdx.x) * ES : t.dummy * ES;
    static_assert(TX == 32 && TY == 32, "secondary stride is 1 or 32");
    float acc = 0.0f;
    const float vw = t.value * W;
    const double vwd = (double)vw;
    const int base_end = base + (b - a) * bstep;
    int ip = a;   // digest only
    // one slab; the loop below runs it twice per trip (the marcher's inline asm
    // is convergent, which rules out the compiler's own unrolling), so the
    // values carried from slab to slab alternate registers instead of moving
    auto slab = [&]() {
        const float pp2 = pp1 + 1.0f;
        float S1, Z1;
        {
            const float sv = sec1 + ks * (pp2 - p1);
            S1 = (DS ? sv / ds : sv) + ns2;
            const float zv = z1 + kz * (pp2 - p1);
            Z1 = (DZ ? zv / G.dz : zv) + G.nz2;
        }
        const float sl = ks_pos ? S0 : S1, sh = ks_pos ? S1 : S0;
        const float zl = kz_pos ? Z0 : Z1, zh = kz_pos ? Z1 : Z0;
        const int cs1 = cvt_flr(sl), cs2 = cvt_flr(sh);
        const int cz1 = cvt_flr(zl), cz2 = cvt_flr(zh);
        const bool bs = cs1 != cs2, bz = cz1 != cz2;
        // LDS addresses of the three terms (dummy where the term is not
        // applied); term 2 is i2a if rs > rz, else i2b.  (Signed 24-bit
        // multiply-adds: a clamped cell index lies in [origin - 1, origin +
        // extent], within [-1, 2^23), and a byte stride is below 2^13;
        // logical operators keep the predicates as lane masks.)
        const bool los = cs1 >= -1, loz = cz1 >= -1;
        int i1, i2a, i2b, i3;
        if (HALO) {
            // clamped into the halo: the tile tests become four med3, each
            // address two 24-bit multiply-adds
            const int c1 = med3_i32(cs1, slo, shi), c2 = med3_i32(cs2, slo, shi);
            const int d1 = med3_i32(cz1, zlo, zhi), d2 = med3_i32(cz2, zlo, zhi);
            const int A = mad_i24(c1, bs_v, base), B = mad_i24(c2, bs_v, base);
            // computed before the selects: inline asm is not speculated, so a
            // multiply-add inside a select would become a branch
            const int j3 = mad_i24(d2, bz_v, B), j2a = mad_i24(d2, bz_v, A), j2b = mad_i24(d1, bz_v, B);
            i1 = mad_i24(d1, bz_v, A);
            i3 = ((bs || bz) && los && loz) ? j3 : dummy;
            i2a = (bs && bz && loz) ? j2a : dummy;
            i2b = (bs && bz && los) ? j2b : dummy;
        } else {
            const bool ts1 = (unsigned)(cs1 - s0) < es, ts2 = (unsigned)(cs2 - s0) < es;
            const bool tz1 = (unsigned)(cz1 - z0) < ez, tz2 = (unsigned)(cz2 - z0) < ez;
            // six signed 24-bit multiply-adds, all computed before the
            // selects (an address is used only when its cells pass the tile
            // tests, so lie in [origin, origin + extent))
            const int A = mad_i24(cs1, bs_v, base), B = mad_i24(cs2, bs_v, base);
            const int j1 = mad_i24(cz1, bz_v, A), j3 = mad_i24(cz2, bz_v, B);
            const int j2a = mad_i24(cz2, bz_v, A), j2b = mad_i24(cz1, bz_v, B);
            i1 = (ts1 && tz1) ? j1 : dummy;
            i3 = ((bs || bz) && ts2 && tz2 && los && loz) ? j3 : dummy;
            i2a = (bs && bz && ts1 && tz2 && loz) ? j2a : dummy;
            i2b = (bs && bz && ts2 && tz1 && los) ? j2b : dummy;
        }
        // materialise the addresses here: sunk past the branch below, their
        // lane-mask operands would be carried across it as 0/1 vectors
        asm volatile("" ::"v"(i1), "v"(i2a), "v"(i2b), "v"(i3));
        // crossing fractions: reciprocal estimates, exact quotients only where
        // the estimate cannot decide rs > rz (which picks the middle voxel);
        // rs = 1 along an axis that is not crossed
        const float dsv = sh - sl, dzv = zh - zl;
        const float ns_ = (float)cs2 - sl, nz_ = (float)cz2 - zl;
        const float es_ = ns_ * __builtin_amdgcn_rcpf(dsv), ez_ = nz_ * __builtin_amdgcn_rcpf(dzv);
        float rs = bs ? es_ : 1.0f, rz = bz ? ez_ : 1.0f;
        float dr = rs - rz;
        // rs, rz lie in (0, 1]: an absolute margin is at least as strict as the
        // relative one the estimates' few-ulp error needs
        if (bs && bz && !(fabsf(dr) > 4e-6f)) {
            rs = ns_ / dsv;
            rz = nz_ / dzv;
            dr = rs - rz;
        }
        asm("" : "+v"(dr));   // keeps |dr| a source modifier of its uses
        const bool sgt = dr > 0.0f;   // rs > rz
        const int i2 = sgt ? i2a : i2b;
        // the three terms' fractions (the reference's branch tree, flattened):
        // r1 = min(rs, rz), r2 = |rs - rz| (both axes cross), r3 = 1 - m with m
        // the largest fraction of an axis that crosses: with one axis crossing
        // that is r1 (the other axis's fraction is 1), with none term 3 is not
        // applied
        const float r1 = min_nonneg(rs, rz);
        const float r2 = fabsf(dr);
        const float m = (bs && bz) ? max_nonneg(rs, rz) : r1;
        float e = 1.0f;
        if (TOF) {
            const float dd = fmaf(pp1, R.tA, R.tB);
            e = __builtin_amdgcn_exp2f(-(dd * dd));
            // the projector computes the weight before its gathers, so that the
            // exponential's latency is not exposed at the accumulation after
            // them
            if (MODE == M_GATHER) asm volatile("" ::"v"(e));
        }
        if (MODE == M_GATHER) {
            const char *L = reinterpret_cast<const char *>(t.lds);
            const float g1 = *reinterpret_cast<const float *>(L + i1);
            const float g2 = *reinterpret_cast<const float *>(L + i2);
            const float g3 = *reinterpret_cast<const float *>(L + i3);
            // g3 * (1 - m) as g3 - g3 * m
            const float sum = fmaf(g1, r1, fmaf(g2, r2, fmaf(-g3, m, g3)));
            acc = TOF ? fmaf(sum, e, acc) : acc + sum;
        } else if (MODE == M_SCATTER) {
            // float LDS atomics (non-finite values only; see k_bp_tiled)
            char *L = reinterpret_cast<char *>(t.lds);
            const float v = vw * e, r3 = 1.0f - m;
            if (i1 != dummy) atomicAdd(reinterpret_cast<float *>(L + i1), v * r1);
            if (i2 != dummy) atomicAdd(reinterpret_cast<float *>(L + i2), v * r2);
            if (i3 != dummy) atomicAdd(reinterpret_cast<float *>(L + i3), v * r3);
        } else if (MODE == M_SCATTER_FX) {
            // 64-bit fixed point: integer LDS atomics run at full rate.  Every
            // lane adds all three terms; rejected ones go to the lane's own
            // dummy slot (no per-term compare and exec-mask save / restore)
            char *L = reinterpret_cast<char *>(t.fx);
            const double v = TOF ? (double)(vw * e) : vwd;
            const float r3 = 1.0f - m;
            atomicAdd(reinterpret_cast<unsigned long long *>(L + i1), fixed_term(v, r1));
            atomicAdd(reinterpret_cast<unsigned long long *>(L + i2), fixed_term(v, r2));
            atomicAdd(reinterpret_cast<unsigned long long *>(L + i3), fixed_term(v, r3));
        } else {
            const int gsp = XD ? 1 : G.nx, gss = XD ? G.nx : 1, gsz = G.nx * G.ny;
            const int pg = ip * gsp;
            const int c2s = sgt ? cs1 : cs2, c2z = sgt ? cz2 : cz1;
            const int bb1 = bs ? (bz ? (sgt ? 5 : 8) : 3) : (bz ? 1 : 0);
            const int bb3 = bs ? (bz ? (sgt ? 7 : 10) : 4) : 2;
            if (i1 != dummy) {
                t.h += mix64(((uint64_t)(uint32_t)(pg + cs1 * gss + cz1 * gsz) << 4) | (uint64_t)bb1);
                t.cnt++;
            }
            if (i2 != dummy) {
                t.h += mix64(((uint64_t)(uint32_t)(pg + c2s * gss + c2z * gsz) << 4) | (uint64_t)(sgt ? 6 : 9));
                t.cnt++;
            }
            if (i3 != dummy) {
                t.h += mix64(((uint64_t)(uint32_t)(pg + cs2 * gss + cz2 * gsz) << 4) | (uint64_t)bb3);
                t.cnt++;
            }
            ++ip;
        }
        pp1 = pp2;
        S0 = S1;
        Z0 = Z1;
        base += bstep;
    };
    while (base < base_end) {
        slab();
        if (!(base < base_end)) break;
        slab();
    }
    if (MODE == M_GATHER) t.acc = acc * W;
}

__device__ __forceinline__ void tile_origin(int tile, const TileCfg &T, int &x0, int &y0, int &z0)
{
    x0 = (tile % T.ntx) * TX;
    y0 = ((tile / T.ntx) % T.nty) * TY;
    z0 = (tile / (T.ntx * T.nty)) * TZ;
}

// Both branches' marchers; a record's branch bit picks one (the records of a
// work item are sorted by branch, so a wavefront mostly runs one of them).
template <bool TOF, bool DS, bool DZ, int MODE>
__device__ __forceinline__ void march_any(const SegRec &R, const LaunchGeom &G, uint32_t ab, float W, TileWalk &t)
{
    if (ab_xd(ab)) march_tile<TOF, DS, DZ, MODE, true>(R, G, ab_a(ab), ab_b(ab), W, t);
    else march_tile<TOF, DS, DZ, MODE, false>(R, G, ab_a(ab), ab_b(ab), W, t);
}

__device__ __forceinline__ void walk_for(TileWalk &t, float *lds, int dummy, int x0, int y0, int z0)
{
    t.lds = lds;
    t.dummy = dummy;
    t.x0 = x0;
    t.y0 = y0;
    t.z0 = z0;
    t.acc = 0.0f;
    t.value = 0.0f;
    t.h = 0;
    t.cnt = 0;
}

// Occupancy.  A 1024-thread workgroup is 4 waves per SIMD and a SIMD holds at
// most 8 waves: 2 workgroups per CU, which the backprojector's 64 KiB
// fixed-point tile also allows.  That takes at most 64 VGPRs per wave (the
// batched backprojector would otherwise use more).
#define BP_OCCUPANCY __attribute__((amdgpu_waves_per_eu(8)))

// Projector LDS: the halo tile plus one zero word, the target of rejected terms.
constexpr int LDS_WORDS = HPAD + 1;
constexpr int STAGE_LOADS = (HX * HY * HZ + TILE_THREADS - 1) / TILE_THREADS;   // 12 per thread

// Projection of one work item: stage its tile in LDS, march every record,
// one partial sum per record.
template <bool TOF, bool DS, bool DZ>
__device__ __forceinline__ void fp_item(const SegRec *__restrict__ recs, const Item &it,
                                        const float *__restrict__ img, const LaunchGeom &G, const TileCfg &T,
                                        float *__restrict__ partial, float *tile)
{
    int x0, y0, z0;
    tile_origin(it.tile, T, x0, y0, z0);
    // every load of the staging is issued before the first LDS store (a
    // load-store loop would wait for each load in turn)
    float v[STAGE_LOADS];
#pragma unroll
    for (int j = 0; j < STAGE_LOADS; ++j) {
        const int i = (int)threadIdx.x + j * TILE_THREADS;
        const int hx = i % HX, hy = (i / HX) % HY, hz = i / (HX * HY);
        const int lx = hx - 1, ly = hy - 1, lz = hz - 1;
        const int gx = x0 + lx, gy = y0 + ly, gz = z0 + lz;
        const bool core = (unsigned)lx < (unsigned)TX && (unsigned)ly < (unsigned)TY && (unsigned)lz < (unsigned)TZ;
        v[j] = (i < HX * HY * HZ && core && gx < G.nx && gy < G.ny && gz < G.nz) ? img[gx + G.nx * (gy + G.ny * gz)]
                                                                               : 0.0f;
    }
#pragma unroll
    for (int j = 0; j < STAGE_LOADS; ++j) {
        const int i = (int)threadIdx.x + j * TILE_THREADS;
        const int hx = i % HX, hy = (i / HX) % HY, hz = i / (HX * HY);
        if (i < HX * HY * HZ) tile[hx + HSY * hy + HSZ * hz] = v[j];
    }
    if (threadIdx.x == 0) tile[HPAD] = 0.0f;
    __syncthreads();
    // the next record is loaded while this one is marched
    const int end = it.start + it.count;
    int r = it.start + threadIdx.x;
    SegRec Rn;
    if (r < end) Rn = recs[r];
    for (; r < end; r += TILE_THREADS) {
        const SegRec R = Rn;
        if (r + TILE_THREADS < end) Rn = recs[r + TILE_THREADS];
        TileWalk t;
        walk_for(t, tile, HPAD, x0, y0, z0);
        march_any<TOF, DS, DZ, M_GATHER>(R, G, R.ab_fp, R.Wfp, t);
        partial[R.slot] = t.acc;
    }
}

template <bool TOF, bool DS, bool DZ>
__global__ __launch_bounds__(TILE_THREADS) void k_fp_tiled(const SegRec *__restrict__ recs,
                                                           const Item *__restrict__ items,
                                                           const int *__restrict__ order,
                                                           const float *__restrict__ img, LaunchGeom G, TileCfg T,
                                                           float *__restrict__ partial)
{
    __shared__ float tile[LDS_WORDS];
    const Item it = items[order[blockIdx.x]];
    fp_item<TOF, DS, DZ>(recs, it, img, G, T, partial, tile);
}

__device__ __forceinline__ float event_value(const float *__restrict__ values, const float *__restrict__ y,
                                             const float *__restrict__ p, int kind, float eps, int e)
{
    if (y) {
        const float d = p[e] + eps;
        float v = y[e] / d;
        if (kind == SRF_RATIO_DIV_NO_NAN && d == 0.0f) v = 0.0f;
        return v;
    }
    return values ? values[e] : 1.0f;
}

// Backprojection of one work item into its LDS tile, written out once as float.
//
// Terms are accumulated in 64-bit fixed point: integer LDS atomics run at full
// rate where float LDS atomics are serialised per lane, and integer addition
// is associative, so the result is bit-reproducible.  The scale is chosen per
// item: a pre-pass (event ids streamed from rec_ev) keeps each lane's event
// values in registers, and the item's largest |value| times its largest weight
// constant W (precomputed per mode; W bounds every term weight of a segment)
// bounds every term, which is scaled below 2^49.  A segment puts at most one term on a voxel and an
// item has at most ITEM = 2^13 segments, so no voxel sum can reach 2^62.
// Items holding a non-finite value use float LDS atomics instead (NaN and inf
// then propagate as through the reference's atomicAdd).
//
// One scale resolves a term to 2^-50 of the item's largest; an item whose
// event values span more than 2^BAND_BITS (the EM ratio y / (p + eps) of an
// event whose projection is ~0 can be 1e8 times the others') would lose its
// small terms.  Its records are then accumulated in exponent bands: band b
// holds the values in [2^(e - 24 (b + 1)), 2^(e - 24 b)), e the exponent of
// the largest, each band with its own scale in its own pass over the records
// (bp_item_bands), and the bands' float tiles are summed in band order (still
// bit-reproducible).  Within a band a value resolves to 2^-25 of itself or
// better, float32's own precision.  An item within one band -- the ordinary
// case -- takes the single pass.
constexpr int FX_BITS = 49;
constexpr int BAND_BITS = 24;
constexpr int MAX_BANDS = 8;   // values below 2^-192 of the largest join the last band
constexpr int RECS_PER_THREAD = ITEM / TILE_THREADS;
static_assert(ITEM % TILE_THREADS == 0, "whole records per thread");

// Band of a nonzero value v for an item whose largest value is below 2^etop.
__device__ __forceinline__ int value_band(float v, int etop, int nb)
{
    int ev = 0;
    (void)frexpf(v, &ev);
    return min((etop - ev) / BAND_BITS, nb - 1);
}

// Shift that scales band `band` of an item (largest value vtop < 2^etop,
// weight constant wmax) below 2^FX_BITS.
__device__ __forceinline__ int band_shift(float vtop, float wmax, int etop, int band)
{
    float c = band == 0 ? vtop * wmax : ldexpf(wmax, etop - BAND_BITS * band);
    if (!isfinite(c)) c = 3.40282347e38f;
    int ex = 0;
    (void)frexpf(c, &ex);   // c < 2^ex
    return FX_BITS - ex;
}

// All bands of a wide-range item (largest value vtop < 2^etop, nb bands), each
// in its own pass over the records: band 0 written to `out`, the others added.
// Only such items run this; it is kept compact (records not unrolled, values
// read again) so that it needs fewer registers than bp_item's single pass.
template <bool TOF, bool DS, bool DZ>
__device__ __forceinline__ void bp_item_bands(const SegRec *__restrict__ recs, const Item &it,
                                              const float *__restrict__ values, const float *__restrict__ y,
                                              const float *__restrict__ p, int kind, float eps, int adjoint,
                                              const LaunchGeom &G, const TileCfg &T,
                                              const int *__restrict__ rec_ev, float *__restrict__ out,
                                              unsigned long long *fx, float vtop, int etop, int nb)
{
    float *fl = reinterpret_cast<float *>(fx);
    int x0, y0, z0;
    tile_origin(it.tile, T, x0, y0, z0);
    const float wmax = adjoint ? it.wmax_fp : it.wmax_bp;
    const int end = it.start + it.count;
    for (int band = 0; band < nb; ++band) {
        const int shift = band_shift(vtop, wmax, etop, band);
        if (band > 0) {
            __syncthreads();   // the previous band's tile has been written out
            for (int i = threadIdx.x; i < TPAD; i += TILE_THREADS) fx[i] = 0ull;
            __syncthreads();
        }
#pragma unroll 1
        for (int r = it.start + threadIdx.x; r < end; r += TILE_THREADS) {
            const float v = event_value(values, y, p, kind, eps, rec_ev[r]);
            if (v == 0.0f || value_band(v, etop, nb) != band) continue;
            const SegRec R = recs[r];
            const uint32_t ab = adjoint ? R.ab_fp : R.ab_bp;
            if (ab_a(ab) >= ab_b(ab)) continue;
            TileWalk t;
            walk_for(t, fl, TPAD, x0, y0, z0);
            t.fx = fx;
            t.value = ldexpf(v, shift);
            march_any<TOF, DS, DZ, M_SCATTER_FX>(R, G, ab, adjoint ? R.Wfp : R.Wbp, t);
        }
        __syncthreads();
        for (int i = threadIdx.x; i < TVOX; i += TILE_THREADS) {
            const float bv = ldexpf((float)(long long)fx[tile_word(i)], -shift);
            out[i] = band == 0 ? bv : out[i] + bv;
        }
    }
}

// Backprojection of one work item into the LDS tile fx (shared reductions in
// s_vmax / s_vmin / s_bad), written to out[0, TVOX) as float.
template <bool TOF, bool DS, bool DZ>
__device__ __forceinline__ void bp_item(const SegRec *__restrict__ recs, const Item &it,
                                        const float *__restrict__ values, const float *__restrict__ y,
                                        const float *__restrict__ p, int kind, float eps, int adjoint,
                                        const LaunchGeom &G, const TileCfg &T, const int *__restrict__ rec_ev,
                                        float *__restrict__ out, unsigned long long *fx, unsigned *s_vmax,
                                        unsigned *s_vmin, int *s_bad)
{
    float *fl = reinterpret_cast<float *>(fx);
    int x0, y0, z0;
    tile_origin(it.tile, T, x0, y0, z0);
    if (threadIdx.x == 0) {
        *s_vmax = 0u;
        *s_vmin = 0x7f800000u;   // +inf: no nonzero value yet
        *s_bad = 0;
    }
    for (int i = threadIdx.x; i < TPAD; i += TILE_THREADS) fx[i] = 0ull;
    const int end = it.start + it.count;
    float vv[RECS_PER_THREAD];
    float vmax = 0.0f, vmin = INFINITY;
    bool bad = false;
#pragma unroll
    for (int k = 0; k < RECS_PER_THREAD; ++k) {
        const int r = it.start + threadIdx.x + k * TILE_THREADS;
        vv[k] = 0.0f;
        if (r < end) {
            const float v = event_value(values, y, p, kind, eps, rec_ev[r]);
            vv[k] = v;
            if (isfinite(v)) {
                vmax = fmaxf(vmax, fabsf(v));
                if (v != 0.0f) vmin = fminf(vmin, fabsf(v));
            } else {
                bad = true;
            }
        }
    }
    __syncthreads();
    // magnitudes of finite non-negative floats order as their bit patterns
    if (vmax > 0.0f) atomicMax(s_vmax, __float_as_uint(vmax));
    if (vmin < INFINITY) atomicMin(s_vmin, __float_as_uint(vmin));
    if (bad) *s_bad = 1;
    __syncthreads();
    const bool fixed = *s_bad == 0;
    const float vtop = __uint_as_float(*s_vmax);
    int nb = 1, etop = 0, shift = 0;
    if (fixed && vtop > 0.0f) {
        int ebot = 0;
        (void)frexpf(vtop, &etop);                       // vtop < 2^etop
        (void)frexpf(__uint_as_float(*s_vmin), &ebot);
        nb = min((etop - ebot) / BAND_BITS + 1, MAX_BANDS);
        shift = band_shift(vtop, adjoint ? it.wmax_fp : it.wmax_bp, etop, 0);
    }
    if (nb > 1) {   // wide-range item (the branch is uniform over the workgroup)
        bp_item_bands<TOF, DS, DZ>(recs, it, values, y, p, kind, eps, adjoint, G, T, rec_ev, out, fx, vtop, etop,
                                   nb);
        return;
    }
    // the next record is loaded while this one is marched
    SegRec Rn;
    if (it.start + (int)threadIdx.x < end) Rn = recs[it.start + threadIdx.x];
#pragma unroll
    for (int k = 0; k < RECS_PER_THREAD; ++k) {
        const float v = vv[k];
        const int r = it.start + threadIdx.x + k * TILE_THREADS;
        if (r >= end) continue;
        const SegRec R = Rn;
        if (r + TILE_THREADS < end) Rn = recs[r + TILE_THREADS];
        const uint32_t ab = adjoint ? R.ab_fp : R.ab_bp;
        if (ab_a(ab) >= ab_b(ab)) continue;
        TileWalk t;
        walk_for(t, fl, TPAD, x0, y0, z0);
        t.fx = fx;
        t.value = fixed ? ldexpf(v, shift) : v;   // scale folded into the value
        if (fixed) march_any<TOF, DS, DZ, M_SCATTER_FX>(R, G, ab, adjoint ? R.Wfp : R.Wbp, t);
        else march_any<TOF, DS, DZ, M_SCATTER>(R, G, ab, adjoint ? R.Wfp : R.Wbp, t);
    }
    __syncthreads();
    if (fixed) {
        for (int i = threadIdx.x; i < TVOX; i += TILE_THREADS)
            out[i] = ldexpf((float)(long long)fx[tile_word(i)], -shift);
    } else {
        for (int i = threadIdx.x; i < TVOX; i += TILE_THREADS) out[i] = fl[tile_word(i)];
    }
}

template <bool TOF, bool DS, bool DZ>
__global__ __launch_bounds__(TILE_THREADS) BP_OCCUPANCY void k_bp_tiled(const SegRec *__restrict__ recs,
                                                           const Item *__restrict__ items,
                                                           const int *__restrict__ order,
                                                           const float *__restrict__ values,
                                                           const float *__restrict__ y, const float *__restrict__ p,
                                                           int kind, float eps, int adjoint, LaunchGeom G, TileCfg T,
                                                           const int *__restrict__ rec_ev,
                                                           float *__restrict__ tile_out)
{
    __shared__ unsigned long long fx[TPAD + TILE_THREADS];   // tile + one dummy slot per lane
    __shared__ unsigned s_vmax, s_vmin;
    __shared__ int s_bad;
    const int k = order[blockIdx.x];
    const Item it = items[k];
    bp_item<TOF, DS, DZ>(recs, it, values, y, p, kind, eps, adjoint, G, T, rec_ev, tile_out + (int64_t)k * TVOX, fx,
                         &s_vmax, &s_vmin, &s_bad);
}

template <bool TOF, bool DS, bool DZ>
__global__ __launch_bounds__(TILE_THREADS) void k_digest_tiled(const SegRec *__restrict__ recs,
                                                               const int *__restrict__ rec_ev,
                                                               const Item *__restrict__ items, int adjoint,
                                                               LaunchGeom G, TileCfg T,
                                                               unsigned long long *__restrict__ hash,
                                                               int *__restrict__ count)
{
    const Item it = items[blockIdx.x];
    int x0, y0, z0;
    tile_origin(it.tile, T, x0, y0, z0);
    for (int r = it.start + threadIdx.x; r < it.start + it.count; r += TILE_THREADS) {
        const SegRec R = recs[r];
        const uint32_t ab = adjoint ? R.ab_fp : R.ab_bp;
        TileWalk t;
        walk_for(t, nullptr, TPAD, x0, y0, z0);   // dummy: never a tile address
        march_any<TOF, DS, DZ, M_DIGEST>(R, G, ab, 0.0f, t);
        if (t.cnt) {
            atomicAdd(hash + rec_ev[r], (unsigned long long)t.h);
            atomicAdd(count + rec_ev[r], t.cnt);
        }
    }
}

// Voxel i of items [k0, k1) of the tile buffer, summed in item order (the
// loads issued four at a time).
__device__ __forceinline__ float sum_items(const float *__restrict__ buf, int k0, int k1, int i)
{
    // up to 8 loads in flight (a tile has ~8 items at the usual sizes); the
    // additions stay in item order
    float s = 0.0f;
    int k = k0;
    for (; k + 8 <= k1; k += 8) {
        float a[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) a[j] = buf[(int64_t)(k + j) * TVOX + i];
#pragma unroll
        for (int j = 0; j < 8; ++j) s += a[j];
    }
    if (k < k1) {
        float a[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) a[j] = k + j < k1 ? buf[(int64_t)(k + j) * TVOX + i] : 0.0f;
#pragma unroll
        for (int j = 0; j < 8; ++j)
            if (k + j < k1) s += a[j];
    }
    return s;
}

// image[v] (+)= sum over the work items of v's tile, in item order.
__global__ __launch_bounds__(256) void k_tile_reduce(const float *__restrict__ tile_out,
                                                     const int *__restrict__ tile_items, LaunchGeom G, TileCfg T,
                                                     int accumulate, float *__restrict__ img)
{
    const int tile = blockIdx.y;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= TVOX) return;
    int x0, y0, z0;
    tile_origin(tile, T, x0, y0, z0);
    const int gx = x0 + (i % TX), gy = y0 + (i / TX) % TY, gz = z0 + i / (TX * TY);
    if (gx >= G.nx || gy >= G.ny || gz >= G.nz) return;
    const float s = sum_items(tile_out, tile_items[tile], tile_items[tile + 1], i);
    const int64_t v = gx + (int64_t)G.nx * (gy + (int64_t)G.ny * gz);
    img[v] = accumulate ? img[v] + s : s;
}

// x[v] = x[v] * (sum over v's tile items) / emap[v] (emap 0 -> guard when
// guard != 0): k_tile_reduce fused with srf_em_update, the same operations.
__global__ __launch_bounds__(256) void k_tile_update(const float *__restrict__ tile_out,
                                                     const int *__restrict__ tile_items, LaunchGeom G, TileCfg T,
                                                     const float *__restrict__ emap, float guard,
                                                     float *__restrict__ x)
{
    const int tile = blockIdx.y;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= TVOX) return;
    int x0, y0, z0;
    tile_origin(tile, T, x0, y0, z0);
    const int gx = x0 + (i % TX), gy = y0 + (i / TX) % TY, gz = z0 + i / (TX * TY);
    if (gx >= G.nx || gy >= G.ny || gz >= G.nz) return;
    const float s = sum_items(tile_out, tile_items[tile], tile_items[tile + 1], i);
    const int64_t v = gx + (int64_t)G.nx * (gy + (int64_t)G.ny * gz);
    float e = emap[v];
    if (guard != 0.0f && e == 0.0f) e = guard;
    x[v] = x[v] * s / e;
}

// x[v] = x[v] * 0 / emap[v]: k_tile_update for a plan with no work items.
__global__ __launch_bounds__(256) void k_update_empty(const float *__restrict__ emap, float guard,
                                                      float *__restrict__ x, int64_t n)
{
    const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (v >= n) return;
    float e = emap[v];
    if (guard != 0.0f && e == 0.0f) e = guard;
    x[v] = x[v] * 0.0f / e;
}

// Per item, the largest weight constant of its records in each mode.
// The records' (Wbp, Wfp) pairs are read from the compact copy the
// permutation writes (8 B per record instead of a 48 B record).
__global__ __launch_bounds__(256) void k_item_wmax(const float2 *__restrict__ w, Item *__restrict__ items)
{
    __shared__ float sb[256], sf[256];
    const Item it = items[blockIdx.x];
    float mb = 0.0f, mf = 0.0f;
    for (int r = it.start + threadIdx.x; r < it.start + it.count; r += 256) {
        const float2 v = w[r];
        mb = fmaxf(mb, fabsf(v.x));
        mf = fmaxf(mf, fabsf(v.y));
    }
    sb[threadIdx.x] = mb;
    sf[threadIdx.x] = mf;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) {
            sb[threadIdx.x] = fmaxf(sb[threadIdx.x], sb[threadIdx.x + k]);
            sf[threadIdx.x] = fmaxf(sf[threadIdx.x], sf[threadIdx.x + k]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        items[blockIdx.x].wmax_bp = sb[0];
        items[blockIdx.x].wmax_fp = sf[0];
    }
}

// p[e] = sum of e's segment partials (in segment order); NaN for non-finite LORs.
// With y != null the EM ratio is applied on the way out: out[e] = ratio(y[e], p[e]).
// The 256 events of a block own one contiguous range of partials: it is
// staged in LDS with coalesced loads (when it fits), then every thread sums
// its event's partials in order -- the same additions as a direct loop.
constexpr int SEG_LOADS = 12;     // staging loads in flight per thread (one round covers 3072 partials)
constexpr int SEG_STAGE = 6144;   // partials per block staged (24 KiB; ~2300 at 9 segments per event)

__global__ __launch_bounds__(256) void k_seg_reduce(const float *__restrict__ partial, const int *__restrict__ ev_off,
                                                    const uint8_t *__restrict__ ev_flag, int64_t n,
                                                    const float *__restrict__ y, int kind, float eps,
                                                    float *__restrict__ out)
{
    __shared__ float st[SEG_STAGE];
    const int64_t e0 = (int64_t)blockIdx.x * 256;
    const int64_t e = e0 + threadIdx.x;
    const int64_t eb = min(e0 + 256, n);
    // the event's own words are loaded first, in flight with the staging
    const bool live = e < n;
    const int k_lo = live ? ev_off[e] : 0, k_hi = live ? ev_off[e + 1] : 0;
    const bool flag = live && ev_flag[e];
    const float ye = (live && y) ? y[e] : 0.0f;
    const int s0 = ev_off[e0], s1 = ev_off[eb];
    const bool staged = s1 - s0 <= SEG_STAGE;
    if (staged) {
        // SEG_LOADS loads per thread issued before their LDS stores, so they
        // are in flight together (a load-store loop waits for each load)
        for (int k0 = s0; k0 < s1; k0 += SEG_LOADS * 256) {
            float v[SEG_LOADS];
#pragma unroll
            for (int j = 0; j < SEG_LOADS; ++j) {
                const int k = k0 + j * 256 + (int)threadIdx.x;
                v[j] = k < s1 ? partial[k] : 0.0f;
            }
#pragma unroll
            for (int j = 0; j < SEG_LOADS; ++j) {
                const int k = k0 + j * 256 + (int)threadIdx.x;
                if (k < s1) st[k - s0] = v[j];
            }
        }
        __syncthreads();
    }
    if (!live) return;
    float s = 0.0f;
    if (staged) {
        for (int k = k_lo; k < k_hi; ++k) s += st[k - s0];
    } else {
        for (int k = k_lo; k < k_hi; ++k) s += partial[k];
    }
    if (flag) s = __builtin_nanf("");
    if (y) {
        const float d = s + eps;
        float q = ye / d;
        if (kind == SRF_RATIO_DIV_NO_NAN && d == 0.0f) q = 0.0f;
        s = q;
    }
    out[e] = s;
}

__global__ void k_fill_u32(unsigned *__restrict__ d, unsigned v, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) d[i] = v;
}

__global__ void k_iota_i32(int *__restrict__ d, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) d[i] = (int)i;
}

// Geometric event order (srf_event_order).  One 64-bit key per entry j of the
// caller's index list: bits 33.. the subset j / nps (clamped to n_sub, so a
// remainder sorts last), then the transaxial angle of the LOR direction folded
// into [0, pi) (7 bits, 128 bins), the axial midpoint (10 bits, 16 mm bins,
// offset 512) and the signed transaxial distance from the axis, negated with
// the folded direction (16 bits, 1 mm bins, offset 32768).  Bins saturate at
// their range; a NaN coordinate gives bin 0.  An index outside [0, n_lors)
// raises *bad and keys only its subset.
__device__ __forceinline__ unsigned order_bin(float v, float hi)
{
    return v != v ? 0u : (unsigned)fminf(fmaxf(v, 0.0f), hi);
}

__global__ void k_event_keys(const float *__restrict__ lors, int64_t n_lors, int64_t lor_stride,
                             const int64_t *__restrict__ pidx, int64_t n, int64_t nps, int64_t n_sub,
                             uint64_t *__restrict__ key, int *__restrict__ bad)
{
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    const int64_t e = pidx[j];
    const int64_t sub = j / nps < n_sub ? j / nps : n_sub;
    uint64_t k = (uint64_t)sub << 33;
    if (e < 0 || e >= n_lors) {
        *bad = 1;
        key[j] = k;
        return;
    }
    const float *L = lors + e * lor_stride;
    const float x1 = L[0], y1 = L[1], z1 = L[2];
    const float dx = L[3] - x1, dy = L[4] - y1;
    float phi = atan2f(dy, dx);
    const bool flip = phi < 0.0f;
    if (flip) phi += 3.14159265358979f;
    const float h = sqrtf(dx * dx + dy * dy);
    float r = (x1 * dy - y1 * dx) / (h > 0.0f ? h : 1.0f);
    if (flip) r = -r;
    const float zm = 0.5f * (z1 + L[5]);
    const unsigned ab = order_bin(floorf(phi * (float)(128.0 / 3.14159265358979323846)), 127.0f);
    const unsigned zb = order_bin(floorf(zm * 0.0625f) + 512.0f, 1023.0f);
    const unsigned rb = order_bin(floorf(r) + 32768.0f, 65535.0f);
    k |= ((uint64_t)ab << 26) | ((uint64_t)zb << 16) | rb;
    key[j] = k;
}

// Records permuted into the order `idx`; their event ids (slot_ev, by slot)
// are written beside them for the backprojector's value gathers.
__global__ void k_permute_recs(const SegRec *__restrict__ recs, const int *__restrict__ slot_ev,
                               const int *__restrict__ idx, int64_t n, SegRec *__restrict__ out,
                               int *__restrict__ out_ev, float2 *__restrict__ out_w)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int j = idx[i];
    const SegRec R = recs[j];
    out[i] = R;
    out_ev[i] = slot_ev[j];
    out_w[i] = make_float2(R.Wbp, R.Wfp);
}

// ---------------------------------------------------------------- batches --
// A batch binds the plans of several scans (one OSEM subset each) and their
// EM buffers.  Its work list holds every (scan, item) pair; the projector and
// backprojector each run as ONE launch over it, either one workgroup per item
// or a persistent grid (workgroups_per_cu x CUs) that pulls items from an
// atomic queue.  Per item the work is exactly the single-plan kernels' (same
// device functions), and the reductions keep their order, so every scan's
// result is bit-identical to running its plan alone.

struct ScanDesc {
    LaunchGeom G;
    TileCfg T;
    const SegRec *recs;
    const Item *items;
    const int *rec_ev;
    const int *tile_items;
    const int *ev_off;
    const uint8_t *ev_flag;
    float *partial;
    float *tile_buf;
    float *x;             // image estimate (kernel layout), updated in place
    const float *emap;    // sensitivity image (kernel layout)
    const float *y;       // event values of the subset
    float *q;             // ratio y / (A x + eps)
    int64_t n;
};

struct Work {
    int scan, item;
};

// Next work index of a workgroup: blockIdx.x first, then from the queue.
__device__ __forceinline__ int next_work(int persistent, unsigned *counter, int *s_next)
{
    if (!persistent) return INT_MAX;
    __syncthreads();   // the whole group is done with the LDS tile
    if (threadIdx.x == 0) *s_next = (int)atomicAdd(counter, 1u) + (int)gridDim.x;
    __syncthreads();
    return *s_next;
}

template <bool TOF, bool DS, bool DZ>
__global__ __launch_bounds__(TILE_THREADS) void k_fp_batch(const ScanDesc *__restrict__ desc,
                                                           const Work *__restrict__ work, int nwork,
                                                           unsigned *__restrict__ counter, int persistent)
{
    __shared__ float tile[LDS_WORDS];
    __shared__ int s_next;
    for (int w = blockIdx.x; w < nwork; w = next_work(persistent, counter, &s_next)) {
        const Work wk = work[w];
        const ScanDesc *d = desc + wk.scan;
        const Item it = d->items[wk.item];
        const LaunchGeom G = d->G;
        const TileCfg T = d->T;
        fp_item<TOF, DS, DZ>(d->recs, it, d->x, G, T, d->partial, tile);
    }
}

template <bool TOF, bool DS, bool DZ>
__global__ __launch_bounds__(TILE_THREADS) BP_OCCUPANCY void k_bp_batch(const ScanDesc *__restrict__ desc,
                                                           const Work *__restrict__ work, int nwork,
                                                           unsigned *__restrict__ counter, int persistent)
{
    __shared__ unsigned long long fx[TPAD + TILE_THREADS];   // tile + one dummy slot per lane
    __shared__ unsigned s_vmax, s_vmin;
    __shared__ int s_bad;
    __shared__ int s_next;
    for (int w = blockIdx.x; w < nwork; w = next_work(persistent, counter, &s_next)) {
        const Work wk = work[w];
        const ScanDesc *d = desc + wk.scan;
        const Item it = d->items[wk.item];
        const LaunchGeom G = d->G;
        const TileCfg T = d->T;
        bp_item<TOF, DS, DZ>(d->recs, it, d->q, nullptr, nullptr, 0, 0.0f, 0, G, T, d->rec_ev,
                             d->tile_buf + (int64_t)wk.item * TVOX, fx, &s_vmax, &s_vmin, &s_bad);
    }
}

// index s with base[s] <= v < base[s + 1] (base non-decreasing, base[0] <= v < base[n])
template <class T>
__device__ __forceinline__ int find_segment(const T *__restrict__ base, int n, T v)
{
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (base[mid] <= v) lo = mid;
        else hi = mid;
    }
    return lo;
}

// q = ratio(y, sum of segment partials) for every event of every scan (the
// order and arithmetic of k_seg_reduce).
__global__ __launch_bounds__(256) void k_seg_reduce_batch(const ScanDesc *__restrict__ desc,
                                                          const int64_t *__restrict__ ev_base, int nscan, int kind,
                                                          float eps)
{
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= ev_base[nscan]) return;
    const int sc = find_segment(ev_base, nscan, e);
    const ScanDesc *d = desc + sc;
    const int64_t le = e - ev_base[sc];
    float s = 0.0f;
    for (int k = d->ev_off[le]; k < d->ev_off[le + 1]; ++k) s += d->partial[k];
    if (d->ev_flag[le]) s = __builtin_nanf("");
    const float den = s + eps;
    float q = d->y[le] / den;
    if (kind == SRF_RATIO_DIV_NO_NAN && den == 0.0f) q = 0.0f;
    d->q[le] = q;
}

// Backprojected tiles summed in item order (k_tile_reduce), then the EM
// update of srf_em_update: x = x * b / (emap == 0 && guard != 0 ? guard : emap).
__global__ __launch_bounds__(256) void k_tile_update_batch(const ScanDesc *__restrict__ desc,
                                                           const int *__restrict__ tile_base, int nscan, float guard)
{
    constexpr int BPT = TVOX / 256;   // blocks per tile
    const int gt = blockIdx.x / BPT;
    const int i = (blockIdx.x % BPT) * 256 + threadIdx.x;
    if (gt >= tile_base[nscan]) return;
    const int sc = find_segment(tile_base, nscan, gt);
    const ScanDesc *d = desc + sc;
    const int tile = gt - tile_base[sc];
    int x0, y0, z0;
    tile_origin(tile, d->T, x0, y0, z0);
    const int gx = x0 + (i % TX), gy = y0 + (i / TX) % TY, gz = z0 + i / (TX * TY);
    if (gx >= d->G.nx || gy >= d->G.ny || gz >= d->G.nz) return;
    const float s = sum_items(d->tile_buf, d->tile_items[tile], d->tile_items[tile + 1], i);
    const int64_t v = gx + (int64_t)d->G.nx * (gy + (int64_t)d->G.ny * gz);
    float e = d->emap[v];
    if (guard != 0.0f && e == 0.0f) e = guard;
    d->x[v] = d->x[v] * s / e;
}

}  // namespace srf

using namespace srf;

// Estimated work of an item: its records times their slab-count class + 1.
static inline int64_t item_work(const Item &it) { return (int64_t)it.count * (it.cls + 1); }

// Launch orders of a plan's items; the results do not depend on them (every
// item writes its own output).  Largest first (estimated work) keeps the tail
// of a launch made of small items: the projector's order.
static std::vector<int> launch_order_largest_first(const std::vector<Item> &items)
{
    std::vector<int> order(items.size());
    for (size_t k = 0; k < items.size(); ++k) order[k] = (int)k;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return item_work(items[(size_t)a]) > item_work(items[(size_t)b]);
    });
    return order;
}

// The backprojector's order.  Workgroup i of a launch runs on XCD i % 8, and
// each of the 8 XCDs has its own L2.  The items, in tile order (z slowest),
// are cut into 8 bands of about equal estimated work; band x takes the launch
// positions i % 8 == x, largest item first, so that one XCD works on a few
// consecutive tile layers, whose segments belong to mostly the same events:
// their ratio gathers hit that XCD's L2 more often.  A band that runs out
// gives its positions to the band with the most items left.
constexpr int N_XCD = 8;
static std::vector<int> launch_order_by_xcd(const std::vector<Item> &items)
{
    int64_t total = 0;
    for (const Item &it : items) total += item_work(it);
    std::vector<std::vector<int>> band(N_XCD);
    int64_t acc = 0;
    for (size_t k = 0; k < items.size(); ++k) {
        const int x = total > 0 ? (int)std::min<int64_t>(N_XCD - 1, acc * N_XCD / total) : 0;
        band[(size_t)x].push_back((int)k);
        acc += item_work(items[k]);
    }
    for (auto &b : band)
        std::stable_sort(b.begin(), b.end(),
                         [&](int a, int c) { return item_work(items[(size_t)a]) > item_work(items[(size_t)c]); });
    std::vector<int> order;
    order.reserve(items.size());
    std::vector<size_t> next(N_XCD, 0);
    while (order.size() < items.size()) {
        for (int x = 0; x < N_XCD && order.size() < items.size(); ++x) {
            int from = x;
            if (next[(size_t)x] == band[(size_t)x].size()) {
                size_t most = 0;
                for (int y = 0; y < N_XCD; ++y) {
                    const size_t left = band[(size_t)y].size() - next[(size_t)y];
                    if (left > most) {
                        most = left;
                        from = y;
                    }
                }
            }
            order.push_back(band[(size_t)from][next[(size_t)from]++]);
        }
    }
    return order;
}

// A device buffer with its capacity in elements (plans that are rebuilt keep
// their buffers and only grow them).
template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t cap = 0;
    bool pooled = false;   // from the plan pool (else hipMalloc)
};

struct srf_plan {
    LaunchGeom G;
    TileCfg T;
    bool tof, ds, dz;
    int64_t n, nseg;
    int nitems;
    bool valid;   // the last build succeeded: item and tile tables match G / T
    DevBuf<SegRec> recs;
    DevBuf<int> rec_ev;
    DevBuf<Item> items;
    DevBuf<int> tile_items;
    DevBuf<int> order;      // projector's launch order of the items: estimated work, largest first
    DevBuf<int> order_bp;   // backprojector's: 8 bands of tile layers, one per XCD
    // one-shot timing events around the next tiled kernel (srf_plan_time_next)
    hipEvent_t ev_before = nullptr, ev_after = nullptr;
    DevBuf<int> ev_off;
    DevBuf<uint8_t> ev_flag;
    DevBuf<float> partial;
    DevBuf<float> tile_buf;
    // binning scratch: freed after srf_plan_create, kept by srf_plan_rebuild
    DevBuf<int> ev_cnt;
    DevBuf<unsigned> key_count;
    DevBuf<char> scan_tmp, sort_tmp;
    DevBuf<unsigned> sk, sk2;
    DevBuf<int> idx, idx2;
    DevBuf<float2> wpair;   // (Wbp, Wfp) per record in the sorted order
    DevBuf<SegRec> recs2;
    DevBuf<int> slot_ev;    // event id per record slot, before the sort
};

namespace {

template <class T>
int dmalloc(T **p, size_t count, size_t &total)
{
    *p = nullptr;
    if (count == 0) count = 1;
    if (hipMalloc((void **)p, count * sizeof(T)) != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        return SRF_ERR_LAUNCH;
    }
    total += count * sizeof(T);
    return SRF_OK;
}

// Plan memory comes from one stream-ordered pool per device that keeps up to
// POOL_KEEP bytes of freed memory for reuse.  Many plans built and destroyed
// in a row (one per OSEM subset of many scans, or chunk after chunk) then
// reuse pool memory instead of calling hipMalloc / hipFree, which map and
// unmap memory and wait for the device each time.
constexpr uint64_t POOL_KEEP = 16ull << 30;
constexpr int POOL_DEVICES = 64;
std::mutex pool_mu;
hipMemPool_t pools[POOL_DEVICES];
bool pool_failed[POOL_DEVICES];

hipMemPool_t plan_pool()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= POOL_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lock(pool_mu);
    if (!pools[dev] && !pool_failed[dev]) {
        hipMemPoolProps props = {};
        props.allocType = hipMemAllocationTypePinned;
        props.handleTypes = hipMemHandleTypeNone;
        props.location.type = hipMemLocationTypeDevice;
        props.location.id = dev;
        hipMemPool_t pool = nullptr;
        if (hipMemPoolCreate(&pool, &props) != hipSuccess) {
            (void)hipGetLastError();
            pool_failed[dev] = true;
            return nullptr;
        }
        uint64_t keep = POOL_KEEP;
        (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
        pools[dev] = pool;
    }
    return pools[dev];
}

// Freed in stream order on s: work already queued on s may still read it.
template <class T>
void dfree(DevBuf<T> &b, hipStream_t s)
{
    if (b.p) {
        if (b.pooled) (void)hipFreeAsync(b.p, s);
        else (void)hipFree(b.p);
    }
    b.p = nullptr;
    b.cap = 0;
    b.pooled = false;
}

// At least `count` elements (1 if count is 0), reallocated only to grow.  A
// buffer that grows again (a rebuild with more segments than any before) gets
// a quarter of headroom, so that a sequence of similar rebuilds reallocates
// (and drains the device) a few times rather than at every new maximum.
template <class T>
int ensure(DevBuf<T> &b, size_t count, hipStream_t s)
{
    if (count == 0) count = 1;
    if (b.cap >= count) return SRF_OK;
    if (b.cap > 0) {
        // a rebuilt plan outgrows this buffer: work on other streams may still
        // read it, so drain the device before it goes back to the pool
        count += count / 4;
        (void)hipDeviceSynchronize();
        dfree(b, s);
    }
    // a device without memory pools falls back to hipMalloc
    hipMemPool_t pool = plan_pool();
    const hipError_t e = pool ? hipMallocFromPoolAsync((void **)&b.p, count * sizeof(T), pool, s)
                              : hipMalloc((void **)&b.p, count * sizeof(T));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        b.p = nullptr;
        return SRF_ERR_LAUNCH;
    }
    b.cap = count;
    b.pooled = pool != nullptr;
    return SRF_OK;
}

void free_scratch(srf_plan *P, hipStream_t s)
{
    dfree(P->ev_cnt, s);
    dfree(P->key_count, s);
    dfree(P->scan_tmp, s);
    dfree(P->sort_tmp, s);
    dfree(P->sk, s);
    dfree(P->sk2, s);
    dfree(P->idx, s);
    dfree(P->idx2, s);
    dfree(P->wpair, s);
    dfree(P->recs2, s);
    dfree(P->slot_ev, s);
}

// Destroy: the plan may have been used on any stream, so the device is
// drained first (as hipFree would), then the memory goes back to the pool.
void plan_free(srf_plan *P)
{
    if (!P) return;
    (void)hipDeviceSynchronize();
    const hipStream_t s = nullptr;
    dfree(P->recs, s);
    dfree(P->rec_ev, s);
    dfree(P->items, s);
    dfree(P->tile_items, s);
    dfree(P->order, s);
    dfree(P->order_bp, s);
    dfree(P->ev_off, s);
    dfree(P->ev_flag, s);
    dfree(P->partial, s);
    dfree(P->tile_buf, s);
    free_scratch(P, s);
    delete P;
}

size_t plan_bytes(const srf_plan *P)
{
    return P->recs.cap * sizeof(SegRec) + P->rec_ev.cap * sizeof(int) + P->items.cap * sizeof(Item) +
           P->tile_items.cap * sizeof(int) + P->order.cap * sizeof(int) +
           P->order_bp.cap * sizeof(int) + P->ev_off.cap * sizeof(int) +
           P->ev_flag.cap + P->partial.cap * sizeof(float) + P->tile_buf.cap * sizeof(float);
}

}  // namespace

// plan_build: return the error code (the caller decides what to free)
#define PB_TRY(x)                 \
    do {                          \
        int rc__ = (x);           \
        if (rc__) return rc__;    \
    } while (0)

#define PB_HIP(x)                                   \
    do {                                            \
        if ((x) != hipSuccess) {                    \
            (void)hipGetLastError();                \
            return SRF_ERR_LAUNCH;                  \
        }                                           \
    } while (0)

// entry points with a local cleanup(): run it, then return the error code
#define SRF_TRY(x)                \
    do {                          \
        int rc__ = (x);           \
        if (rc__) {               \
            cleanup();            \
            return rc__;          \
        }                         \
    } while (0)

#define SRF_HIP(x)                                  \
    do {                                            \
        if ((x) != hipSuccess) {                    \
            (void)hipGetLastError();                \
            cleanup();                              \
            return SRF_ERR_LAUNCH;                  \
        }                                           \
    } while (0)

namespace {

// Bins an event list into P (whose buffers are grown as needed and otherwise
// reused); on failure P holds no valid list.
// v = 2^k with 2^-64 <= v <= 2^64, k != 0: dividing by v is multiplying by
// the exact 1 / v (no subnormal for the coordinates and slopes involved).
static bool pow2_edge(float v)
{
    int e = 0;
    return v != 1.0f && v >= 0x1p-64f && v <= 0x1p64f && frexpf(v, &e) == 0.5f;
}

int plan_build(srf_plan *P, const float *lors, int64_t n, int64_t lor_stride, const float *tof_values,
               int64_t tof_stride, srf_grid_t grid, const srf_tof_t *tof, const srf_plan_opts_t *opts,
               hipStream_t s)
{
    const int32_t item_records = (opts && opts->item_records) ? opts->item_records : ITEM;
    const float twin = opts ? opts->tof_window : 0.0f;
    if (item_records < 256 || item_records > ITEM) return SRF_ERR_INVALID_ARG;
    if (!(twin >= 0.0f) || twin > 1000.0f) return SRF_ERR_INVALID_ARG;
    const unsigned item_max = (unsigned)item_records;
    if (n < 0 || (n > 0 && (!lors || lor_stride < 6))) return SRF_ERR_INVALID_ARG;
    if (tof && n > 0 && (!tof_values || tof_stride < 1)) return SRF_ERR_INVALID_ARG;
    // n + 1 event offsets are scanned as a 32-bit count
    if (n >= (int64_t)INT_MAX) return SRF_ERR_UNSUPPORTED;
    LaunchGeom G;
    int rc = make_geom(grid, tof, G);
    if (rc) return rc;
    if (grid.nx > 32767 || grid.ny > 32767) return SRF_ERR_UNSUPPORTED;   // slab index packs in 15 bits
    const int prescale = (pow2_edge(G.dy) ? 1 : 0) | (pow2_edge(G.dz) ? 2 : 0);
    TileCfg T;
    T.ntx = (grid.nx + TX - 1) / TX;
    T.nty = (grid.ny + TY - 1) / TY;
    T.ntz = (grid.nz + TZ - 1) / TZ;
    T.ntiles = T.ntx * T.nty * T.ntz;
    T.nkeys = T.ntiles * NCLASS;
    if (T.nkeys > MAX_KEYS) return SRF_ERR_UNSUPPORTED;

    // the plan holds no list until this build succeeds; its grid and kernel
    // variant are committed only then
    P->n = 0;
    P->nseg = 0;
    P->nitems = 0;
    P->valid = false;
    const bool use_tof = tof != nullptr;
    PB_TRY(ensure(P->ev_off, (size_t)n + 1, s));
    PB_TRY(ensure(P->ev_flag, (size_t)n, s));
    PB_TRY(ensure(P->ev_cnt, (size_t)n + 1, s));
    PB_TRY(ensure(P->key_count, (size_t)T.nkeys, s));
    hipLaunchKernelGGL(k_fill_u32, dim3((T.nkeys + 255) / 256), dim3(256), 0, s, P->key_count.p, 0u,
                       (int64_t)T.nkeys);
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((n + 1 + 255) / 256)), dim3(256), 0, s,
                       (unsigned *)P->ev_cnt.p, 0u, n + 1);
    const unsigned nb = (unsigned)((n + BIN_THREADS - 1) / BIN_THREADS);
    const size_t shm = (size_t)T.nkeys * sizeof(unsigned);
    if (n > 0) {
        if (use_tof)
            hipLaunchKernelGGL(k_bin_count<true>, dim3(nb), dim3(BIN_THREADS), shm, s, lors, n, (int)lor_stride,
                               tof_values, (int)tof_stride, G, T, twin, P->ev_cnt.p, P->ev_flag.p, P->key_count.p);
        else
            hipLaunchKernelGGL(k_bin_count<false>, dim3(nb), dim3(BIN_THREADS), shm, s, lors, n, (int)lor_stride,
                               tof_values, (int)tof_stride, G, T, twin, P->ev_cnt.p, P->ev_flag.p, P->key_count.p);
    }
    PB_HIP(hipGetLastError());
    // event offsets: exclusive scan of the per-event counts (ev_off[n] = total)
    size_t tmp_bytes = 0;
    PB_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, P->ev_cnt.p, P->ev_off.p, (int)(n + 1), s));
    PB_TRY(ensure(P->scan_tmp, tmp_bytes, s));
    PB_HIP(hipcub::DeviceScan::ExclusiveSum(P->scan_tmp.p, tmp_bytes, P->ev_cnt.p, P->ev_off.p, (int)(n + 1), s));
    std::vector<unsigned> kc((size_t)T.nkeys);
    int total = 0;
    PB_HIP(hipMemcpyAsync(kc.data(), P->key_count.p, kc.size() * sizeof(unsigned), hipMemcpyDeviceToHost, s));
    PB_HIP(hipMemcpyAsync(&total, P->ev_off.p + n, sizeof(int), hipMemcpyDeviceToHost, s));
    PB_HIP(hipStreamSynchronize(s));
    // segment slots are 32-bit: a list with more than 2^31 - 1 segments (the
    // 32-bit prefix sum above would have wrapped) is refused
    uint64_t kc_sum = 0;
    for (unsigned c : kc) kc_sum += c;
    if (kc_sum > (uint64_t)INT_MAX) return SRF_ERR_UNSUPPORTED;
    const int64_t nseg = total;
    // work items from the key counts (keys are tile-major, so a tile's items are contiguous)
    std::vector<Item> items;
    std::vector<int> order, order_bp;
    std::vector<int> tile_items((size_t)T.ntiles + 1, 0);
    unsigned acc = 0;
    for (int k = 0; k < T.nkeys; ++k) {
        const int tile = k / NCLASS;
        if (k % NCLASS == 0) tile_items[tile] = (int)items.size();
        for (unsigned st = 0; st < kc[k]; st += item_max) {
            items.push_back(Item{(int)(acc + st), (int)std::min<unsigned>(item_max, kc[k] - st), tile, k % NCLASS,
                                 0.0f, 0.0f, {0.0f, 0.0f}});
        }
        acc += kc[k];
    }
    tile_items[T.ntiles] = (int)items.size();
    if ((int64_t)acc != nseg) return SRF_ERR_LAUNCH;
    const int nitems = (int)items.size();
    PB_TRY(ensure(P->recs, (size_t)nseg, s));
    PB_TRY(ensure(P->rec_ev, (size_t)nseg, s));
    PB_TRY(ensure(P->partial, (size_t)nseg, s));
    PB_TRY(ensure(P->items, items.size(), s));
    PB_TRY(ensure(P->tile_items, tile_items.size(), s));
    PB_TRY(ensure(P->order, items.size(), s));
    PB_TRY(ensure(P->order_bp, items.size(), s));
    PB_TRY(ensure(P->tile_buf, (size_t)std::max(1, nitems) * TVOX, s));
    if (!items.empty())
        PB_HIP(hipMemcpyAsync(P->items.p, items.data(), items.size() * sizeof(Item), hipMemcpyHostToDevice, s));
    PB_HIP(hipMemcpyAsync(P->tile_items.p, tile_items.data(), tile_items.size() * sizeof(int),
                           hipMemcpyHostToDevice, s));
    order = launch_order_largest_first(items);
    order_bp = launch_order_by_xcd(items);
    if (!order.empty()) {
        PB_HIP(hipMemcpyAsync(P->order.p, order.data(), order.size() * sizeof(int), hipMemcpyHostToDevice, s));
        PB_HIP(hipMemcpyAsync(P->order_bp.p, order_bp.data(), order_bp.size() * sizeof(int),
                               hipMemcpyHostToDevice, s));
    }
    if (n > 0) {
        PB_TRY(ensure(P->sk, (size_t)nseg, s));
        PB_TRY(ensure(P->sk2, (size_t)nseg, s));
        PB_TRY(ensure(P->idx, (size_t)nseg, s));
        PB_TRY(ensure(P->idx2, (size_t)nseg, s));
        PB_TRY(ensure(P->recs2, (size_t)nseg, s));
        PB_TRY(ensure(P->wpair, (size_t)nseg, s));
        PB_TRY(ensure(P->slot_ev, (size_t)nseg, s));
        if (use_tof)
            hipLaunchKernelGGL(k_bin_emit<true>, dim3(nb), dim3(BIN_THREADS), 0, s, lors, n, (int)lor_stride,
                               tof_values, (int)tof_stride, G, T, twin, prescale, P->ev_off.p, P->recs.p, P->slot_ev.p,
                               P->sk.p);
        else
            hipLaunchKernelGGL(k_bin_emit<false>, dim3(nb), dim3(BIN_THREADS), 0, s, lors, n, (int)lor_stride,
                               tof_values, (int)tof_stride, G, T, twin, prescale, P->ev_off.p, P->recs.p, P->slot_ev.p,
                               P->sk.p);
        // Record order: a stable sort of the slot-ordered records by (key,
        // branch, slab count) -- within one, slot order.  The order (and with
        // it the split into work items and each item's fixed-point scale) is a
        // function of the events alone, so every plan of the same events is
        // bit-identical, and each wavefront gets one branch's segments of
        // similar length.
        if (nseg > 0) {
            const unsigned nbs = (unsigned)((nseg + 255) / 256);
            hipLaunchKernelGGL(k_iota_i32, dim3(nbs), dim3(256), 0, s, P->idx.p, nseg);
            int key_bits = 1;
            while ((1 << key_bits) < T.nkeys) ++key_bits;
            if (7 + key_bits > 32) return SRF_ERR_UNSUPPORTED;
            size_t sort_bytes = 0;
            PB_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, P->sk.p, P->sk2.p, P->idx.p, P->idx2.p,
                                                       (int)nseg, 0, 7 + key_bits, s));
            PB_TRY(ensure(P->sort_tmp, sort_bytes, s));
            PB_HIP(hipcub::DeviceRadixSort::SortPairs(P->sort_tmp.p, sort_bytes, P->sk.p, P->sk2.p, P->idx.p,
                                                       P->idx2.p, (int)nseg, 0, 7 + key_bits, s));
            hipLaunchKernelGGL(k_permute_recs, dim3(nbs), dim3(256), 0, s, P->recs.p, P->slot_ev.p, P->idx2.p,
                               nseg, P->recs2.p, P->rec_ev.p, P->wpair.p);
            PB_HIP(hipGetLastError());
            std::swap(P->recs, P->recs2);
        }
    }
    if (nitems > 0)
        hipLaunchKernelGGL(k_item_wmax, dim3((unsigned)nitems), dim3(256), 0, s, P->wpair.p, P->items.p);
    PB_HIP(hipGetLastError());
    PB_HIP(hipStreamSynchronize(s));
    P->G = G;
    P->T = T;
    P->tof = use_tof;
    P->ds = G.dy != 1.0f && !(prescale & 1);
    P->dz = G.dz != 1.0f && !(prescale & 2);
    P->n = n;
    P->nseg = nseg;
    P->nitems = nitems;
    P->valid = true;
    return SRF_OK;
}

}  // namespace

extern "C" {

int srf_plan_create(const float *lors, int64_t n, int64_t lor_stride, const float *tof_values, int64_t tof_stride,
                    srf_grid_t grid, const srf_tof_t *tof, srf_plan **out, void *stream)
{
    return srf_plan_create_ex(lors, n, lor_stride, tof_values, tof_stride, grid, tof, nullptr, out, stream);
}

int srf_plan_create_ex(const float *lors, int64_t n, int64_t lor_stride, const float *tof_values,
                       int64_t tof_stride, srf_grid_t grid, const srf_tof_t *tof, const srf_plan_opts_t *opts,
                       srf_plan **out, void *stream)
{
    if (!out) return SRF_ERR_INVALID_ARG;
    *out = nullptr;
    srf_plan *P = new srf_plan();
    const int rc = plan_build(P, lors, n, lor_stride, tof_values, tof_stride, grid, tof, opts, (hipStream_t)stream);
    if (rc) {
        (void)hipStreamSynchronize((hipStream_t)stream);
        (void)hipGetLastError();
        plan_free(P);
        return rc;
    }
    free_scratch(P, (hipStream_t)stream);
    *out = P;
    return SRF_OK;
}

int srf_plan_rebuild(srf_plan *plan, const float *lors, int64_t n, int64_t lor_stride, const float *tof_values,
                     int64_t tof_stride, srf_grid_t grid, const srf_tof_t *tof, const srf_plan_opts_t *opts,
                     void *stream)
{
    if (!plan) return SRF_ERR_INVALID_ARG;
    const int rc = plan_build(plan, lors, n, lor_stride, tof_values, tof_stride, grid, tof, opts,
                              (hipStream_t)stream);
    if (rc) {
        (void)hipStreamSynchronize((hipStream_t)stream);
        (void)hipGetLastError();
        // plan_build committed nothing: the plan keeps its previous grid and
        // holds an empty list (projections write nothing, backprojections
        // zeros); it cannot join a batch until a rebuild succeeds
        plan->n = plan->nseg = 0;
        plan->nitems = 0;
        plan->valid = false;
    }
    return rc;
}

int srf_plan_destroy(srf_plan *plan)
{
    plan_free(plan);
    return SRF_OK;
}

int srf_plan_pool_trim(size_t keep_bytes)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= POOL_DEVICES) {
        (void)hipGetLastError();
        return SRF_ERR_LAUNCH;
    }
    hipMemPool_t pool;
    {
        std::lock_guard<std::mutex> lock(pool_mu);
        pool = pools[dev];
    }
    if (!pool) return SRF_OK;   // no plan was built on this device
    // memory freed in stream order becomes releasable once that work is done
    if (hipDeviceSynchronize() != hipSuccess || hipMemPoolTrimTo(pool, keep_bytes) != hipSuccess) {
        (void)hipGetLastError();
        return SRF_ERR_LAUNCH;
    }
    return SRF_OK;
}

int srf_event_order(const float *lors, int64_t n_lors, int64_t lor_stride, const int64_t *pidx, int64_t n,
                    int64_t n_sub, int64_t *out, void *stream)
{
    if (n < 0 || n_lors < 0 || n_sub < 1 || lor_stride < 6) return SRF_ERR_INVALID_ARG;
    if (n == 0) return SRF_OK;
    if (!lors || !pidx || !out || out == pidx) return SRF_ERR_INVALID_ARG;
    if (n >= INT32_MAX || n_sub >= ((int64_t)1 << 30)) return SRF_ERR_UNSUPPORTED;
    const hipStream_t s = (hipStream_t)stream;
    const int64_t nps = n / n_sub > 0 ? n / n_sub : 1;
    int sub_bits = 1;
    while (((int64_t)1 << sub_bits) <= n_sub) ++sub_bits;
    DevBuf<uint64_t> key, key2;
    DevBuf<unsigned char> tmp;
    DevBuf<int> bad;
    auto cleanup = [&]() {
        dfree(key, s);
        dfree(key2, s);
        dfree(tmp, s);
        dfree(bad, s);
    };
    SRF_TRY(ensure(key, (size_t)n, s));
    SRF_TRY(ensure(key2, (size_t)n, s));
    SRF_TRY(ensure(bad, 1, s));
    SRF_HIP(hipMemsetAsync(bad.p, 0, sizeof(int), s));
    hipLaunchKernelGGL(k_event_keys, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, lors, n_lors, lor_stride,
                       pidx, n, nps, n_sub, key.p, bad.p);
    SRF_HIP(hipGetLastError());
    size_t bytes = 0;
    SRF_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, key.p, key2.p, pidx, out, (int)n, 0, 33 + sub_bits, s));
    SRF_TRY(ensure(tmp, bytes, s));
    SRF_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, bytes, key.p, key2.p, pidx, out, (int)n, 0, 33 + sub_bits, s));
    int h_bad = 0;
    SRF_HIP(hipMemcpyAsync(&h_bad, bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
    SRF_HIP(hipStreamSynchronize(s));
    cleanup();
    return h_bad ? SRF_ERR_INVALID_ARG : SRF_OK;
}

int srf_plan_info(const srf_plan *plan, int64_t *n_events, int64_t *n_segments, int64_t *n_items,
                  size_t *device_bytes)
{
    if (!plan) return SRF_ERR_INVALID_ARG;
    if (n_events) *n_events = plan->n;
    if (n_segments) *n_segments = plan->nseg;
    if (n_items) *n_items = plan->nitems;
    if (device_bytes) *device_bytes = plan_bytes(plan);
    return SRF_OK;
}

#define PLAN_DISPATCH(KERNEL, ...)                                                                   \
    do {                                                                                             \
        if (P->tof) {                                                                                \
            if (P->ds) { if (P->dz) KERNEL<true, true, true> __VA_ARGS__; else KERNEL<true, true, false> __VA_ARGS__; } \
            else { if (P->dz) KERNEL<true, false, true> __VA_ARGS__; else KERNEL<true, false, false> __VA_ARGS__; } \
        } else {                                                                                     \
            if (P->ds) { if (P->dz) KERNEL<false, true, true> __VA_ARGS__; else KERNEL<false, true, false> __VA_ARGS__; } \
            else { if (P->dz) KERNEL<false, false, true> __VA_ARGS__; else KERNEL<false, false, false> __VA_ARGS__; } \
        }                                                                                            \
    } while (0)

// Records the pending timing events (if any) before / after the tiled kernel;
// they are consumed by that one launch.
static void mark_before(srf_plan *P, hipStream_t s)
{
    if (P->ev_before) (void)hipEventRecord(P->ev_before, s);
}
static void mark_after(srf_plan *P, hipStream_t s)
{
    if (P->ev_after) (void)hipEventRecord(P->ev_after, s);
    P->ev_before = P->ev_after = nullptr;
}

int srf_plan_time_next(srf_plan *P, void *before, void *after)
{
    if (!P || !before || !after) return SRF_ERR_INVALID_ARG;
    P->ev_before = (hipEvent_t)before;
    P->ev_after = (hipEvent_t)after;
    return SRF_OK;
}

int srf_plan_project(srf_plan *P, const float *image, float *out, void *stream)
{
    if (!P) return SRF_ERR_INVALID_ARG;
    if (P->n == 0) return SRF_OK;
    if (!image || !out) return SRF_ERR_INVALID_ARG;
    hipStream_t s = (hipStream_t)stream;
    mark_before(P, s);
    if (P->nitems > 0)
        PLAN_DISPATCH(k_fp_tiled, <<<P->nitems, TILE_THREADS, 0, s>>>(P->recs.p, P->items.p, P->order.p, image, P->G, P->T,
                                                                       P->partial.p));
    mark_after(P, s);
    hipLaunchKernelGGL(k_seg_reduce, dim3((unsigned)((P->n + 255) / 256)), dim3(256), 0, s, P->partial.p, P->ev_off.p,
                       P->ev_flag.p, P->n, (const float *)nullptr, 0, 0.0f, out);
    return launch_status();
}

int srf_plan_project_ratio(srf_plan *P, const float *image, const float *y, srf_ratio_kind_t kind, float eps,
                           float *q, void *stream)
{
    if (!P) return SRF_ERR_INVALID_ARG;
    if (kind != SRF_RATIO_PLAIN && kind != SRF_RATIO_DIV_NO_NAN) return SRF_ERR_INVALID_ARG;
    if (P->n == 0) return SRF_OK;
    if (!image || !y || !q) return SRF_ERR_INVALID_ARG;
    hipStream_t s = (hipStream_t)stream;
    mark_before(P, s);
    if (P->nitems > 0)
        PLAN_DISPATCH(k_fp_tiled, <<<P->nitems, TILE_THREADS, 0, s>>>(P->recs.p, P->items.p, P->order.p, image, P->G, P->T,
                                                                       P->partial.p));
    mark_after(P, s);
    hipLaunchKernelGGL(k_seg_reduce, dim3((unsigned)((P->n + 255) / 256)), dim3(256), 0, s, P->partial.p, P->ev_off.p,
                       P->ev_flag.p, P->n, y, (int)kind, eps, q);
    return launch_status();
}

static int64_t plan_voxels(const srf_plan *P) { return (int64_t)P->G.nx * P->G.ny * P->G.nz; }

static int plan_bp(srf_plan *P, const float *values, const float *y, const float *p, int kind, float eps,
                   srf_bp_mode_t mode, int accumulate, float *image, hipStream_t s)
{
    if (!P || !image) return SRF_ERR_INVALID_ARG;
    if (mode != SRF_BP_REFERENCE && mode != SRF_BP_ADJOINT) return SRF_ERR_INVALID_ARG;
    if (P->nitems == 0) {
        // an empty list (also after a failed rebuild): the backprojection is
        // zero, and no tile buffer or item table is read
        if (!accumulate) {
            const int64_t nv = plan_voxels(P);
            hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((nv + 255) / 256)), dim3(256), 0, s, (unsigned *)image,
                               0u, nv);
        }
        return launch_status();
    }
    mark_before(P, s);
    PLAN_DISPATCH(k_bp_tiled, <<<P->nitems, TILE_THREADS, 0, s>>>(P->recs.p, P->items.p, P->order_bp.p, values, y, p,
                                                                   kind, eps, (int)(mode == SRF_BP_ADJOINT), P->G, P->T,
                                                                   P->rec_ev.p, P->tile_buf.p));
    mark_after(P, s);
    dim3 grid((TVOX + 255) / 256, (unsigned)P->T.ntiles);
    hipLaunchKernelGGL(k_tile_reduce, grid, dim3(256), 0, s, P->tile_buf.p, P->tile_items.p, P->G, P->T, accumulate,
                       image);
    return launch_status();
}

int srf_plan_digest(srf_plan *P, srf_bp_mode_t mode, uint64_t *hash, int32_t *count, void *stream)
{
    if (!P) return SRF_ERR_INVALID_ARG;
    if (mode != SRF_BP_REFERENCE && mode != SRF_BP_ADJOINT) return SRF_ERR_INVALID_ARG;
    if (P->n == 0) return SRF_OK;
    if (!hash || !count) return SRF_ERR_INVALID_ARG;
    hipStream_t s = (hipStream_t)stream;
    const unsigned nb = (unsigned)((2 * P->n + 255) / 256);
    hipLaunchKernelGGL(k_fill_u32, dim3(nb), dim3(256), 0, s, (unsigned *)hash, 0u, 2 * P->n);
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((P->n + 255) / 256)), dim3(256), 0, s, (unsigned *)count, 0u,
                       P->n);
    if (P->nitems > 0)
        PLAN_DISPATCH(k_digest_tiled, <<<P->nitems, TILE_THREADS, 0, s>>>(P->recs.p, P->rec_ev.p, P->items.p,
                                                                           (int)(mode == SRF_BP_ADJOINT), P->G,
                                                                           P->T, (unsigned long long *)hash, count));
    return launch_status();
}

int srf_plan_backproject(srf_plan *P, const float *values, srf_bp_mode_t mode, int accumulate, float *image,
                         void *stream)
{
    return plan_bp(P, values, nullptr, nullptr, 0, 0.0f, mode, accumulate, image, (hipStream_t)stream);
}

int srf_plan_backproject_update(srf_plan *P, const float *values, float *x, const float *emap, float zero_guard,
                                void *stream)
{
    if (!P || !x || !emap) return SRF_ERR_INVALID_ARG;
    hipStream_t s = (hipStream_t)stream;
    if (P->nitems == 0) {
        // empty list: x * 0 / emap, k_tile_update's operations with a zero sum
        const int64_t nv = plan_voxels(P);
        hipLaunchKernelGGL(k_update_empty, dim3((unsigned)((nv + 255) / 256)), dim3(256), 0, s, emap, zero_guard, x,
                           nv);
        return launch_status();
    }
    mark_before(P, s);
    PLAN_DISPATCH(k_bp_tiled, <<<P->nitems, TILE_THREADS, 0, s>>>(P->recs.p, P->items.p, P->order_bp.p, values,
                                                                   nullptr, nullptr, 0, 0.0f, 0, P->G, P->T,
                                                                   P->rec_ev.p, P->tile_buf.p));
    mark_after(P, s);
    dim3 grid((TVOX + 255) / 256, (unsigned)P->T.ntiles);
    hipLaunchKernelGGL(k_tile_update, grid, dim3(256), 0, s, P->tile_buf.p, P->tile_items.p, P->G, P->T, emap,
                       zero_guard, x);
    return launch_status();
}

int srf_plan_backproject_ratio(srf_plan *P, const float *y, const float *p, srf_ratio_kind_t kind, float eps,
                               srf_bp_mode_t mode, int accumulate, float *image, void *stream)
{
    if (P && P->n > 0 && (!y || !p)) return SRF_ERR_INVALID_ARG;
    if (kind != SRF_RATIO_PLAIN && kind != SRF_RATIO_DIV_NO_NAN) return SRF_ERR_INVALID_ARG;
    return plan_bp(P, nullptr, y, p, (int)kind, eps, mode, accumulate, image, (hipStream_t)stream);
}

}  // extern "C"

// Scans whose plans use different kernel variants (TOF, voxel aspect) share a
// batch: the work list is grouped by variant and each group is one launch.
constexpr int NVARIANT = 8;   // tof << 2 | ds << 1 | dz

struct srf_batch {
    int nscan;
    ScanDesc *desc;
    Work *work;
    int nwork;
    int64_t *ev_base;
    int *tile_base;
    unsigned *counter;   // [2 g]: projector queue of group g, [2 g + 1]: backprojector queue
    int64_t total_ev;
    int total_tiles;
    int persistent;
    int grid;            // largest group's grid
    int g_start[NVARIANT], g_count[NVARIANT], g_grid[NVARIANT];
    size_t bytes;
};

static void batch_free(srf_batch *B)
{
    if (!B) return;
    (void)hipFree(B->desc);
    (void)hipFree(B->work);
    (void)hipFree(B->ev_base);
    (void)hipFree(B->tile_base);
    (void)hipFree(B->counter);
    delete B;
}

#define BATCH_DISPATCH(V, KERNEL, ...)                                                               \
    do {                                                                                             \
        switch (V) {                                                                                 \
        case 0: KERNEL<false, false, false> __VA_ARGS__; break;                                      \
        case 1: KERNEL<false, false, true> __VA_ARGS__; break;                                       \
        case 2: KERNEL<false, true, false> __VA_ARGS__; break;                                       \
        case 3: KERNEL<false, true, true> __VA_ARGS__; break;                                        \
        case 4: KERNEL<true, false, false> __VA_ARGS__; break;                                       \
        case 5: KERNEL<true, false, true> __VA_ARGS__; break;                                        \
        case 6: KERNEL<true, true, false> __VA_ARGS__; break;                                        \
        default: KERNEL<true, true, true> __VA_ARGS__; break;                                        \
        }                                                                                            \
    } while (0)

static int plan_variant(const srf_plan *P) { return (P->tof ? 4 : 0) | (P->ds ? 2 : 0) | (P->dz ? 1 : 0); }

extern "C" {

int srf_batch_create(srf_plan *const *plans, const srf_batch_scan_t *scans, int32_t n, const srf_batch_cfg_t *cfg,
                     srf_batch **out, void *stream)
{
    if (!out) return SRF_ERR_INVALID_ARG;
    *out = nullptr;
    if (!plans || !scans || !cfg || n < 1) return SRF_ERR_INVALID_ARG;
    if ((cfg->persistent != 0 && cfg->persistent != 1) || cfg->workgroups_per_cu < 1 ||
        cfg->workgroups_per_cu > 16 || (cfg->order != 0 && cfg->order != 1))
        return SRF_ERR_INVALID_ARG;
    for (int i = 0; i < n; ++i) {
        const srf_plan *P = plans[i];
        if (!P || !P->valid || !scans[i].image || !scans[i].emap) return SRF_ERR_INVALID_ARG;
        if (P->n > 0 && (!scans[i].y || !scans[i].q)) return SRF_ERR_INVALID_ARG;
    }
    hipStream_t s = (hipStream_t)stream;
    srf_batch *B = new srf_batch();
    memset(B, 0, sizeof(*B));
    B->nscan = n;
    B->persistent = cfg->persistent;
    auto cleanup = [&]() { batch_free(B); };
    std::vector<ScanDesc> desc((size_t)n);
    std::vector<int64_t> ev_base((size_t)n + 1, 0);
    std::vector<int> tile_base((size_t)n + 1, 0);
    std::vector<Work> work;
    std::vector<int64_t> cost;
    std::vector<int> variant;
    for (int i = 0; i < n; ++i) {
        const srf_plan *P = plans[i];
        ScanDesc &d = desc[(size_t)i];
        d.G = P->G;
        d.T = P->T;
        d.recs = P->recs.p;
        d.items = P->items.p;
        d.rec_ev = P->rec_ev.p;
        d.tile_items = P->tile_items.p;
        d.ev_off = P->ev_off.p;
        d.ev_flag = P->ev_flag.p;
        d.partial = P->partial.p;
        d.tile_buf = P->tile_buf.p;
        d.x = scans[i].image;
        d.emap = scans[i].emap;
        d.y = scans[i].y;
        d.q = scans[i].q;
        d.n = P->n;
        ev_base[(size_t)i + 1] = ev_base[(size_t)i] + P->n;
        tile_base[(size_t)i + 1] = tile_base[(size_t)i] + P->T.ntiles;
        if (P->nitems > 0) {
            std::vector<Item> items((size_t)P->nitems);
            SRF_HIP(hipMemcpyAsync(items.data(), P->items.p, items.size() * sizeof(Item), hipMemcpyDeviceToHost, s));
            SRF_HIP(hipStreamSynchronize(s));
            for (int k = 0; k < P->nitems; ++k) {
                work.push_back(Work{i, k});
                cost.push_back(item_work(items[(size_t)k]));
                variant.push_back(plan_variant(P));
            }
        }
    }
    {
        // grouped by kernel variant; within a group in scan order, or longest
        // items first (the tail of the launch is then made of short items)
        std::vector<size_t> idx(work.size());
        for (size_t k = 0; k < idx.size(); ++k) idx[k] = k;
        std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) {
            if (variant[a] != variant[b]) return variant[a] < variant[b];
            return cfg->order == 1 && cost[a] > cost[b];
        });
        std::vector<Work> sorted(work.size());
        for (size_t k = 0; k < idx.size(); ++k) sorted[k] = work[idx[k]];
        for (int g = 0; g < NVARIANT; ++g) B->g_start[g] = B->g_count[g] = B->g_grid[g] = 0;
        for (size_t k = 0; k < idx.size(); ++k) {
            const int g = variant[idx[k]];
            if (B->g_count[g] == 0) B->g_start[g] = (int)k;
            ++B->g_count[g];
        }
        work.swap(sorted);
    }
    B->nwork = (int)work.size();
    B->total_ev = ev_base[(size_t)n];
    B->total_tiles = tile_base[(size_t)n];
    int dev = 0, cus = 0;
    SRF_HIP(hipGetDevice(&dev));
    SRF_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    B->grid = 0;
    for (int g = 0; g < NVARIANT; ++g) {
        const int c = B->g_count[g];
        B->g_grid[g] = c == 0 ? 0 : (cfg->persistent ? std::max(1, std::min(c, cus * cfg->workgroups_per_cu)) : c);
        B->grid = std::max(B->grid, B->g_grid[g]);
    }
    SRF_TRY(dmalloc(&B->desc, desc.size(), B->bytes));
    SRF_TRY(dmalloc(&B->work, std::max<size_t>(1, work.size()), B->bytes));
    SRF_TRY(dmalloc(&B->ev_base, ev_base.size(), B->bytes));
    SRF_TRY(dmalloc(&B->tile_base, tile_base.size(), B->bytes));
    SRF_TRY(dmalloc(&B->counter, 2 * NVARIANT, B->bytes));
    SRF_HIP(hipMemcpyAsync(B->desc, desc.data(), desc.size() * sizeof(ScanDesc), hipMemcpyHostToDevice, s));
    if (!work.empty())
        SRF_HIP(hipMemcpyAsync(B->work, work.data(), work.size() * sizeof(Work), hipMemcpyHostToDevice, s));
    SRF_HIP(hipMemcpyAsync(B->ev_base, ev_base.data(), ev_base.size() * sizeof(int64_t), hipMemcpyHostToDevice, s));
    SRF_HIP(hipMemcpyAsync(B->tile_base, tile_base.data(), tile_base.size() * sizeof(int), hipMemcpyHostToDevice,
                           s));
    SRF_HIP(hipStreamSynchronize(s));
    *out = B;
    return SRF_OK;
}

int srf_batch_destroy(srf_batch *batch)
{
    batch_free(batch);
    return SRF_OK;
}

int srf_batch_info(const srf_batch *B, int32_t *n_scans, int32_t *n_work, int32_t *grid, int64_t *n_events,
                   size_t *device_bytes)
{
    if (!B) return SRF_ERR_INVALID_ARG;
    if (n_scans) *n_scans = B->nscan;
    if (n_work) *n_work = B->nwork;
    if (grid) *grid = B->grid;
    if (n_events) *n_events = B->total_ev;
    if (device_bytes) *device_bytes = B->bytes;
    return SRF_OK;
}

int srf_batch_project_ratio(srf_batch *B, srf_ratio_kind_t kind, float eps, void *stream)
{
    if (!B) return SRF_ERR_INVALID_ARG;
    if (kind != SRF_RATIO_PLAIN && kind != SRF_RATIO_DIV_NO_NAN) return SRF_ERR_INVALID_ARG;
    hipStream_t s = (hipStream_t)stream;
    if (B->persistent && B->nwork > 0)
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(256), 0, s, B->counter, 0u, (int64_t)(2 * NVARIANT));
    for (int g = 0; g < NVARIANT; ++g) {
        if (B->g_count[g] == 0) continue;
        BATCH_DISPATCH(g, k_fp_batch, <<<B->g_grid[g], TILE_THREADS, 0, s>>>(B->desc, B->work + B->g_start[g],
                                                                             B->g_count[g], B->counter + 2 * g,
                                                                             B->persistent));
    }
    if (B->total_ev > 0)
        hipLaunchKernelGGL(k_seg_reduce_batch, dim3((unsigned)((B->total_ev + 255) / 256)), dim3(256), 0, s, B->desc,
                           B->ev_base, B->nscan, (int)kind, eps);
    return launch_status();
}

int srf_batch_backproject_update(srf_batch *B, float zero_guard, void *stream)
{
    if (!B) return SRF_ERR_INVALID_ARG;
    hipStream_t s = (hipStream_t)stream;
    if (B->persistent && B->nwork > 0)
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(256), 0, s, B->counter, 0u, (int64_t)(2 * NVARIANT));
    for (int g = 0; g < NVARIANT; ++g) {
        if (B->g_count[g] == 0) continue;
        BATCH_DISPATCH(g, k_bp_batch, <<<B->g_grid[g], TILE_THREADS, 0, s>>>(B->desc, B->work + B->g_start[g],
                                                                             B->g_count[g], B->counter + 2 * g + 1,
                                                                             B->persistent));
    }
    if (B->total_tiles > 0)
        hipLaunchKernelGGL(k_tile_update_batch, dim3((unsigned)B->total_tiles * (TVOX / 256)), dim3(256), 0, s,
                           B->desc, B->tile_base, B->nscan, zero_guard);
    return launch_status();
}

int srf_batch_em_step(srf_batch *B, srf_ratio_kind_t kind, float eps, float zero_guard, void *stream)
{
    int rc = srf_batch_project_ratio(B, kind, eps, stream);
    if (rc) return rc;
    return srf_batch_backproject_update(B, zero_guard, stream);
}

}  // extern "C"

