// marcher.hpp -- the device ray marcher shared by every projector/backprojector
// kernel of libsrf_hip.so (gfx950).
//
// Semantics: the Siddon-type slab marcher of srfnef/cc/siddon/siddon.cu.cc:12-656
// and srfnef/cc/siddon_tof/siddon_tof.cu.cc:14-707.  Every expression that
// decides WHICH voxel a term lands in is evaluated with the reference's own
// float operations in the reference's order (this translation unit is built
// with -ffp-contract=off; HIP's default float division and sqrt are correctly
// rounded), so voxel indices and branch choices are bit-identical to an
// un-contracted IEEE build of the reference.  Weights are re-associated
// (per-event constants hoisted, TOF Gaussian evaluated in fp32 through
// v_exp_f32) and agree to a few ulp.
//
// Pieces:
//   ray_setup     endpoints -> voxel units, early-out tests, L, dominance tests
//   make_branch   the primary/secondary description of one reference branch
//                 (x branch siddon.cu.cc:46-191, y branch :192-335)
//   clip_range    slabs whose secondary/axial coordinates can lie in the image
//                 (conservative; the exact per-slab tests still decide)
//   fp_cut_range  the projector's segment cuts (siddon.cu.cc:56-57,203-204;
//                 siddon_tof.cu.cc:67-74,228-235) as an exact slab interval
//   march_slabs   the slab loop: carried boundary values, flattened branch tree,
//                 up to three predicated terms per slab handed to a sink
//   march_event   all of the above for one event (one-event-per-lane kernels)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace srf {

enum : int { V_FP = 0, V_BP = 1 };

// Per-launch constants.  Computed on the host with the same float expressions
// as siddon.cu.cc:19-22,39-41 and siddon_tof.cu.cc:46-47,57 (bit-identical).
struct LaunchGeom {
    int nx, ny, nz;
    float cx, cy, cz;
    float dx_;            // sx / nx
    float dy, dz;         // sy / ny / dx_, sz / nz / dx_
    float nx2, ny2, nz2;  // n / 2.0f
    float tof_bin_;       // tof_bin / dx_
    float tof_sigma2_;    // tof_sigma2 / dx_ / dx_
    float tse;            // tof_sigma2_ + tof_bin_^2 / 12
    float tof_c0;         // (float)(tof_bin_ / sqrt(2*PI*tse))   (double, as the reference)
    float tof_c1;         // -0.5 * log2(e) / tse                 (exp(-t2/2) == exp2(c1 * d^2))
};

__device__ __forceinline__ int floor_to_int(float v)
{
    // v_floor_f32 + v_cvt_i32_f32: NaN -> 0, saturating (the oracle mirrors it)
    return (int)floorf(v);
}

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Branch-independent part of one event.
struct RaySetup {
    float x1, y1, z1, x2, y2, z2;
    float xd, yd, zd, L;
    float ratio1;     // TOF position of the emission point along p1->p2 (siddon_tof.cu.cc:44)
    bool nan;         // non-finite endpoint: projection NaN, backprojection skips
    bool live;        // passes the 10-voxel transaxial cut (siddon.cu.cc:34)
    bool q2;          // non-TOF projector's axial cut fires (siddon.cu.cc:38)
    bool xdom_ge;     // |xd| >= |yd|  (non-TOF projector, siddon.cu.cc:46)
    bool xdom_gt;     // |xd| >  |yd|  (everything else)
};

template <bool TOF>
__device__ __forceinline__ void ray_setup(const float *__restrict__ lor, float tof_value, const LaunchGeom &G,
                                          RaySetup &r)
{
    r.x1 = (lor[0] - G.cx) / G.dx_;
    r.y1 = (lor[1] - G.cy) / G.dx_;
    r.z1 = (lor[2] - G.cz) / G.dx_;
    r.x2 = (lor[3] - G.cx) / G.dx_;
    r.y2 = (lor[4] - G.cy) / G.dx_;
    r.z2 = (lor[5] - G.cz) / G.dx_;
    r.nan = !(isfinite(r.x1) && isfinite(r.y1) && isfinite(r.z1) && isfinite(r.x2) && isfinite(r.y2) &&
              isfinite(r.z2));
    r.xd = r.x2 - r.x1;
    r.yd = r.y2 - r.y1;
    r.zd = r.z2 - r.z1;
    const float t2d = r.xd * r.xd + r.yd * r.yd;
    const float tlen = sqrtf(t2d);
    r.live = !r.nan && !(tlen < 10.0f);
    r.q2 = tlen < fabsf(r.zd);
    r.L = sqrtf(t2d + r.zd * r.zd);
    const float axd = fabsf(r.xd), ayd = fabsf(r.yd);
    r.xdom_ge = axd >= ayd;
    r.xdom_gt = axd > ayd;
    r.ratio1 = 0.0f;
    if (TOF) r.ratio1 = (float)((1.0 - (double)(tof_value / G.dx_ / r.L)) / 2);
}

// One primary-axis branch.
struct Branch {
    bool xdom;
    float p1, p2, sec1, ks, kz;
    float ds;              // secondary divisor: dy (x branch) or 1 (y branch; x / 1.0f == x)
    float ns2, np2;
    float pd_tof;          // dx (=1) on the x branch, dy on the y branch (siddon_tof.cu.cc:55,215)
    int np, ns, stride_p, stride_s;
    float sq;              // sqrt(1 + ks^2 + kz^2)
};

__device__ __forceinline__ void make_branch(const RaySetup &r, bool xdom, const LaunchGeom &G, Branch &B)
{
    B.xdom = xdom;
    if (xdom) {
        B.p1 = r.x1; B.p2 = r.x2; B.sec1 = r.y1;
        B.ks = r.yd / r.xd; B.kz = r.zd / r.xd;
        B.np = G.nx; B.ns = G.ny; B.np2 = G.nx2; B.ns2 = G.ny2;
        B.stride_p = 1; B.stride_s = G.nx;
        B.pd_tof = 1.0f;
        B.ds = G.dy;
    } else {
        B.p1 = r.y1; B.p2 = r.y2; B.sec1 = r.x1;
        B.ks = r.xd / r.yd; B.kz = r.zd / r.yd;
        B.np = G.ny; B.ns = G.nx; B.np2 = G.ny2; B.ns2 = G.nx2;
        B.stride_p = G.nx; B.stride_s = 1;
        B.pd_tof = G.dy;
        B.ds = 1.0f;
    }
    B.sq = sqrtf(1.0f + B.ks * B.ks + B.kz * B.kz);
}

// Term weight constant of a branch: sqrt(1+k^2+kz^2)*dx_ [/L/L] [* tof norm]
// (siddon.cu.cc:94 projector, :418 backprojector).
template <bool TOF>
__device__ __forceinline__ float branch_weight(const Branch &B, const RaySetup &r, const LaunchGeom &G, bool fp)
{
    float W = B.sq * G.dx_;
    if (fp) W = W / r.L / r.L;
    if (TOF) W = W * G.tof_c0;
    return W;
}

// Interval of primary coordinate x on which value(x) = (c + k*(x - p1))/d + n2
// lies in [lo_v, hi_v]; empty => a > b.
__device__ __forceinline__ void axis_window(double c, double k, double p1, double d, double n2, double lo_v,
                                            double hi_v, double &a, double &b)
{
    if (k == 0.0) {
        double v = c / d + n2;
        if (v >= lo_v && v <= hi_v) { a = -1e30; b = 1e30; }
        else { a = 1e30; b = -1e30; }
    } else {
        double xa = p1 + ((lo_v - n2) * d - c) / k;
        double xb = p1 + ((hi_v - n2) * d - c) / k;
        a = fmin(xa, xb);
        b = fmax(xa, xb);
    }
}

// Slabs [lo, hi) that can hold a term: the secondary and axial coordinates of
// a slab must reach [-1, n] (widened by a margin far above float round-off, and
// by one slab on each side).
__device__ __forceinline__ void clip_range(const Branch &B, const RaySetup &r, const LaunchGeom &G, int &lo, int &hi)
{
    const double m = 0.01;
    double sa, sb, za, zb;
    axis_window((double)B.sec1, (double)B.ks, (double)B.p1, (double)B.ds, (double)B.ns2, -1.0 - m,
                (double)B.ns + m, sa, sb);
    axis_window((double)r.z1, (double)B.kz, (double)B.p1, (double)G.dz, (double)G.nz2, -1.0 - m,
                (double)G.nz + m, za, zb);
    const double xl = fmax(fmax(-(double)B.np2, sa - 1.0), za - 1.0);
    const double xh = fmin(fmin((double)B.np2, sb + 1.0), zb + 1.0);
    if (xl > xh) {
        lo = 0; hi = 0;
    } else {
        lo = max(0, (int)floor(xl + (double)B.np2) - 1);
        hi = min(B.np, (int)ceil(xh + (double)B.np2) + 1);
    }
}

// (int) of a double clamped far inside the int range.
__device__ __forceinline__ int dbl_to_int(double v)
{
    return (int)fmin(fmax(v, -1.0e9), 1.0e9);
}

// First ip in [lo, hi] with keep(ip) true for a predicate that is false then
// true along ip; `guess` only sets where the exact search starts.
template <class P>
__device__ __forceinline__ int first_true(P keep, int lo, int hi, int guess)
{
    int ip = min(max(guess, lo), hi);
    while (ip > lo && keep(ip - 1)) --ip;
    while (ip < hi && !keep(ip)) ++ip;
    return ip;
}

// The projector's segment cut as the exact slab interval [a, b) within [lo, hi).
// Each reference predicate is monotone in the slab index; the interval ends are
// located with the reference's own float comparisons (slab boundaries
// pp1 = ip - np2 and pp2 = pp1 + 1 are exact in float).
template <bool TOF>
__device__ __forceinline__ void fp_cut_range(const Branch &B, int lo, int hi, int &a, int &b)
{
    const float p1 = B.p1, p2 = B.p2, np2 = B.np2;
    a = lo;
    b = hi;
    if (!TOF) {
        // siddon.cu.cc:56-57 / 203-204: skip (p1<p2 && pp2>p2) || (p1>=p2 && pp1<p2)
        if (p1 < p2) {
            auto cut = [&](int ip) { return ((float)ip - np2) + 1.0f > p2; };
            b = first_true(cut, lo, hi, dbl_to_int(floor((double)p2 + (double)np2)));
        } else {
            auto keep = [&](int ip) { return (float)ip - np2 >= p2; };
            a = first_true(keep, lo, hi, dbl_to_int(ceil((double)p2 + (double)np2)));
        }
    } else if (B.xdom) {
        // siddon_tof.cu.cc:67-74: skip (p2>p1 && pp1<=p1) || (p2<p1 && pp2>=p1)
        if (p2 > p1) {
            auto keep = [&](int ip) { return (float)ip - np2 > p1; };
            a = first_true(keep, lo, hi, dbl_to_int(floor((double)p1 + (double)np2)) + 1);
        } else if (p2 < p1) {
            auto cut = [&](int ip) { return ((float)ip - np2) + 1.0f >= p1; };
            b = first_true(cut, lo, hi, dbl_to_int(ceil((double)p1 + (double)np2)) - 1);
        }
    } else {
        // siddon_tof.cu.cc:228-235: skip (p2>p1 && pp1<p1) || (p2<p1 && pp2>p1)
        if (p2 > p1) {
            auto keep = [&](int ip) { return (float)ip - np2 >= p1; };
            a = first_true(keep, lo, hi, dbl_to_int(ceil((double)p1 + (double)np2)));
        } else if (p2 < p1) {
            auto cut = [&](int ip) { return ((float)ip - np2) + 1.0f > p1; };
            b = first_true(cut, lo, hi, dbl_to_int(floor((double)p1 + (double)np2)));
        }
    }
    if (b < a) b = a;
}

// TOF Gaussian parameters of one branch: d(ip) = ((pc - p1) * inv_pd - ratio1) * L
struct TofLine {
    float inv_pd, ratio1, L;
};

// The slab loop over [a, b).  Act provides
//   void term(int ip, int cs, int cz, int branch, float weight)
// with (ip, cs, cz) the primary slab, secondary and axial cell of the term.
template <bool TOF, bool DIV_S, bool DIV_Z, class Act>
__device__ __forceinline__ void march_slabs(const Branch &B, float z1, const LaunchGeom &G, int a, int b, float W,
                                            const TofLine &T, Act &act)
{
    const bool ks_pos = B.ks >= 0.0f;
    const bool kz_pos = B.kz >= 0.0f;
    const int ns = B.ns, nz = G.nz;
    float pp1 = (float)a - B.np2;
    float S0, Z0;
    {
        float s = B.sec1 + B.ks * (pp1 - B.p1);
        S0 = (DIV_S ? s / B.ds : s) + B.ns2;
        float z = z1 + B.kz * (pp1 - B.p1);
        Z0 = (DIV_Z ? z / G.dz : z) + G.nz2;
    }
    for (int ip = a; ip < b; ++ip) {
        const float pp2 = pp1 + 1.0f;
        float S1, Z1;
        {
            float s = B.sec1 + B.ks * (pp2 - B.p1);
            S1 = (DIV_S ? s / B.ds : s) + B.ns2;
            float z = z1 + B.kz * (pp2 - B.p1);
            Z1 = (DIV_Z ? z / G.dz : z) + G.nz2;
        }
        const float sl = ks_pos ? S0 : S1, sh = ks_pos ? S1 : S0;
        const float zl = kz_pos ? Z0 : Z1, zh = kz_pos ? Z1 : Z0;
        const int cs1 = floor_to_int(sl), cs2 = floor_to_int(sh);
        const int cz1 = floor_to_int(zl), cz2 = floor_to_int(zh);
        const bool bs = cs1 != cs2, bz = cz1 != cz2;
        // top-level tests of siddon.cu.cc:86-189 (x) / 231-333 (y)
        const bool ok = cs1 >= (bs ? -1 : 0) && cs1 < ns && cz1 >= (bz ? -1 : 0) && cz1 < nz;
        if (ok) {
            float wg = W;
            if (TOF) {
                // siddon_tof.cu.cc:55-59: Gaussian at the slab centre on the primary axis
                const float pc = ((float)(ip - B.np / 2) + 0.5f) * B.pd_tof;
                const float d = ((pc - B.p1) * T.inv_pd - T.ratio1) * T.L;
                wg = W * __builtin_amdgcn_exp2f(d * d * G.tof_c1);
            }
            // exact quotients: the rs > rz comparison picks a voxel
            const float rs = bs ? ((float)cs2 - sl) / (sh - sl) : 1.0f;
            const float rz = bz ? ((float)cz2 - zl) / (zh - zl) : 1.0f;
            const bool sgt = rs > rz;
            {
                const float r1 = bs ? (bz ? (sgt ? rz : rs) : rs) : (bz ? rz : 1.0f);
                const int br = bs ? (bz ? (sgt ? 5 : 8) : 3) : (bz ? 1 : 0);
                if (cs1 >= 0 && cz1 >= 0) act.term(ip, cs1, cz1, br, r1 * wg);
            }
            if (bs && bz) {
                const int c_s = sgt ? cs1 : cs2, c_z = sgt ? cz2 : cz1;
                const bool c2 = sgt ? (cs1 >= 0 && cz2 < nz) : (cs2 < ns && cz1 >= 0);
                const float r2 = sgt ? (rs - rz) : (rz - rs);
                if (c2) act.term(ip, c_s, c_z, sgt ? 6 : 9, r2 * wg);
            }
            if (bs || bz) {
                const float r3 = bs ? (bz ? (sgt ? 1.0f - rs : 1.0f - rz) : 1.0f - rs) : 1.0f - rz;
                const int br = bs ? (bz ? (sgt ? 7 : 10) : 4) : 2;
                if (cs2 < ns && cz2 < nz) act.term(ip, cs2, cz2, br, r3 * wg);
            }
        }
        pp1 = pp2;
        S0 = S1;
        Z0 = Z1;
    }
}

// Adapts a flat-index sink  void term(int voxel, int branch, float w)  to march_slabs.
template <class Flat>
struct FlatIndex {
    Flat &f;
    int stride_p, stride_s, stride_z;
    __device__ __forceinline__ void term(int ip, int cs, int cz, int br, float w)
    {
        f.term(ip * stride_p + cs * stride_s + cz * stride_z, br, w);
    }
};

// One event, as one reference device function.  Act must provide
//   void term(int voxel, int branch, float weight)   (flat kernel-layout index)
//   void nan_event()                                  (non-finite endpoints)
template <bool TOF, int VAR, bool DIV_S, bool DIV_Z, class Act>
__device__ __forceinline__ void march_event(const float *__restrict__ lor, float tof_value, const LaunchGeom &G,
                                            Act &act)
{
    RaySetup r;
    ray_setup<TOF>(lor, tof_value, G, r);
    if (r.nan) {
        act.nan_event();
        return;
    }
    if (!r.live) return;
    if (!TOF && VAR == V_FP && r.q2) return;
    const bool xdom = (!TOF && VAR == V_FP) ? r.xdom_ge : r.xdom_gt;
    Branch B;
    make_branch(r, xdom, G, B);
    const float W = branch_weight<TOF>(B, r, G, VAR == V_FP);
    int lo, hi;
    clip_range(B, r, G, lo, hi);
    int a = lo, b = hi;
    if (VAR == V_FP) fp_cut_range<TOF>(B, lo, hi, a, b);
    TofLine T{0.0f, r.ratio1, r.L};
    if (TOF) T.inv_pd = 1.0f / (B.p2 - B.p1);
    FlatIndex<Act> fi{act, B.stride_p, B.stride_s, G.nx * G.ny};
    march_slabs<TOF, DIV_S, DIV_Z>(B, r.z1, G, a, b, W, T, fi);
}

}  // namespace srf
