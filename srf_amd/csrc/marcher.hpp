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
// Structure (what differs from the reference, and why):
//  * one event per lane, no grid-stride loop: launches carry >> 256 workgroups;
//  * the slab loop is clipped to the slabs whose secondary coordinates can lie
//    inside the image (the reference walks all nx or ny slabs, siddon.cu.cc:51);
//    the window is widened by a margin so the exact per-slab tests below still
//    decide every term;
//  * the secondary/axial boundary values are computed once per slab boundary and
//    carried to the next slab (the reference evaluates both ends per slab; the
//    values are identical because pp2(ip) == pp1(ip+1) exactly);
//  * the reference's nested branch tree (siddon.cu.cc:86-189) is flattened into
//    three predicated terms per slab, so divergent lanes of a wavefront issue the
//    same instruction stream.
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

// Interval of primary coordinate x on which value(x) = (c + k*(x - p1))/d + n2
// lies in [lo_v, hi_v]; empty => a > b.
__device__ __forceinline__ void axis_window(double c, double k, double p1, double d, double n2,
                                            double lo_v, double hi_v, double &a, double &b)
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

// March one event.  Act must provide
//   void term(int voxel, int branch, float weight)   weight = full term weight
//   void nan_event()                                  non-finite endpoints
// TOF: siddon_tof family; VAR: V_FP (projector predicates, 1/L^2 weight) or
// V_BP (backprojector predicates).  DIV_S / DIV_Z: the secondary (dy) / axial
// (dz) divisors differ from 1.0f (x / 1.0f == x, so the division is skipped
// when the host proved the divisor is exactly 1).
template <bool TOF, int VAR, bool DIV_S, bool DIV_Z, class Act>
__device__ __forceinline__ void march_event(const float *__restrict__ lor, float tof_value,
                                            const LaunchGeom &G, Act &act)
{
    const float x1 = (lor[0] - G.cx) / G.dx_;
    const float y1 = (lor[1] - G.cy) / G.dx_;
    const float z1 = (lor[2] - G.cz) / G.dx_;
    const float x2 = (lor[3] - G.cx) / G.dx_;
    const float y2 = (lor[4] - G.cy) / G.dx_;
    const float z2 = (lor[5] - G.cz) / G.dx_;
    if (!(isfinite(x1) && isfinite(y1) && isfinite(z1) && isfinite(x2) && isfinite(y2) &&
          isfinite(z2))) {
        act.nan_event();
        return;
    }
    const float xd = x2 - x1;
    const float yd = y2 - y1;
    const float zd = z2 - z1;
    const float t2d = xd * xd + yd * yd;
    const float tlen = sqrtf(t2d);
    // siddon.cu.cc:34, siddon_tof.cu.cc:37 (10-voxel transaxial cut)
    if (tlen < 10.0f) return;
    // siddon.cu.cc:38 (non-TOF projector only)
    if (!TOF && VAR == V_FP) {
        if (tlen < fabsf(zd)) return;
    }
    const float L = sqrtf(t2d + zd * zd);
    const float axd = fabsf(xd), ayd = fabsf(yd);
    // siddon.cu.cc:46 (>=, non-TOF projector) vs :372 and siddon_tof.cu.cc:48,405 (>)
    const bool xdom = (!TOF && VAR == V_FP) ? (axd >= ayd) : (axd > ayd);

    float p1, p2, sec1, ks, kz, ns2, np2, pd_tof;
    int np, ns, stride_p, stride_s;
    if (xdom) {
        p1 = x1; p2 = x2; sec1 = y1;
        ks = yd / xd; kz = zd / xd;
        np = G.nx; ns = G.ny; np2 = G.nx2; ns2 = G.ny2;
        stride_p = 1; stride_s = G.nx;
        pd_tof = 1.0f;             // dx
    } else {
        p1 = y1; p2 = y2; sec1 = x1;
        ks = xd / yd; kz = zd / yd;
        np = G.ny; ns = G.nx; np2 = G.ny2; ns2 = G.nx2;
        stride_p = G.nx; stride_s = 1;
        pd_tof = G.dy;             // siddon_tof.cu.cc:215 yc uses dy
    }
    // secondary divisor: dy on the x branch, none on the y branch (siddon.cu.cc:207)
    const float ds = xdom ? G.dy : 1.0f;
    const int stride_z = G.nx * G.ny;

    // per-event weight constant: sqrt(1+k^2+kz^2) * dx_ [/ L / L] [* tof norm]
    const float sq = sqrtf(1.0f + ks * ks + kz * kz);
    float W = sq * G.dx_;
    if (VAR == V_FP) W = W / L / L;
    float ratio1 = 0.0f, inv_pd = 0.0f;
    if (TOF) {
        // siddon_tof.cu.cc:44 (double, as the reference)
        ratio1 = (float)((1.0 - (double)(tof_value / G.dx_ / L)) / 2);
        inv_pd = 1.0f / (p2 - p1);
        W = W * G.tof_c0;
    }

    // ---- clip the slab range (conservative; exact tests below decide) -------
    int lo, hi;
    {
        const double m = 0.01;
        double sa, sb, za, zb;
        axis_window((double)sec1, (double)ks, (double)p1, (double)ds, (double)ns2, -1.0 - m,
                    (double)ns + m, sa, sb);
        axis_window((double)z1, (double)kz, (double)p1, (double)G.dz, (double)G.nz2, -1.0 - m,
                    (double)G.nz + m, za, zb);
        double xl = fmax(fmax(-(double)np2, sa - 1.0), za - 1.0);
        double xh = fmin(fmin((double)np2, sb + 1.0), zb + 1.0);
        if (xl > xh) {
            lo = 0; hi = 0;
        } else {
            lo = max(0, (int)floor(xl + (double)np2) - 1);
            hi = min(np, (int)ceil(xh + (double)np2) + 1);
        }
    }

    const bool ks_pos = ks >= 0.0f;
    const bool kz_pos = kz >= 0.0f;
    float pp1 = (float)lo - np2;
    float S0, Z0;
    {
        float s = sec1 + ks * (pp1 - p1);
        S0 = (DIV_S ? s / ds : s) + ns2;
        float z = z1 + kz * (pp1 - p1);
        Z0 = (DIV_Z ? z / G.dz : z) + G.nz2;
    }
    for (int ip = lo; ip < hi; ++ip) {
        const float pp2 = pp1 + 1.0f;
        float S1, Z1;
        {
            float s = sec1 + ks * (pp2 - p1);
            S1 = (DIV_S ? s / ds : s) + ns2;
            float z = z1 + kz * (pp2 - p1);
            Z1 = (DIV_Z ? z / G.dz : z) + G.nz2;
        }
        bool cut = false;
        if (VAR == V_FP) {
            if (!TOF) {
                // siddon.cu.cc:56-57 / 203-204
                cut = (p1 < p2 && pp2 > p2) || (p1 >= p2 && pp1 < p2);
            } else if (xdom) {
                // siddon_tof.cu.cc:67-74
                cut = (p2 > p1 && pp1 <= p1) || (p2 < p1 && pp2 >= p1);
            } else {
                // siddon_tof.cu.cc:228-235
                cut = (p2 > p1 && pp1 < p1) || (p2 < p1 && pp2 > p1);
            }
        }
        const float sl = ks_pos ? S0 : S1, sh = ks_pos ? S1 : S0;
        const float zl = kz_pos ? Z0 : Z1, zh = kz_pos ? Z1 : Z0;
        const int cs1 = floor_to_int(sl), cs2 = floor_to_int(sh);
        const int cz1 = floor_to_int(zl), cz2 = floor_to_int(zh);
        const bool bs = cs1 != cs2, bz = cz1 != cz2;
        const bool ok = !cut && cs1 >= (bs ? -1 : 0) && cs1 < ns && cz1 >= (bz ? -1 : 0) &&
                        cz1 < G.nz;
        if (ok) {
            float wg = W;
            if (TOF) {
                // siddon_tof.cu.cc:55-59: Gaussian at the slab centre on the primary axis
                const float pc = ((float)(ip - np / 2) + 0.5f) * pd_tof;
                const float d = ((pc - p1) * inv_pd - ratio1) * L;
                wg = W * __builtin_amdgcn_exp2f(d * d * G.tof_c1);
            }
            // exact quotients: the rs > rz comparison picks a voxel
            const float rs = bs ? ((float)cs2 - sl) / (sh - sl) : 1.0f;
            const float rz = bz ? ((float)cz2 - zl) / (zh - zl) : 1.0f;
            const bool sgt = rs > rz;
            const int base = ip * stride_p;
            // term 1: (cs1, cz1)
            {
                const float r1 = bs ? (bz ? (sgt ? rz : rs) : rs) : (bz ? rz : 1.0f);
                const int br = bs ? (bz ? (sgt ? 5 : 8) : 3) : (bz ? 1 : 0);
                if (cs1 >= 0 && cz1 >= 0) act.term(base + cs1 * stride_s + cz1 * stride_z, br, r1 * wg);
            }
            // term 2: both secondary axes crossed
            if (bs && bz) {
                const int c_s = sgt ? cs1 : cs2, c_z = sgt ? cz2 : cz1;
                const bool c2 = sgt ? (cs1 >= 0 && cz2 < G.nz) : (cs2 < ns && cz1 >= 0);
                const float r2 = sgt ? (rs - rz) : (rz - rs);
                if (c2) act.term(base + c_s * stride_s + c_z * stride_z, sgt ? 6 : 9, r2 * wg);
            }
            // term 3: (cs2, cz2)
            if (bs || bz) {
                const float r3 = bs ? (bz ? (sgt ? 1.0f - rs : 1.0f - rz) : 1.0f - rs) : 1.0f - rz;
                const int br = bs ? (bz ? (sgt ? 7 : 10) : 4) : 2;
                if (cs2 < ns && cz2 < G.nz) act.term(base + cs2 * stride_s + cz2 * stride_z, br, r3 * wg);
            }
        }
        pp1 = pp2;
        S0 = S1;
        Z0 = Z1;
    }
}

}  // namespace srf
