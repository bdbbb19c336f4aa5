/*
 * srf_oracle.c -- CPU restatement of the SRF-NEF Siddon ray marchers.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle and the CPU
 * baseline.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg load it; the product path (srf_amd/, libsrf_hip.so) never does.
 *
 * It restates, in plain C with IEEE single precision, no FP contraction and the
 * reference's own double-precision TOF expressions, the four device functions:
 *
 *   non-TOF projector      srfnef/cc/siddon/siddon.cu.cc:12-338
 *   non-TOF backprojector  srfnef/cc/siddon/siddon.cu.cc:342-656
 *   TOF projector          srfnef/cc/siddon_tof/siddon_tof.cu.cc:14-367
 *   TOF backprojector      srfnef/cc/siddon_tof/siddon_tof.cu.cc:371-707
 *
 * The reference writes each of those as two hand-copied branch bodies (x- and
 * y-dominant).  Here one generic body is parameterised by the primary axis;
 * the arithmetic of each branch is kept operation for operation (see the
 * comments in og_march), so the voxel indices and the float weights are those
 * of an un-contracted IEEE build of the reference.
 *
 * "variant" selects the predicate/weight set:
 *   OG_FP  projector:   dominant test >= (non-TOF, siddon.cu.cc:46) or > (TOF,
 *                       siddon_tof.cu.cc:48); non-TOF cut sqrt(xd^2+yd^2)<|zd|
 *                       (siddon.cu.cc:38); segment cuts (siddon.cu.cc:56-57,
 *                       203-204; siddon_tof.cu.cc:67-74,228-235); weight /L/L.
 *   OG_BP  backprojector: dominant test > (siddon.cu.cc:372,
 *                       siddon_tof.cu.cc:405); no cuts besides the 10-voxel
 *                       one; weight without /L/L.
 * "action" selects what is done with each term: gather (projection), scatter
 * (backprojection), digest (term hash) or list (term dump).  Projector
 * predicates + scatter is the exact adjoint of the projector.
 *
 * Image layout is the reference kernels' own: flat index ix + iy*nx + iz*nx*ny
 * (siddon.cu.cc:95), i.e. x fastest.  LORs are AoS rows of `stride` floats
 * [x1,y1,z1,x2,y2,z2,(tof)] in mm (srfnef/data/lors.py:12-22).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

enum { OG_FP = 0, OG_BP = 1 };
enum { ACT_GATHER = 0, ACT_SCATTER = 1, ACT_DIGEST = 2, ACT_LIST = 3 };

/* siddon_tof.cu.cc:10  const float PI = 3.1415926; */
static const float OG_PI = 3.1415926f;

typedef struct {
    int nx, ny, nz;
    float cx, cy, cz, sx, sy, sz;
} og_grid;

typedef struct {
    int tof;            /* 1: siddon_tof.cu.cc family, 0: siddon.cu.cc family */
    int variant;        /* OG_FP or OG_BP */
    float tof_bin, tof_sigma2;
} og_kind;

/* Per-event result sink. */
typedef struct {
    const float *img_in;    /* gather source */
    float *img_out;         /* scatter target (f32) */
    double *img_out64;      /* scatter target (f64), used when non-null */
    float value;            /* scatter value (the listmode entry) */
    float acc;              /* gather accumulator, float as the reference */
    double acc64;           /* gather accumulator, double */
    uint64_t hash;          /* digest */
    int64_t count;
    /* term list */
    int64_t *l_vox;
    int32_t *l_slot;
    float *l_w;
    float *l_wtof;
    int64_t l_cap;
} og_sink;

/* (int)floor(x) with the GPU conversion semantics (NaN -> 0, saturating), so
 * huge secondary coordinates of steep rays behave identically on both sides. */
static inline int og_floor_i(float x)
{
    float f = floorf(x);
    if (f != f) return 0;
    if (f >= 2147483648.0f) return 2147483647;
    if (f <= -2147483648.0f) return -2147483647 - 1;
    return (int)f;
}

static inline uint64_t og_mix64(uint64_t z)
{
    /* splitmix64 finaliser */
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/*
 * One term.  `rw` is the reference's per-branch fraction (1, r, 1-r, r1-r2),
 * `sq` = sqrt(1 + k*k + kz*kz), `w_tof` is 1 for non-TOF.
 * Weight and accumulation order follow siddon.cu.cc:94-95 / 418-419 and
 * siddon_tof.cu.cc:110-111 / 460-461:
 *   FP: weight = rw * sq * dx_ / L / L;  vproj += image[v] * weight [* w_tof]
 *   BP: weight = rw * sq * dx_;          image[v] += vproj * weight [* w_tof]
 */
static inline __attribute__((always_inline)) void
og_term(og_sink *s, int act, const og_kind *k, int64_t v, int slot, float rw,
        float sq, float dx_, float L, float w_tof)
{
    float weight = rw * sq * dx_;
    if (k->variant == OG_FP) weight = weight / L / L;
    switch (act) {
    case ACT_GATHER: {
        float t = s->img_in[v] * weight;
        if (k->tof) t = t * w_tof;
        s->acc += t;
        s->acc64 += (double)t;
        break;
    }
    case ACT_SCATTER: {
        float t = s->value * weight;
        if (k->tof) t = t * w_tof;
        if (s->img_out64) s->img_out64[v] += (double)t;
        else s->img_out[v] += t;
        break;
    }
    case ACT_DIGEST:
        s->hash += og_mix64(((uint64_t)v << 4) | (uint64_t)slot);
        s->count += 1;
        break;
    case ACT_LIST:
        if (s->count < s->l_cap) {
            s->l_vox[s->count] = v;
            s->l_slot[s->count] = slot;
            s->l_w[s->count] = weight;
            s->l_wtof[s->count] = k->tof ? w_tof : 1.0f;
        }
        s->count += 1;
        break;
    }
}

/*
 * The marcher.  Returns after the early-out tests exactly where the reference
 * returns (the output stays at its memset value 0).
 */
static inline __attribute__((always_inline)) void
og_march(const float *lor, float tof_value, const og_grid *g, const og_kind *k,
         int act, og_sink *s)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    /* siddon.cu.cc:19-28 -- everything in units of dx_ */
    const float dx_ = g->sx / nx;
    const float dx = 1.0f;
    const float dy = g->sy / ny / dx_;
    const float dz = g->sz / nz / dx_;
    const float x1 = (lor[0] - g->cx) / dx_;
    const float x2 = (lor[3] - g->cx) / dx_;
    const float y1 = (lor[1] - g->cy) / dx_;
    const float y2 = (lor[4] - g->cy) / dx_;
    const float z1 = (lor[2] - g->cz) / dx_;
    const float z2 = (lor[5] - g->cz) / dx_;
    const float xd = x2 - x1;
    const float yd = y2 - y1;
    const float zd = z2 - z1;

    /* Non-finite endpoints: the reference would march a NaN ray (its float->int
     * conversion maps NaN to 0).  Contract here and in the HIP kernels: the
     * projection is NaN and the backprojector skips the event. */
    if (!(isfinite(x1) && isfinite(y1) && isfinite(z1) && isfinite(x2) && isfinite(y2) && isfinite(z2))) {
        if (act == ACT_GATHER) { s->acc = NAN; s->acc64 = NAN; }
        return;
    }

    /* siddon.cu.cc:34 / 364, siddon_tof.cu.cc:37 / 395 */
    if (sqrtf(xd * xd + yd * yd) < 10.0f) return;
    /* siddon.cu.cc:38 (non-TOF projector only) */
    if (!k->tof && k->variant == OG_FP) {
        float azd = zd < 0 ? -zd : zd;
        if (sqrtf(xd * xd + yd * yd) < azd) return;
    }
    const float nx2 = nx / 2.0f;
    const float ny2 = ny / 2.0f;
    const float nz2 = nz / 2.0f;
    const float L = sqrtf(xd * xd + yd * yd + zd * zd);

    /* siddon_tof.cu.cc:44-47 (TOF constants; double where the reference is) */
    float ratio1 = 0.0f, tof_bin_ = 0.0f, tof_sigma2_ = 0.0f;
    if (k->tof) {
        ratio1 = (float)((1.0 - (double)(tof_value / dx_ / L)) / 2);
        tof_bin_ = k->tof_bin / dx_;
        tof_sigma2_ = k->tof_sigma2 / dx_ / dx_;
    }

    /* dominant axis: >= for the non-TOF projector, > everywhere else */
    const float axd = xd < 0 ? -xd : xd;
    const float ayd = yd < 0 ? -yd : yd;
    int xdom = (!k->tof && k->variant == OG_FP) ? (axd >= ayd) : (axd > ayd);

    /* Generic primary/secondary description of the two reference branches.
     * x-branch: siddon.cu.cc:46-191; y-branch: siddon.cu.cc:192-335. */
    float p1, p2, sec1, ds, ns2, k_s, kz, pd_tof;
    int np, ns;
    int64_t stride_p, stride_s;
    if (xdom) {
        p1 = x1; p2 = x2; sec1 = y1; ds = dy; ns2 = ny2;
        np = nx; ns = ny; stride_p = 1; stride_s = nx;
        k_s = yd / xd; kz = zd / xd;
        pd_tof = dx;                       /* xc = (ix - nx/2 + 0.5) * dx */
    } else {
        p1 = y1; p2 = y2; sec1 = x1; ds = 1.0f; ns2 = nx2;   /* x/1.0f == x exactly */
        np = ny; ns = nx; stride_p = nx; stride_s = 1;
        k_s = xd / yd; kz = zd / yd;
        pd_tof = dy;                       /* yc = (iy - ny/2 + 0.5) * dy */
    }
    const float np2 = xdom ? nx2 : ny2;
    const int64_t stride_z = (int64_t)nx * ny;
    const float sq = sqrtf(1 + k_s * k_s + kz * kz);

    for (int ip = 0; ip < np; ++ip) {
        float w_tof = 1.0f;
        if (k->tof) {
            /* siddon_tof.cu.cc:55-59 / 215-219 */
            float pc = (float)(((double)(ip - np / 2) + 0.5) * (double)pd_tof);
            float d2_tof = ((pc - p1) / (p2 - p1) - ratio1) * L;
            float tse = tof_sigma2_ + (tof_bin_ * tof_bin_) / 12;
            float t2 = d2_tof * d2_tof / tse;
            w_tof = (float)((double)tof_bin_ * exp(-0.5 * (double)t2) /
                            sqrt(2.0 * (double)OG_PI * (double)tse));
        }
        const float pp1 = (float)ip - np2;
        const float pp2 = pp1 + 1.0f;
        if (k->variant == OG_FP) {
            if (!k->tof) {
                /* siddon.cu.cc:56-57 (x) and 203-204 (y): same form */
                if (p1 < p2 && pp2 > p2) continue;
                if (p1 >= p2 && pp1 < p2) continue;
            } else if (xdom) {
                /* siddon_tof.cu.cc:67-74 */
                if (p2 > p1 && pp1 <= p1) continue;
                if (p2 < p1 && pp2 >= p1) continue;
            } else {
                /* siddon_tof.cu.cc:228-235 */
                if (p2 > p1 && pp1 < p1) continue;
                if (p2 < p1 && pp2 > p1) continue;
            }
        }
        float s_lo, s_hi, z_lo, z_hi;
        if (k_s >= 0.0f) {
            s_lo = (sec1 + k_s * (pp1 - p1)) / ds + ns2;
            s_hi = (sec1 + k_s * (pp2 - p1)) / ds + ns2;
        } else {
            s_lo = (sec1 + k_s * (pp2 - p1)) / ds + ns2;
            s_hi = (sec1 + k_s * (pp1 - p1)) / ds + ns2;
        }
        const int cs1 = og_floor_i(s_lo);
        const int cs2 = og_floor_i(s_hi);
        if (kz >= 0.0f) {
            z_lo = (z1 + kz * (pp1 - p1)) / dz + nz2;
            z_hi = (z1 + kz * (pp2 - p1)) / dz + nz2;
        } else {
            z_lo = (z1 + kz * (pp2 - p1)) / dz + nz2;
            z_hi = (z1 + kz * (pp1 - p1)) / dz + nz2;
        }
        const int cz1 = og_floor_i(z_lo);
        const int cz2 = og_floor_i(z_hi);

#define OG_VOX(cs, czz) ((int64_t)ip * stride_p + (int64_t)(cs) * stride_s + (int64_t)(czz) * stride_z)
#define OG_T(cs, czz, slot, rw) og_term(s, act, k, OG_VOX(cs, czz), slot, rw, sq, dx_, L, w_tof)
        /* branch tree of siddon.cu.cc:86-189 (x) / 231-333 (y) */
        if (cs1 == cs2) {
            if (0 <= cs1 && cs1 < ns) {
                if (cz1 == cz2) {
                    if (0 <= cz1 && cz1 < nz) OG_T(cs1, cz1, 0, 1.0f);
                } else if (-1 <= cz1 && cz1 < nz) {
                    float rz = (cz2 - z_lo) / (z_hi - z_lo);
                    if (cz1 >= 0) OG_T(cs1, cz1, 1, rz);
                    if (cz2 < nz) OG_T(cs1, cz2, 2, 1 - rz);
                }
            }
        } else if (-1 <= cs1 && cs1 < ns) {
            if (cz1 == cz2) {
                if (0 <= cz1 && cz1 < nz) {
                    float rs = (cs2 - s_lo) / (s_hi - s_lo);
                    if (cs1 >= 0) OG_T(cs1, cz1, 3, rs);
                    if (cs2 < ns) OG_T(cs2, cz1, 4, 1 - rs);
                }
            } else if (-1 <= cz1 && cz1 < nz) {
                float rs = (cs2 - s_lo) / (s_hi - s_lo);
                float rz = (cz2 - z_lo) / (z_hi - z_lo);
                if (rs > rz) {
                    if (cs1 >= 0 && cz1 >= 0) OG_T(cs1, cz1, 5, rz);
                    if (cs1 >= 0 && cz2 < nz) OG_T(cs1, cz2, 6, rs - rz);
                    if (cs2 < ns && cz2 < nz) OG_T(cs2, cz2, 7, 1 - rs);
                } else {
                    if (cs1 >= 0 && cz1 >= 0) OG_T(cs1, cz1, 8, rs);
                    if (cs2 < ns && cz1 >= 0) OG_T(cs2, cz1, 9, rz - rs);
                    if (cs2 < ns && cz2 < nz) OG_T(cs2, cz2, 10, 1 - rz);
                }
            }
        }
#undef OG_T
#undef OG_VOX
    }
}

static void og_fill(og_grid *g, og_kind *k, const int32_t grid[3], const float center[3],
                    const float size[3], int tof, float tof_bin, float tof_sigma2, int variant)
{
    g->nx = grid[0]; g->ny = grid[1]; g->nz = grid[2];
    g->cx = center[0]; g->cy = center[1]; g->cz = center[2];
    g->sx = size[0]; g->sy = size[1]; g->sz = size[2];
    k->tof = tof; k->variant = variant; k->tof_bin = tof_bin; k->tof_sigma2 = tof_sigma2;
}

static inline float og_tofv(const float *tofv, int64_t tof_stride, int64_t e)
{
    return tofv ? tofv[e * tof_stride] : 0.0f;
}

static int og_threads(int nthreads)
{
#ifdef _OPENMP
    return nthreads > 0 ? nthreads : omp_get_max_threads();
#else
    (void)nthreads;
    return 1;
#endif
}

/* Projection: out32[e] (float accumulation in slab order, as the reference),
 * out64[e] (same terms accumulated in double).  Either may be NULL. */
int og_project(const float *lors, int64_t n, int64_t stride, const float *tofv, int64_t tof_stride,
               const float *image, const int32_t grid[3], const float center[3], const float size[3],
               int tof, float tof_bin, float tof_sigma2, int variant,
               float *out32, double *out64, int nthreads)
{
    og_grid g; og_kind k;
    og_fill(&g, &k, grid, center, size, tof, tof_bin, tof_sigma2, variant);
    int nt = og_threads(nthreads);
    (void)nt;
#pragma omp parallel for schedule(static) num_threads(nt)
    for (int64_t e = 0; e < n; ++e) {
        og_sink s;
        memset(&s, 0, sizeof(s));
        s.img_in = image;
        og_march(lors + e * stride, og_tofv(tofv, tof_stride, e), &g, &k, ACT_GATHER, &s);
        if (out32) out32[e] = s.acc;
        if (out64) out64[e] = s.acc64;
    }
    return 0;
}

/* Backprojection into image32 (float) or image64 (double; preferred as the
 * parity target, since the reference's atomicAdd order is unspecified).  The
 * target is accumulated into (the caller zeroes it, as the reference's
 * cudaMemset at siddon.cc:128 does).  With several threads each one owns a
 * private image over a static chunk of events and the images are summed in
 * thread order, so the result is deterministic for a given thread count. */
int og_backproject(const float *lors, int64_t n, int64_t stride, const float *tofv, int64_t tof_stride,
                   const float *values, const int32_t grid[3], const float center[3], const float size[3],
                   int tof, float tof_bin, float tof_sigma2, int variant,
                   float *image32, double *image64, int nthreads)
{
    og_grid g; og_kind k;
    og_fill(&g, &k, grid, center, size, tof, tof_bin, tof_sigma2, variant);
    const int64_t nvox = (int64_t)g.nx * g.ny * g.nz;
    int nt = og_threads(nthreads);
    if (nt <= 1) {
        og_sink s;
        memset(&s, 0, sizeof(s));
        s.img_out = image32; s.img_out64 = image64;
        for (int64_t e = 0; e < n; ++e) {
            s.value = values[e];
            og_march(lors + e * stride, og_tofv(tofv, tof_stride, e), &g, &k, ACT_SCATTER, &s);
        }
        return 0;
    }
    size_t esz = image64 ? sizeof(double) : sizeof(float);
    char *priv = (char *)calloc((size_t)nt * (size_t)nvox, esz);
    if (!priv) return -1;
#pragma omp parallel num_threads(nt)
    {
#ifdef _OPENMP
        int t = omp_get_thread_num();
        int ntt = omp_get_num_threads();
#else
        int t = 0, ntt = 1;
#endif
        int64_t chunk = (n + ntt - 1) / ntt;
        int64_t e0 = (int64_t)t * chunk, e1 = e0 + chunk < n ? e0 + chunk : n;
        og_sink s;
        memset(&s, 0, sizeof(s));
        if (image64) s.img_out64 = (double *)(priv + (size_t)t * nvox * esz);
        else s.img_out = (float *)(priv + (size_t)t * nvox * esz);
        for (int64_t e = e0; e < e1; ++e) {
            s.value = values[e];
            og_march(lors + e * stride, og_tofv(tofv, tof_stride, e), &g, &k, ACT_SCATTER, &s);
        }
#pragma omp barrier
#pragma omp for schedule(static)
        for (int64_t v = 0; v < nvox; ++v) {
            if (image64) {
                double a = image64[v];
                for (int q = 0; q < ntt; ++q) a += ((double *)priv)[(size_t)q * nvox + v];
                image64[v] = a;
            } else {
                float a = image32[v];
                for (int q = 0; q < ntt; ++q) a += ((float *)priv)[(size_t)q * nvox + v];
                image32[v] = a;
            }
        }
    }
    free(priv);
    return 0;
}

/* Per-event term digest: hash[e] = sum over terms of mix64(voxel<<4 | slot)
 * (wrapping), count[e] = number of terms.  Commutative, so any kernel that
 * visits the same (voxel, branch) terms in any order gives the same digest. */
int og_digest(const float *lors, int64_t n, int64_t stride, const float *tofv, int64_t tof_stride,
              const int32_t grid[3], const float center[3], const float size[3],
              int tof, float tof_bin, float tof_sigma2, int variant,
              uint64_t *hash, int32_t *count, int nthreads)
{
    og_grid g; og_kind k;
    og_fill(&g, &k, grid, center, size, tof, tof_bin, tof_sigma2, variant);
    int nt = og_threads(nthreads);
    (void)nt;
#pragma omp parallel for schedule(static) num_threads(nt)
    for (int64_t e = 0; e < n; ++e) {
        og_sink s;
        memset(&s, 0, sizeof(s));
        og_march(lors + e * stride, og_tofv(tofv, tof_stride, e), &g, &k, ACT_DIGEST, &s);
        hash[e] = s.hash;
        count[e] = (int32_t)s.count;
    }
    return 0;
}

/* Term list of one event (for tests): returns the number of terms (which may
 * exceed cap; only the first cap are written). */
int64_t og_terms(const float *lor, float tof_value, const int32_t grid[3], const float center[3],
                 const float size[3], int tof, float tof_bin, float tof_sigma2, int variant,
                 int64_t *vox, int32_t *slot, float *w, float *wtof, int64_t cap)
{
    og_grid g; og_kind k;
    og_fill(&g, &k, grid, center, size, tof, tof_bin, tof_sigma2, variant);
    og_sink s;
    memset(&s, 0, sizeof(s));
    s.l_vox = vox; s.l_slot = slot; s.l_w = w; s.l_wtof = wtof; s.l_cap = cap;
    og_march(lor, tof_value, &g, &k, ACT_LIST, &s);
    return s.count;
}

int og_num_threads(void)
{
    return og_threads(0);
}
