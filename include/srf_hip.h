/*
 * srf_hip.h -- C ABI of libsrf_hip.so, the MI355X (gfx950) list-mode PET hot path.
 *
 * Conventions shared by every entry point
 *  - All pointers are device pointers (HBM) unless the comment says otherwise.
 *    The caller owns every buffer; the library allocates device memory only in
 *    srf_plan_create[_ex], srf_plan_rebuild and srf_batch_create (freed by the
 *    matching *_destroy) and transient scratch in srf_sinogram_lors and
 *    srf_event_order.
 *  - LORs are AoS rows of `lor_stride` floats [x1,y1,z1,x2,y2,z2,(tof,...)] in
 *    mm, exactly Lors.data of srfnef/data/lors.py:12-22 (stride 6 or 7).
 *  - TOF values are read from `tof_values[e * tof_stride]`; pass
 *    `lors + 6` with `tof_stride = 7` to read column 6 in place
 *    (Lors.tof_values, srfnef/data/lors.py:20-22).
 *  - Images use the reference kernels' layout: flat index ix + iy*nx + iz*nx*ny
 *    (srfnef/cc/siddon/siddon.cu.cc:95), i.e. the tf.transpose(image.data) that
 *    srfnef/functions/project.py:27 hands to the op.
 *  - `stream` is a hipStream_t (NULL = the legacy default stream).  Every call is
 *    asynchronous on that stream and host-validated before any launch.
 *  - Return value: SRF_OK (0) or a negative SRF_ERR_* code; srf_strerror()
 *    names it.
 */
#ifndef SRF_HIP_H
#define SRF_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Image grid: number of voxels, centre and full size in mm.
 * Replaces the device tensors grid/center/size of the TF ops
 * (srfnef/cc/siddon/siddon.cc:11-21) -- passed by value, so no D2H copies
 * (srfnef/cc/siddon/siddon.cu.cc:715-717). */
typedef struct {
    int32_t nx, ny, nz;
    float cx, cy, cz;
    float sx, sy, sz;
} srf_grid_t;

/* TOF kernel parameters: attributes tof_bin / tof_sigma2 of ProjectionTof
 * (srfnef/cc/siddon_tof/siddon_tof.cc:18-19), TofConfig defaults 3.66 / 1000
 * (srfnef/geometry/tof_config.py:16-17).  NULL selects the non-TOF kernels. */
typedef struct {
    float tof_bin;
    float tof_sigma2;
} srf_tof_t;

/* Backprojection weight set. */
typedef enum {
    /* Bit-compatible with Backprojection / BackprojectionTof: weight without
     * 1/L^2, dominant test '>', no segment cuts (siddon.cu.cc:342-656,
     * siddon_tof.cu.cc:371-707). */
    SRF_BP_REFERENCE = 0,
    /* Exact transpose of srf_project: the projector's predicates, cuts and 1/L^2
     * weight, scattered instead of gathered.  Backward of the autograd projector. */
    SRF_BP_ADJOINT = 1
} srf_bp_mode_t;

/* Ratio kinds of the EM drivers (reference file:line per kind). */
typedef enum {
    SRF_RATIO_PLAIN = 0,      /* q = y / (p + eps): mlem.py:41-42 (eps = mean(y)*1e-8), osem_full.py:74-76 (1e-8) */
    SRF_RATIO_DIV_NO_NAN = 1  /* q = (p+eps)==0 ? 0 : y/(p+eps): mlem_full.py:68 (eps 0), mlem_full_tof.py:64 (1e-5) */
} srf_ratio_kind_t;

#define SRF_OK 0
#define SRF_ERR_INVALID_ARG (-1)   /* null pointer, negative size, bad stride or grid */
#define SRF_ERR_UNSUPPORTED (-2)   /* e.g. image larger than the 2^31 voxel index space */
#define SRF_ERR_LAUNCH (-3)        /* a HIP launch or API call failed */
#define SRF_ERR_WORKSPACE (-4)     /* workspace too small */

const char *srf_strerror(int code);
int srf_version(void);

/* Forward projection (Projection / ProjectionTof, siddon.cc:11-21,
 * siddon_tof.cc:11-24; host fn projection(), siddon.cu.cc:706-728 and
 * siddon_tof.cu.cc:761-785).  out[e] = sum_v image[v] * w_ev for every event;
 * events the reference skips get 0.  `out` is fully overwritten. */
int srf_project(const float *lors, int64_t n, int64_t lor_stride,
                const float *tof_values, int64_t tof_stride,
                const float *image, srf_grid_t grid, const srf_tof_t *tof,
                float *out, void *stream);

/* Backprojection (Backprojection / BackprojectionTof, siddon.cc:23-35,
 * siddon_tof.cc:26-41; host fn backprojection(), siddon.cu.cc:731-753 and
 * siddon_tof.cu.cc:788-812).  image[v] (+)= sum_e values[e] * w_ev.
 * accumulate = 0 zeroes `image` first (the op's cudaMemset, siddon.cc:128);
 * 1 adds into it.  `values` may be NULL, meaning all ones (the sensitivity
 * image, emap_generator.py:169-201). */
int srf_backproject(const float *lors, int64_t n, int64_t lor_stride,
                    const float *tof_values, int64_t tof_stride,
                    const float *values, srf_grid_t grid, const srf_tof_t *tof,
                    srf_bp_mode_t mode, int accumulate,
                    float *image, void *stream);

/* Backprojection of the EM ratio q[e] = ratio(y[e], p[e]) without
 * materialising q: the ratio of mlem.py:41-42 / mlem_full_tof.py:64 /
 * osem_full.py:74-76 fused into the backprojector's event load. */
int srf_backproject_ratio(const float *lors, int64_t n, int64_t lor_stride,
                          const float *tof_values, int64_t tof_stride,
                          const float *y, const float *p, srf_ratio_kind_t kind, float eps,
                          srf_grid_t grid, const srf_tof_t *tof,
                          srf_bp_mode_t mode, int accumulate,
                          float *image, void *stream);

/* q[e] = ratio(y[e], p[e]) (see srf_ratio_kind_t). */
int srf_ratio(const float *y, const float *p, float *q, int64_t n,
              srf_ratio_kind_t kind, float eps, void *stream);

/* EM multiplicative update x <- x * bp / e, e = emap (zero_guard == 0,
 * osem_full.py:77) or e = (emap == 0 ? zero_guard : emap) (mlem.py:33-35,43,
 * zero_guard = 1e8).  16 B/voxel, one pass. */
int srf_em_update(float *x, const float *bp, const float *emap, int64_t nvox,
                  float zero_guard, void *stream);

/* Per-event digest of the (voxel, branch) terms the projector
 * (mode = SRF_BP_ADJOINT: projector predicates) or backprojector
 * (SRF_BP_REFERENCE) visits: hash[e] = sum_terms mix64(voxel << 4 | branch),
 * count[e] = number of terms.  Runs the same device marcher as srf_project /
 * srf_backproject; used to prove voxel indexing bit-exact. */
int srf_lor_digest(const float *lors, int64_t n, int64_t lor_stride,
                   const float *tof_values, int64_t tof_stride,
                   srf_grid_t grid, const srf_tof_t *tof, srf_bp_mode_t mode,
                   uint64_t *hash, int32_t *count, void *stream);

/* Separable Gaussian smoothing of a C-order (n0, n1, n2) image (sigmas in
 * voxels, <= 64; weights renormalised where the kernel leaves the image), then
 * out *= mask (mask may be NULL).  x, tmp and out are distinct buffers of
 * n0*n1*n2 floats.  The scatter estimate's emission input in the OSEM pipeline
 * (no counterpart in the reference). */
int srf_smooth(const float *x, int32_t n0, int32_t n1, int32_t n2, float sigma0, float sigma1, float sigma2,
               const float *mask, float *tmp, float *out, void *stream);
/* srf_smooth without the tmp argument, for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  The intermediate image
 * lives in a static device array of 2^23 floats (32 MiB) that the library
 * holds on every device it runs on until the process ends; n0*n1*n2 above 2^23
 * (a 256 x 256 x 128 image) returns SRF_ERR_WORKSPACE.  Nothing is allocated
 * and nothing needs clearing between calls.  All calls on one device share
 * that array: calls that overlap on one device (several streams or host
 * threads) race and must be serialised by the caller.  Same kernels and the
 * same results as srf_smooth, bit for bit. */
int srf_smooth_noalloc(const float *x, int32_t n0, int32_t n1, int32_t n2, float sigma0, float sigma1,
                       float sigma2, const float *mask, float *out, void *stream);

/* In-plane PSF blur of x (nx ny, nz; slices fastest) with the sparse K_xy of
 * srfnef's Deconvolute (corrections/psf/deconvolute.py:56-62, built by
 * xy_main :163-213) held as per-source-pixel patch weights
 * weights[p][(a + half) (2 half + 1) + b + half] on target pixel p + (a, b).
 * adjoint 0: out = K_xy x, 1: out = K_xy^T x; d != NULL: out = d / (. + eps).
 * out must not alias x or d. */
int srf_psf_xy(const float *x, int32_t nx, int32_t ny, int32_t nz, const float *weights, int32_t half,
               int32_t adjoint, const float *d, float eps, float *out, void *stream);
/* Layout conversions between the Python (nx,ny,nz) C-order image (z fastest)
 * and the kernel layout (x fastest): dst = transpose(src). */
int srf_image_to_kernel_layout(const float *src_xyz, float *dst_kernel, int32_t nx, int32_t ny,
                               int32_t nz, void *stream);
int srf_image_from_kernel_layout(const float *src_kernel, float *dst_xyz, int32_t nx, int32_t ny,
                                 int32_t nz, void *stream);

/* ---------------------------------------------------------------- plans --
 * A plan bins a fixed event list (one OSEM subset, a sensitivity LOR set) into
 * per-tile segments once, so that every later projection and backprojection of
 * the same events runs the LDS-tiled kernels.  Same semantics and the same
 * reference entry points as srf_project / srf_backproject above
 * (siddon.cu.cc:706-753, siddon_tof.cu.cc:761-812); results agree with them to
 * float round-off (summation order differs).
 *
 * srf_plan_create, srf_plan_rebuild and srf_plan_destroy are the plan entry
 * points that allocate or free device memory (about 56 B per event-tile
 * segment: the 48-byte record, its event id and its partial sum; plus per-item
 * tile buffers); srf_plan_info reports the total.  The LOR array is read
 * during create / rebuild only, which synchronise `stream`.  An event list may
 * bin into at most 2^31 - 1 segments, and a plan's grid may have at most 32767
 * cells along x and along y (SRF_ERR_UNSUPPORTED beyond: split the list into
 * several plans, as OSEM subsets are).  The memory comes from a stream-ordered pool per
 * device that keeps up to 16 GiB of freed plan memory for the next plans
 * (hipMalloc where a device has no pools; srf_plan_pool_trim releases it);
 * srf_plan_destroy first waits for the device, as hipFree does. */
typedef struct srf_plan srf_plan;

int srf_plan_create(const float *lors, int64_t n, int64_t lor_stride,
                    const float *tof_values, int64_t tof_stride,
                    srf_grid_t grid, const srf_tof_t *tof, srf_plan **plan, void *stream);
/* Plan options (srf_plan_create_ex; NULL = the srf_plan_create defaults). */
typedef struct {
    /* Largest work item in segments, 256..8192 (0 = 8192).  Smaller items
     * spread a crowded tile over more workgroups (more tile partials to sum);
     * the batch autotuner chooses it. */
    int32_t item_records;
    /* TOF window: 0 keeps every term, as the reference (siddon_tof.cu.cc
     * weights every slab, :55-59).  w > 0 bins only the slabs whose TOF
     * Gaussian factor is at least 2^-w of the largest factor among the
     * event's slabs in the image (plus one slab each side).  Where the TOF
     * centre lies in the image the dropped weight is erfc(sqrt(w ln 2)) of the
     * event's: w = 22 drops < 4e-8 (below one float32 ulp of the total) and
     * about half of the terms at 60 mm FWHM in a 512 mm field of view; w = 26,
     * < 2e-9.  Where it lies outside, the window keeps the in-image tail
     * around the largest term, so the EM ratio y / sum(terms) still spreads
     * the event as the reference does. */
    float tof_window;
} srf_plan_opts_t;

int srf_plan_create_ex(const float *lors, int64_t n, int64_t lor_stride,
                       const float *tof_values, int64_t tof_stride,
                       srf_grid_t grid, const srf_tof_t *tof, const srf_plan_opts_t *opts,
                       srf_plan **plan, void *stream);
/* Bins a new event list into an existing plan, reusing its device buffers:
 * they are grown when the new list needs more and otherwise kept, with the
 * binning scratch, so a loop over chunks of one size (the sensitivity image
 * over all crystal pairs) allocates only on its first pass.  Arguments as
 * srf_plan_create_ex.  On failure the plan keeps its previous grid and holds
 * an empty list: projections write nothing, backprojections write zeros (add
 * nothing when accumulating), srf_plan_backproject_update sets x to
 * x * 0 / emap, and srf_batch_create refuses it until a rebuild succeeds.
 * No reference counterpart (srfnef re-runs its TF op per call). */
int srf_plan_rebuild(srf_plan *plan, const float *lors, int64_t n, int64_t lor_stride,
                     const float *tof_values, int64_t tof_stride, srf_grid_t grid, const srf_tof_t *tof,
                     const srf_plan_opts_t *opts, void *stream);
int srf_plan_destroy(srf_plan *plan);
/* Returns the current device's pooled plan memory above keep_bytes to the
 * system (hipMemPoolTrimTo after draining the device), so that other HIP users
 * of the process (torch's caching allocator) can have it; 0 releases all
 * memory no live plan holds.  No reference counterpart (the plan pool is this
 * library's). */
int srf_plan_pool_trim(size_t keep_bytes);
/* Geometric event order within OSEM subsets.  pidx holds n event indices
 * into the LOR array (n_lors rows of lor_stride floats) in n_sub contiguous
 * subsets of n / n_sub entries (a remainder forms one more group, last).
 * out = pidx with each subset's entries stably sorted by the geometry of
 * their LORs: transaxial angle of the direction folded into [0, pi) (128
 * bins), axial midpoint (16 mm bins), signed transaxial distance from the
 * axis (1 mm bins).  Subset membership is unchanged, and with it every EM
 * update; events that cross the same image tile get nearby positions, so a
 * plan built from the ordered subset gathers and scatters its per-event
 * values over few cache lines.  The reference keeps its subsets in the order
 * of the shuffled permutation (srfnef/functions/osem_full.py:63-73): this
 * reorders within a subset only.  n < 2^31; out must not alias pidx.  Scratch
 * (about 32 B per entry) comes from the plan pool; the call synchronises
 * `stream`.  An index outside [0, n_lors) returns SRF_ERR_INVALID_ARG (out
 * is then unspecified). */
int srf_event_order(const float *lors, int64_t n_lors, int64_t lor_stride, const int64_t *pidx, int64_t n,
                    int64_t n_sub, int64_t *out, void *stream);
int srf_plan_info(const srf_plan *plan, int64_t *n_events, int64_t *n_segments,
                  int64_t *n_items, size_t *device_bytes);
/* out[e] = projection of event e (srf_project semantics). */
int srf_plan_project(srf_plan *plan, const float *image, float *out, void *stream);
/* q[e] = ratio(y[e], p[e]) with p = the projection (srf_plan_project then
 * srf_ratio, fused: the ratio is applied in the per-event reduction).  The EM
 * sub-iteration's forward half (mlem.py:40-42, mlem_full_tof.py:63-64,
 * osem_full.py:74-75). */
int srf_plan_project_ratio(srf_plan *plan, const float *image, const float *y, srf_ratio_kind_t kind,
                           float eps, float *q, void *stream);
/* image (+)= backprojection (srf_backproject semantics; values NULL = ones). */
int srf_plan_backproject(srf_plan *plan, const float *values, srf_bp_mode_t mode,
                         int accumulate, float *image, void *stream);
/* One EM update from a plan: x = x * srf_plan_backproject(values, reference
 * mode) / emap (emap 0 -> zero_guard when it is not 0), the tile sums fused
 * with srf_em_update's operations (bit-identical to the two calls). */
int srf_plan_backproject_update(srf_plan *plan, const float *values, float *x, const float *emap,
                                float zero_guard, void *stream);
/* image (+)= backprojection of ratio(y, p) (srf_backproject_ratio semantics). */
int srf_plan_backproject_ratio(srf_plan *plan, const float *y, const float *p,
                               srf_ratio_kind_t kind, float eps, srf_bp_mode_t mode,
                               int accumulate, float *image, void *stream);
/* srf_lor_digest semantics, computed by the tiled kernels' own segment walk:
 * proves that every term lands in exactly one tile segment. */
int srf_plan_digest(srf_plan *plan, srf_bp_mode_t mode, uint64_t *hash, int32_t *count,
                    void *stream);
/* Measurement hook, no counterpart in the reference: the next projection or
 * backprojection of this plan records the two hipEvent_t handles on its stream
 * immediately before and after its tiled kernel (not the per-event reduction
 * or tile sums around it), so a caller can time that kernel alone.  One-shot:
 * the events are dropped after that launch. */
int srf_plan_time_next(srf_plan *plan, void *before_event, void *after_event);

/* -------------------------------------------------------------- batches --
 * Many reconstructions (scans of different sizes and event counts) as one
 * launch per direction.  No counterpart in the reference, which reconstructs
 * one scan per OsemFull call (srfnef/functions/osem_full.py:49-87); per scan
 * the results are bit-identical to the single-plan calls
 * (srf_plan_project_ratio + srf_plan_backproject(SRF_BP_REFERENCE) +
 * srf_em_update).
 *
 * A batch binds one plan per scan (typically one OSEM subset each) and the
 * scan's EM buffers.  Plans may differ in grid, TOF and voxel aspect: the work
 * list is grouped by kernel variant, one launch per non-empty group (one
 * launch per direction when all scans share a variant).  Every call's kernel
 * arguments are fixed at
 * creation, so a sequence of srf_batch_em_step calls can be captured in a HIP
 * graph and replayed.  srf_batch_create allocates the work list (freed by
 * srf_batch_destroy) and synchronises `stream`; plans and buffers must outlive
 * the batch. */
typedef struct srf_batch srf_batch;

typedef struct {
    float *image;         /* estimate x (kernel layout), updated by the EM step */
    const float *emap;    /* sensitivity image (kernel layout) */
    const float *y;       /* event values, one per plan event */
    float *q;             /* workspace, one float per plan event */
} srf_batch_scan_t;

typedef struct {
    int32_t persistent;          /* 0: one workgroup per work item; 1: a persistent
                                  * grid that takes items from an atomic queue */
    int32_t workgroups_per_cu;   /* persistent grid = this x compute units (1..16) */
    int32_t order;               /* 0: scan order; 1: work items by decreasing estimated work (records x slab-count class) */
} srf_batch_cfg_t;

int srf_batch_create(srf_plan *const *plans, const srf_batch_scan_t *scans, int32_t n_scans,
                     const srf_batch_cfg_t *cfg, srf_batch **batch, void *stream);
int srf_batch_destroy(srf_batch *batch);
int srf_batch_info(const srf_batch *batch, int32_t *n_scans, int32_t *n_work, int32_t *grid,
                   int64_t *n_events, size_t *device_bytes);
/* q = ratio(y, A x) for every scan (srf_plan_project_ratio semantics). */
int srf_batch_project_ratio(srf_batch *batch, srf_ratio_kind_t kind, float eps, void *stream);
/* x = x * A^T q / emap for every scan (srf_plan_backproject with
 * SRF_BP_REFERENCE, then srf_em_update with zero_guard, fused). */
int srf_batch_backproject_update(srf_batch *batch, float zero_guard, void *stream);
/* One EM sub-iteration for every scan: the two calls above. */
int srf_batch_em_step(srf_batch *batch, srf_ratio_kind_t kind, float eps, float zero_guard, void *stream);

/* ----------------------------------------------------- scatter estimate --
 * Single-scatter simulation on a virtual-detector sinogram: the per-LOR
 * Klein-Nishina sum of srfnef/corrections/scattering/scatter_fraction/
 * scatter_fraction.py:12-136 (loop_all_lors / loop_all_s) with the line
 * integrals of scatter/preprocess.py:21-43 supplied by the projector.
 *   det     (n_det, 4)   detector x, y, z (mm), normal angle (rad)
 *   points  (n_points,4) scatter point x, y, z (mm), mu at the point (1/mm)
 *   em_int, mu_int (n_det, n_points) emission / attenuation line integrals
 *   rows    workspace of srf_sss_rows_bytes() bytes (16-byte aligned)
 *   bins    (n_bins, 2)  detector pairs; out (n_bins)
 *   tables  543 floats: 511/E', window efficiency, KN/sigma_KN per integer
 *           scattering angle 0..180 degrees.
 * Below 32768 bins (a few rings, e.g. a surrogate's check subset) each bin's
 * sum over points is split across a wavefront and combined in a fixed order;
 * above, one lane sums one bin serially.  The two agree to float rounding. */
size_t srf_sss_rows_bytes(int32_t n_det, int32_t n_points);
int srf_sss_prepare(const float *det, int32_t n_det, const float *points, int32_t n_points,
                    const float *em_int, const float *mu_int, void *rows, size_t rows_bytes, void *stream);
int srf_sss_sinogram(const void *rows, int32_t n_det, const float *points, int32_t n_points,
                     const int32_t *bins, int32_t n_bins, const float *tables, float scale, float *out,
                     void *stream);

/* Scatter correction of list-mode values (ScatterCorrect.__call__,
 * srfnef/corrections/scattering/scatter.py:51-53):
 *   out[e] = (y[e] - sino[bins[e]] / S * Y * fraction) * weight[e]
 * with S = sum_e sino[bins[e]], Y = sum_e y[e] (double sums in work[2]);
 * y NULL = ones, weight NULL = ones (pass 1/attenuation to fold in the
 * attenuation correction); bins outside [0, n_bins) count as zero scatter.
 * S = 0 subtracts nothing.  out may alias y.  No memset: graph-safe. */
int srf_scatter_correct(const float *y, const float *weight, const int32_t *bins, const float *sino, int64_t n_bins,
                        int64_t n, float fraction, double *work, float *out, void *stream);

/* ---- Scatter surrogate CNN (BF16 MFMA, FP32 accumulation) -------------------
 * No counterpart in the reference: these train and run the 3D encoder-decoder
 * that stands in for srf_sss_sinogram (ScatterCorrect's estimate,
 * srfnef/corrections/scattering/scatter.py:25-55) inside the EM loop.
 * Tensors are channels-last, act[v][c] with v = (z*H + y)*W + x; bf16 tensors
 * are uint16 bit patterns.  Padded channel counts: inputs (cin_p, row stride of
 * x) are multiples of 8 with zero padding; weight tensors are [taps][cout_p][cin_p]
 * with cout_p a multiple of 32 and zero rows/columns beyond the real counts.
 * taps is 27 (3x3x3, zero padding, stride 1) or 1 (1x1x1). */

/* y[v][out_off + n] = act(bias[n] + sum_t sum_c x[v + o_t][c] w[t][n][c]) for
 * n < cout; act = leaky ReLU of `slope` (slope < 0: identity).  Writes y_bf16
 * and/or y_f32 (row stride out_stride).  bias may be NULL.  x and w 16-byte
 * aligned.  The _batch form convolves N samples of D x H x W stacked along z
 * (N*D slices; taps never reach across a sample's z faces). */
int srf_cnn_conv(const void *x, const void *w, const float *bias, int32_t D, int32_t H, int32_t W,
                 int32_t cin_p, int32_t cout, int32_t cout_p, int32_t taps, int32_t out_stride, int32_t out_off,
                 float slope, void *y_bf16, float *y_f32, void *stream);
int srf_cnn_conv_batch(const void *x, const void *w, const float *bias, int32_t N, int32_t D, int32_t H, int32_t W,
                       int32_t cin_p, int32_t cout, int32_t cout_p, int32_t taps, int32_t out_stride, int32_t out_off,
                       float slope, void *y_bf16, float *y_f32, void *stream);
/* srf_cnn_conv_batch of the concatenation [upsample2x(coarse), fine] read in
 * place: input channels [0, cc) are coarse (rows of cc channels on the
 * (D/2, H/2, W/2) grid of each sample, nearest neighbour), [cc, cin_p) are
 * fine (rows of cin_p - cc channels).  D, H, W even; cc a multiple of 8. */
int srf_cnn_conv_up_batch(const void *fine, const void *coarse, int32_t cc, const void *w, const float *bias,
                          int32_t N, int32_t D, int32_t H, int32_t W, int32_t cin_p, int32_t cout, int32_t cout_p,
                          int32_t taps, int32_t out_stride, int32_t out_off, float slope, void *y_bf16, float *y_f32,
                          void *stream);
/* srf_cnn_conv_batch into y_bf16 (rows of cout channels) followed by a 1x1
 * convolution to one channel in the same pass: head_out[v] = head_b[0] +
 * sum_n y_bf16[v][n] head_w[n] (head_w: cout bf16 weights; float sums, so
 * equal to srf_cnn_conv of y_bf16 with taps 1 up to float rounding).  cout <=
 * 64 with every channel in one workgroup; SRF_ERR_UNSUPPORTED otherwise. */
int srf_cnn_conv_head_batch(const void *x, const void *w, const float *bias, int32_t N, int32_t D, int32_t H,
                            int32_t W, int32_t cin_p, int32_t cout, int32_t cout_p, int32_t taps, float slope,
                            void *y_bf16, const void *head_w, const float *head_b, float *head_out, void *stream);
/* dw[t][n][c] += sum_v dy[v][n] x[v + o_t][c]   (dy bf16 rows of dy_stride);
 * the _batch form over N stacked samples, and with db != NULL also
 * db[n] += sum_v dy[v][n] (the bias gradient, as srf_cnn_bias_grad). */
int srf_cnn_conv_wgrad(const void *x, const void *dy, int32_t dy_stride, int32_t D, int32_t H, int32_t W,
                       int32_t cin_p, int32_t cout, int32_t cout_p, int32_t taps, float *dw, void *stream);
int srf_cnn_conv_wgrad_batch(const void *x, const void *dy, int32_t dy_stride, int32_t N, int32_t D, int32_t H,
                             int32_t W, int32_t cin_p, int32_t cout, int32_t cout_p, int32_t taps, float *dw,
                             float *db, void *stream);
/* wf[t][c][n] = w[taps-1-t][n][c]: the kernel whose srf_cnn_conv of dy is the
 * gradient with respect to x (wf is [taps][fcout_p][fcin_p]). */
int srf_cnn_flip_weights(const void *w, int32_t taps, int32_t cout, int32_t cout_p, int32_t cin_p, int32_t cin,
                         void *wf, int32_t fcout_p, int32_t fcin_p, void *stream);
/* 2x2x2 average pooling (bf16) and its gradient (f32; written, not added). */
int srf_cnn_pool(const void *x, int32_t D, int32_t H, int32_t W, int32_t C, int32_t xs, void *y, int32_t ys,
                 void *stream);
int srf_cnn_pool_bwd(const float *dy, int32_t D, int32_t H, int32_t W, int32_t C, int32_t dys, float *dx,
                     int32_t dxs, void *stream);
/* Nearest 2x upsampling of x (D,H,W) into channels [off, off+C) of y, and its
 * gradient (sum over the 8 children, f32, written). */
int srf_cnn_upsample(const void *x, int32_t D, int32_t H, int32_t W, int32_t C, int32_t xs, void *y, int32_t ys,
                     int32_t off, void *stream);
int srf_cnn_upsample_bwd(const float *dy, int32_t D, int32_t H, int32_t W, int32_t C, int32_t dys, int32_t off,
                         float *dx, int32_t dxs, void *stream);
/* y[v][yoff + c] (=|+=) x[v][xoff + c], c < C; f32 != 0: float tensors (only
 * those accumulate), else bf16. */
int srf_cnn_copy_channels(const void *x, int64_t V, int32_t C, int32_t xs, int32_t xoff, void *y, int32_t ys,
                          int32_t yoff, int32_t f32, int32_t accumulate, void *stream);
/* dp[v][c] = bf16(da[v][c] * (a[v][c] > 0 ? 1 : slope)) for c < C, 0 for
 * C <= c < Cp (slope < 0: no activation). */
int srf_cnn_act_bwd(const float *da, int32_t das, const void *a, int32_t as, int64_t V, int32_t C, int32_t Cp,
                    float slope, void *dp, void *stream);
/* db[c] += sum_v dp[v][c] for c < C (dp bf16 rows of Cp). */
int srf_cnn_bias_grad(const void *dp, int64_t V, int32_t C, int32_t Cp, float *db, void *stream);
/* loss[0] += w mean((p - t)^2); grad = 2 w (p - t) / n  (w = 1 / mean(t^2): a
 * relative loss). */
int srf_cnn_mse(const float *p, const float *t, int64_t n, float weight, float *grad, float *loss, void *stream);
/* One Adam step (Kingma & Ba, 2015) on FP32 master weights w with gradient
 * g * gscale; step counts from 1; wbf (may be NULL) receives bf16(w). */
int srf_cnn_adam(float *w, const float *g, float *m, float *v, int64_t n, float lr, float b1, float b2, float eps,
                 int32_t step, float gscale, void *wbf, void *stream);
/* Surrogate input: x[i][0..7] = bf16 of
 * [e, exp(-mu), e exp(-mu), mu/4, extra[i] * scale[1], 0, 0, 0] with
 * e = em_int[i] * scale[0] (scale is a device pointer; extra may be NULL: 0). */
int srf_cnn_features(const float *em_int, const float *mu_int, const float *extra, int64_t n, const float *scale,
                     void *x, void *stream);
/* f32 <-> bf16 (round to nearest even). */
int srf_cnn_cast(const void *x, int64_t n, int32_t to_bf16, void *y, void *stream);

/* ---- crystal geometry ---------------------------------------------------
 * LORs of the ordered crystal pairs p in [pair_start, pair_start + n_pairs):
 * lors[k] = (pos[p / N], pos[p % N]) (6 floats), N = n_crystals, from the
 * crystal-centre table crystal_pos (N x 3).  Replaces the host LOR arrays that
 * srfnef's EmapGenerator builds with np.kron (emap_generator.py:28-119,
 * 122-243); the pair order is theirs. */
int srf_pair_lors(const float *crystal_pos, int64_t n_crystals, int64_t pair_start, int64_t n_pairs, float *lors,
                  void *stream);

/* Block (ECAT) scanner as srfnef's PetEcatScanner properties
 * (ops/geometry/pet_ecat_scanner.py:13-47); lengths in mm. */
typedef struct {
    double angle_per_block;   /* 2 pi / nb_blocks_per_ring */
    double axial_length, gap;
    double block_size_y, block_size_z;
    double unit_y, unit_z;    /* block size / crystal grid */
    int32_t nb_blocks_per_ring, shape_y, nb_crystals_per_block, nb_crystals_per_ring;
} srf_ecat_t;
/* Cylindrical scanner as PetCylindricalScanner's properties
 * (ops/geometry/pet_cylindrical_scanner.py:13-74), [transaxial, z] pairs. */
typedef struct {
    double angle_per_rsector;
    double sz_rsector[2], mv_module[2], sz_module[2], mv_submodule[2], sz_submodule[2], mv_crystal[2];
    int32_t nb_rsector, nb_crystal[2], nb_submodule[2], nb_module[2];
    int32_t nb_crystal_per_submodule, nb_crystal_per_module, nb_crystal_per_rsector;
} srf_cyl_t;
/* ind[2e], ind[2e + 1] = crystal index of LOR e's endpoints (kind 0: ECAT,
 * EcatCrystalPosToIndex, functions/geometry/pet_ecat_scanner.py:15-62; kind 1:
 * cylindrical, CylindricalCrystalPosToIndex, pet_cylindrical_scanner.py:15-110),
 * in the reference's precision on float32 positions (the block / sector angle
 * and the ECAT ring in float32, the rest in float64; NumPy rounding).  Indices
 * may fall outside [0, nb_crystals) as the reference's do; a non-finite
 * position gives INT64_MIN. */
int srf_crystal_index(const float *lors, int64_t n, int64_t lor_stride, int32_t kind, const srf_ecat_t *ecat,
                      const srf_cyl_t *cyl, int64_t *ind, void *stream);
/* out = lors with both endpoints replaced by their crystals' centres (columns
 * past the sixth, e.g. TOF, copied): PetEcatListmodeTrans /
 * PetCylindricalListmodeTrans (functions/data_transform/pet_ecat.py:43-60,
 * pet_cylindrical.py:43-60).  An index outside the table makes the LOR NaN and
 * adds one to *outside (a device counter). */
int srf_snap_lors(const float *lors, int64_t n, int64_t lor_stride, const int64_t *ind, const float *crystal_pos,
                  int64_t n_crystals, float *out, uint64_t *outside, void *stream);
/* sino[(i - 1) i / 2 + j] += value (1 if values is NULL) for each event, with
 * (i, j) its endpoints' indices ordered i >= j: lm2sino's crystal-pair
 * sinogram over the N (N - 1) / 2 pairs of get_all_lors_id
 * (corrections/scattering/listmode2sinogram.py:17-35, 76-85).  Events whose
 * bin is outside the table add one to *outside (a device counter).  Double
 * atomics: sums of integer-valued weights are exact. */
int srf_pair_sinogram(const int64_t *ind, const float *values, int64_t n, int64_t n_crystals, double *sino,
                      uint64_t *outside, void *stream);
/* sino2lm (listmode2sinogram.py:88-99): the bins with 0 < value < 1e8, in bin
 * order, as LORs from crystal i to crystal j (6 floats) and their values.
 * *count receives the number of such bins; with lors == NULL nothing else is
 * written, otherwise capacity must hold *count.  Allocates its scratch;
 * synchronises `stream`. */
int srf_sinogram_lors(const double *sino, int64_t n_crystals, const float *crystal_pos, float *lors, double *values,
                      int64_t capacity, int64_t *count, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* SRF_HIP_H */
