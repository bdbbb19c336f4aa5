"""End-to-end TOF list-mode OSEM of many scans with scatter estimated inside
the loop: the plain per-scan method and the batched method that replaces it.

Per scan (both methods): values y_e = 1 / a_e (attenuation factors of the
mu-map along each event, AttenuationCorrect semantics with the physical line
integral); OSEM with n_sub subsets (OsemEngine semantics: seeded permutation,
q = y / (A x + 1e-8), emap zeros guarded by 1e8); before every iteration after the first,
a scatter sinogram is estimated from the current image and the mu-map and the
values become (1 - s_e / sum(s) * N * fraction) / a_e (ScatterCorrect,
srfnef/corrections/scattering/scatter.py:51-53; srf_scatter_correct).

  plain     per scan: OsemEngine launches per sub-iteration, SSS
            (SingleScatterSimulation) for every scatter estimate
  batched   one BatchOsem over all scans (one launch per direction per subset,
            the iteration captured in a HIP graph, autotuned schedule); the CNN
            surrogate for the scatter estimates, every estimate checked against
            SSS on a few rings -- an estimate whose check error exceeds the
            tolerance is replaced by the full SSS, and that scan uses SSS from
            then on (the fallback)

Scans follow the batched-reconstruction workload: event counts log-uniform in
[1e5, 2e7], in-plane sizes {96, 128, 192, 256}, 64-128 slices, isotropic
2-4 mm voxels, the random ellipsoid phantoms of the OSEM workload (water body,
mu 0.0096/mm), TOF FWHM 60 mm (400 ps).
"""
import math
import time

import numpy as np
import torch

from . import ops, phantom
from .batch import BatchConfig, BatchOsem, Scan, scan_permutation
from .data import TofConfig, Umap
from .functions import OsemEngine, uniform_cylinder_lors
from .sss import SingleScatterSimulation, SssGeometry, scatter_correct
from .surrogate import ScatterSurrogate, SinogramFeatures, relative_l2

RADIUS, AXIAL = 424.0, 300.0
TOF_BIN = 3.66
FWHM = 60.0
# Image grids may extend past the scanner's axial length, where the sensitivity
# image is 0: those voxels take the guard of mlem.py:33-35 instead of 0/0.
ZERO_GUARD = 1e8
# TOF window of the plans (srf_plan_opts_t.tof_window): terms below 2^-22 of
# the event's TOF peak are not binned (< 4e-8 of its weight: below one float32
# ulp of the event's total).
TOF_WINDOW = 22.0


def _log(*a):
    print(*a, flush=True)


def make_scans(n_scans=32, seed=5, ev_range=(1e5, 2e7), sizes=(96, 128, 192, 256), slices=(64, 128),
               voxel=(2.0, 4.0), n_emap=2_000_000, device=None, log=None):
    """The batched workload's scans on the GPU (Scan objects with extra
    fields in .meta: mu (x,y,z), inv_att, bins)."""
    from .corrections import attenuation_factors
    dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
    rng = np.random.default_rng(seed)
    g = SssGeometry()
    scans = []
    for j in range(n_scans):
        n = int(round(math.exp(rng.uniform(math.log(ev_range[0]), math.log(ev_range[1])))))
        nxy = int(rng.choice(sizes))
        nz = int(rng.integers(slices[0], slices[1] + 1))
        vox = float(rng.uniform(voxel[0], voxel[1]))
        shape = (nxy, nxy, nz)
        act, mu, center, size = phantom.random_ellipsoid_phantom(shape, vox, 1000 * seed + j, device=dev)
        lors = phantom.sample_events(act, center, size, n, RADIUS, AXIAL, 77 + 1000 * seed + j, tof_fwhm_mm=FWHM,
                                     device=dev)
        inv_att = (1.0 / attenuation_factors(Umap(mu, center, size), lors, physical=True)).contiguous()
        geom = list(center) + list(size)
        emap_k = torch.zeros((nz, nxy, nxy), dtype=torch.float32, device=dev)
        chunk = 1 << 21
        for s0 in range(0, n_emap, chunk):
            el = uniform_cylinder_lors(min(chunk, n_emap - s0), RADIUS, AXIAL, 5000 + s0 + 17 * j, dev)
            emap_k += ops.backproject(el, None, list(shape), geom, False, 0.0, 0.0, 0)
        emap_k /= float(n_emap)
        emap = ops.from_kernel_layout(emap_k)
        bins = g.bin_of_lors(lors).to(torch.int32).contiguous()
        sc = Scan(lors, shape, center, size, emap, values=inv_att.clone(), seed=j,
                  meta={"mu": mu, "inv_att": inv_att, "bins": bins, "voxel": vox, "act": act})
        scans.append(sc)
        if log:
            log(f"scan {j}: {n} events grid {shape} voxel {vox:.2f} mm")
    return scans


# Gaussian FWHM (mm) of the emission estimate that scatter is computed from.
SCATTER_SMOOTH_MM = 20.0


def body_estimate(x_xyz, mu_xyz, voxel=None, fwhm=SCATTER_SMOOTH_MM):
    """The emission estimate that scatter is computed from: restricted to the
    body (mu > 0; activity outside tissue is unphysical, and early OSEM iterates
    spread some there) and, with `voxel` (mm), smoothed with a Gaussian of
    `fwhm` mm (srf_smooth, weights renormalised at the image edge) -- early
    iterates hold isolated hot voxels that scatter, a smooth functional of the
    activity, does not resolve."""
    from ._lib import check, lib, stream_handle
    body = (mu_xyz > 0).to(torch.float32).contiguous()
    x = (x_xyz.to(torch.float32) * body).contiguous()
    if voxel is None or not fwhm:
        return x
    sig = float(fwhm) / 2.3548200450309493 / float(voxel)
    tmp = torch.empty_like(x)
    out = torch.empty_like(x)
    n0, n1, n2 = x.shape
    check(lib().srf_smooth(x.data_ptr(), n0, n1, n2, sig, sig, sig, body.data_ptr(), tmp.data_ptr(), out.data_ptr(),
                           stream_handle(x.device)), "srf_smooth")
    return out


class ScatterEstimator:
    """Scatter sinograms for scans of any grid on one virtual-detector
    geometry: full SSS, or the surrogate with an SSS check on `check_rings`,
    both from the emission estimate restricted to the body (body_estimate).
    SSS is cell-volume normalised (the same volume integral on every grid;
    identical to the plain sum on the surrogate's 4 mm training grid, where
    the scatter point cell is exactly 8 mm)."""

    def __init__(self, geometry=None, net=None, kappa=None, check_rings=(16,), tol=0.02, fraction=0.3,
                 coarse_spacing=None):
        self.g = geometry or SssGeometry()
        self.net, self.kappa = net, kappa
        self.coarse_spacing = coarse_spacing or None
        self.check_rings, self.tol, self.fraction = tuple(check_rings), float(tol), float(fraction)
        self._sss, self._chk, self._feat = {}, {}, {}
        self.fallback = {}        # scan key -> True once a check failed
        self.checks = []          # (scan key, estimate index of that scan, error)
        self._n_est = {}          # scan key -> estimates made

    @staticmethod
    def _key(sc):
        return (tuple(sc.shape), tuple(sc.center), tuple(sc.size))

    def sss(self, sc, x_xyz):
        return self._sss_of(sc, body_estimate(x_xyz, sc.meta["mu"], sc.meta.get("voxel")))

    def _sss_of(self, sc, x_xyz):
        """SSS of an already regularised estimate."""
        k = self._key(sc)
        if k not in self._sss:
            self._sss[k] = SingleScatterSimulation(self.g, sc.shape, sc.center, sc.size, cell_normalized=True)
        return self._sss[k](x_xyz, sc.meta["mu"])

    def surrogate(self, sc, x_xyz, scan_id):
        """One scan's surrogate estimate, checked (surrogate_batch of one)."""
        return self.surrogate_batch([(scan_id, sc, x_xyz)])[0]

    def _features(self, sc):
        k = self._key(sc)
        if k not in self._feat:
            self._feat[k] = SinogramFeatures(self.g, sc.shape, sc.center, sc.size, self.net.dev,
                                             coarse_spacing=self.coarse_spacing)
        return self._feat[k]

    def _checker(self, sc):
        k = self._key(sc)
        if k not in self._chk:
            self._chk[k] = SingleScatterSimulation(self.g, sc.shape, sc.center, sc.size, rings=self.check_rings,
                                                   cell_normalized=True)
        return self._chk[k]

    def surrogate_batch(self, items):
        """Surrogate estimates of several scans at once: per scan the body
        estimate, the features (written into one stacked input) and the
        check-ring SSS; then one CNN forward over all of them, every check
        error on the GPU and a single transfer of the errors to the host.
        items: [(scan_id, scan, x_xyz)].  Returns [(sinogram, fell_back)] in
        item order; scans whose check fails (now or before) get the full SSS."""
        res = [None] * len(items)
        todo = []
        for j, (scan_id, sc, x_xyz) in enumerate(items):
            xb = body_estimate(x_xyz, sc.meta["mu"], sc.meta.get("voxel"))
            if self.fallback.get(scan_id):
                res[j] = (self._sss_of(sc, xb), True)
            else:
                todo.append((j, scan_id, sc, xb))
        if not todo:
            return res
        dims = (self.g.n_rings, self.g.n_angles, self.g.n_radial)
        n = dims[0] * dims[1] * dims[2]
        X = torch.empty((len(todo) * n, 8), dtype=torch.int16, device=self.net.dev)
        ms, refs = [], []
        for i, (j, scan_id, sc, xb) in enumerate(todo):
            _, m = self._features(sc)(xb, sc.meta["mu"], out=X[i * n:(i + 1) * n])
            ms.append(m)
            if self.check_rings:
                refs.append(self._checker(sc)(xb, sc.meta["mu"]))
        out = self.net.forward(X, dims, len(todo))
        preds = out.reshape(len(todo), *dims) * (torch.cat(ms) / self.kappa).reshape(-1, 1, 1, 1)
        if self.check_rings:
            rings = list(self.check_rings)
            errs = torch.stack([torch.linalg.vector_norm((preds[i, rings] - r).double())
                                / torch.linalg.vector_norm(r.double()).clamp_min(1e-300)
                                for i, r in enumerate(refs)]).cpu().tolist()
        for i, (j, scan_id, sc, xb) in enumerate(todo):
            fell = False
            if self.check_rings:
                nn_ = self._n_est.get(scan_id, 0)
                self._n_est[scan_id] = nn_ + 1
                self.checks.append((scan_id, nn_, errs[i]))
                if not errs[i] <= self.tol:
                    self.fallback[scan_id] = True
                    fell = True
            res[j] = (self._sss_of(sc, xb), True) if fell else (preds[i], False)
        return res

    def values(self, sc, sino):
        return scatter_correct(None, sc.meta["inv_att"], sc.meta["bins"], sino, self.fraction)


def reconstruct_plain(scans, n_sub=8, n_iter=4, estimator=None, scatter="sss", item_records=8192, log=None):
    """Per-scan OsemEngine reconstructions (scan after scan; plans freed after
    each).  scatter: 'sss', 'surrogate' (same policy as the batched method) or
    None.  Returns (images (x,y,z) list, seconds)."""
    s2 = phantom.tof_sigma2_from_fwhm(FWHM)
    tof = TofConfig(TOF_BIN, s2)
    imgs = []
    t_sc = 0.0
    torch.cuda.synchronize()
    t0 = time.time()
    for i, sc in enumerate(scans):
        n = sc.lors.shape[0]
        perm = scan_permutation(n, sc.seed, sc.lors.device)
        eng = OsemEngine(sc.lors, sc.meta["inv_att"], sc.emap, sc.shape, sc.geom, n_sub, perm, tof=tof,
                         ratio_kind=0, eps=1e-8, zero_guard=ZERO_GUARD, item_records=item_records,
                         tof_window=TOF_WINDOW)
        pidx = eng.event_index
        for it in range(n_iter):
            if it >= 1 and scatter is not None:
                torch.cuda.synchronize()
                ts = time.time()
                x = eng.image()
                if scatter == "sss":
                    sino = estimator.sss(sc, x)
                else:
                    sino, _ = estimator.surrogate(sc, x, i)
                eng.y.copy_(estimator.values(sc, sino).index_select(0, pidx))
                torch.cuda.synchronize()
                t_sc += time.time() - ts
            eng.run(n_sub)
        imgs.append(eng.image().clone())
        for p in eng.plans:
            p.close()
        del eng
    torch.cuda.synchronize()
    t = time.time() - t0
    if log:
        log(f"plain ({scatter}): {t:.2f} s, of which scatter estimates {t_sc:.2f} s")
    return imgs, t


def reconstruct_batched(scans, n_sub=8, n_iter=4, estimator=None, config=BatchConfig(), graph=True, log=None,
                        scatter="surrogate"):
    """One BatchOsem over all scans.  Returns (images, seconds, BatchOsem info)."""
    s2 = phantom.tof_sigma2_from_fwhm(FWHM)
    torch.cuda.synchronize()
    t0 = time.time()
    for sc in scans:
        sc.values = sc.meta["inv_att"]
    b = BatchOsem(scans, n_sub, TOF_BIN, s2, True, config=config, zero_guard=ZERO_GUARD, tof_window=TOF_WINDOW)
    t_setup = time.time() - t0
    if graph:
        try:
            b.capture()
        except RuntimeError:
            pass   # b.advance runs eager launches and warns (graph_error)
    t_sc = 0.0
    for it in range(n_iter):
        if it >= 1 and scatter is not None:
            torch.cuda.synchronize()
            ts = time.time()
            if scatter == "sss":
                for i, sc in enumerate(scans):
                    b.set_values(i, estimator.values(sc, estimator.sss(sc, b.image(i))))
            else:
                est = estimator.surrogate_batch([(i, sc, b.image(i)) for i, sc in enumerate(scans)])
                for i, sc in enumerate(scans):
                    b.set_values(i, estimator.values(sc, est[i][0]))
            torch.cuda.synchronize()
            t_sc += time.time() - ts
        b.advance(graph)
    imgs = [b.image(i).clone() for i in range(len(scans))]
    torch.cuda.synchronize()
    t = time.time() - t0
    info = b.info(0)
    info["setup_s"] = round(t_setup, 3)
    info["scatter_s"] = round(t_sc, 3)
    b.close()
    if log:
        log(f"batched ({scatter}, graph={graph}): {t:.2f} s (plans {t_setup:.2f} s, scatter estimates {t_sc:.2f} s)")
    return imgs, t, info


def _main(argv=None):
    """Time scatter estimates on the batched workload's scans: per scan the
    plain SSS (its steps separately: scatter points and attenuation
    integrals, emission line integrals, the pair and bin kernels), then one
    surrogate estimate of all the scans together.  Each estimate runs twice;
    the second (warm: cached attenuation integrals, built plans) is timed."""
    import argparse
    import os
    from ._lib import check, lib, stream_handle
    from .surrogate import UNet
    ap = argparse.ArgumentParser()
    ap.add_argument("--scans", type=int, default=8)
    ap.add_argument("--seed", type=int, default=5)
    a = ap.parse_args(argv)
    dev = torch.device("cuda", 0)
    scans = make_scans(n_scans=32, seed=a.seed, device=dev)[:a.scans]
    here = os.path.dirname(os.path.abspath(__file__))
    net, meta = UNet.load(os.path.join(here, "weights", "scatter_unet_in_loop.npz"), device=dev)
    est = ScatterEstimator(net=net, kappa=meta["kappa"], coarse_spacing=meta.get("coarse_spacing") or None)
    xs = [torch.full(tuple(int(v) for v in sc.shape), 0.5, device=dev) for sc in scans]
    T = {}

    def tick(k, t0):
        torch.cuda.synchronize()
        T[k] = T.get(k, 0.0) + time.time() - t0
        return time.time()

    for rep in range(2):
        T.clear()
        for sc, x in zip(scans, xs):
            torch.cuda.synchronize()
            t = time.time()
            xb = body_estimate(x, sc.meta["mu"], sc.meta.get("voxel"))
            k = est._key(sc)
            if k not in est._sss:
                est._sss[k] = SingleScatterSimulation(est.g, sc.shape, sc.center, sc.size, cell_normalized=True)
            s = est._sss[k]
            t = tick("body estimate (smoothing)", t)
            pts, mu = s._mu_part(sc.meta["mu"])
            t = tick("sss: scatter points + attenuation integrals (cached after the first)", t)
            em = s.line_integrals(xb, pts)
            t = tick("sss: emission line integrals", t)
            rows = s._rows_for(pts.shape[0])
            st = stream_handle(dev)
            check(lib().srf_sss_prepare(s.det.data_ptr(), s.nd, pts.data_ptr(), pts.shape[0], em.data_ptr(),
                                        mu.data_ptr(), rows.data_ptr(), rows.numel() * 4, st), "srf_sss_prepare")
            t = tick("sss: srf_sss_prepare (k_sss_pairs)", t)
            out = torch.zeros(s.bins.shape[0], device=dev)
            check(lib().srf_sss_sinogram(rows.data_ptr(), s.nd, pts.data_ptr(), pts.shape[0], s.bins.data_ptr(),
                                         s.bins.shape[0], s.tables.data_ptr(), float(s.scale), out.data_ptr(), st),
                  "srf_sss_sinogram")
            t = tick("sss: srf_sss_sinogram (k_sss_bins)", t)
        torch.cuda.synchronize()
        t = time.time()
        est.surrogate_batch([(i, sc, x) for i, (sc, x) in enumerate(zip(scans, xs))])
        tick(f"surrogate: all {len(scans)} scans in one batch (features, check ring, CNN)", t)
    for k, v in T.items():
        per = v * 1e3 / (1 if k.startswith("surrogate") else len(scans))
        print(f"{k:75s} {per:9.2f} ms{'' if k.startswith('surrogate') else ' per scan'}", flush=True)


if __name__ == "__main__":
    _main()


__all__ = ["make_scans", "ScatterEstimator", "reconstruct_plain", "reconstruct_batched"]
