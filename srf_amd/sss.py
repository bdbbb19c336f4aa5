"""Single-scatter simulation (SSS) and scatter correction.

The physics is the reference's (srfnef/corrections/scattering/):
  * scatter points on a regular grid of the emission image
    (preprocess.py:6-18 pre_all_scatter_position: step = floor(spacing/voxel));
  * emission and attenuation line integrals from every scatter point to every
    detector, computed by the same projector that reconstructs
    (preprocess.py:21-43 uses Project / UmapProject for them);
  * per LOR a Klein-Nishina weighted sum over scatter points
    (scatter_fraction/scatter_fraction.py:40-131, kn_formula.py), evaluated by
    the HIP kernel srf_sss_sinogram;
  * correction by subtracting the normalised scatter estimate scaled to a
    fixed scatter fraction of the counts (scatter.py:51-53).

Where the reference runs over every crystal pair of a block scanner, this
evaluates a coarse virtual-detector sinogram: n_det detectors per ring on
n_rings rings of the scanner cylinder, direct planes only, n_angles x
n_radial bins per ring; bins snap to the nearest detectors.  Attenuation uses
the physical line integral of mu (the reference's UmapProject scaling
L^2/unit_size[0] is exact only for 1 mm voxels, quirk Q15).
"""
import math
import os

import numpy as np
import torch

from . import ops
from ._lib import check, lib, stream_handle
from .data import Listmode, nef_class

SIGMA_KN = math.pi * (40.0 / 9.0 - 3.0 * math.log(3.0))


def energy_tables(low, high, resolution, physical_kn=False):
    """(543,) float32: per integer scattering angle 0..180 deg: 511/E',
    window efficiency eff(E'), KN(c)/sigma_KN with c the quantised angle used
    as a cosine (kn_formula.py:12,16; quirk Q16) or cos of it (physical_kn)."""
    k = np.arange(181, dtype=np.float64)
    th = k / 180.0 * math.pi
    c = np.cos(th)
    Ep = 511.0 / (2.0 - c)
    cc = c if physical_kn else th
    kn = 1.0 / (2.0 - cc) ** 2 * ((1.0 + cc ** 2) / 2.0 + (1.0 - cc) ** 2 / 2.0 / (2.0 - cc)) / SIGMA_KN
    s = math.sqrt(511.0) * resolution
    Ei = np.arange(int(low), int(high), 5, dtype=np.float64)
    eff = np.sum(np.exp(-(Ei[None, :] - Ep[:, None]) ** 2 / 2.0 / s ** 2), axis=1)
    return np.concatenate([511.0 / Ep, eff, kn]).astype(np.float32)


def sss_scale(low, high, resolution):
    """5/sqrt(1022 pi)/resolution * eff(511) / (4 pi).  The reference's constant
    (scatter_fraction.py:27-33) is eff_ws * area^2 * 5/sqrt(1022 pi)/resolution
    / (4 pi) with eff_ws = eff(511) * 5/sqrt(1022 pi)/resolution: this one drops
    a factor 5/sqrt(1022 pi)/resolution and the crystal area squared (virtual
    detectors have no crystal area).  Both cancel wherever the estimate is
    normalised (ScatterCorrect divides by its sum); the crystal-pair path
    (srf_amd.scatter_pairs) applies the full constant."""
    s = math.sqrt(511.0) * resolution
    Ei = np.arange(int(low), int(high), 5, dtype=np.float64)
    e511 = float(np.sum(np.exp(-(Ei - 511.0) ** 2 / 2.0 / s ** 2)))
    return 5.0 / math.sqrt(1022.0 * math.pi) / resolution * e511 / (4 * math.pi)


@nef_class
class SssGeometry:
    """Virtual detectors and the coarse sinogram (defaults: 424 mm radius,
    300 mm axial, 128 detectors x 32 rings, 64 angles x 64 radial bins)."""
    radius: float = 424.0
    axial_length: float = 300.0
    n_det: int = 128
    n_rings: int = 32
    n_angles: int = 64
    n_radial: int = 64
    fov_radius: float = 256.0
    spacing: float = 8.0
    low_energy: float = 400.0
    high_energy: float = 650.0
    resolution: float = 0.15

    def detectors(self, device):
        k = torch.arange(self.n_det, dtype=torch.float64)
        r = torch.arange(self.n_rings, dtype=torch.float64)
        phi = 2 * math.pi * k / self.n_det
        z = (r + 0.5) * self.axial_length / self.n_rings - self.axial_length / 2
        det = torch.zeros((self.n_rings, self.n_det, 4), dtype=torch.float64)
        det[:, :, 0] = self.radius * torch.cos(phi)[None, :]
        det[:, :, 1] = self.radius * torch.sin(phi)[None, :]
        det[:, :, 2] = z[:, None]
        det[:, :, 3] = phi[None, :]
        return det.reshape(-1, 4).to(torch.float32).to(device)

    def bin_detectors(self):
        """(n_rings*n_angles*n_radial, 2) detector pairs of the sinogram bins
        (ring-major, then angle, then radial)."""
        a = (np.arange(self.n_angles) + 0.0) * math.pi / self.n_angles
        ds = 2.0 * self.fov_radius / self.n_radial
        s = (np.arange(self.n_radial) + 0.5 - self.n_radial / 2) * ds
        A, S = np.meshgrid(a, s, indexing="ij")
        t = np.sqrt(np.maximum(self.radius ** 2 - S ** 2, 0.0))
        nx, ny = np.cos(A), np.sin(A)
        dx, dy = -ny, nx
        psi1 = np.arctan2(S * ny - t * dy, S * nx - t * dx)
        psi2 = np.arctan2(S * ny + t * dy, S * nx + t * dx)
        k1 = np.mod(np.rint(psi1 / (2 * math.pi) * self.n_det), self.n_det).astype(np.int64)
        k2 = np.mod(np.rint(psi2 / (2 * math.pi) * self.n_det), self.n_det).astype(np.int64)
        rings = np.arange(self.n_rings)[:, None, None] * self.n_det
        b = np.stack([rings + k1[None], rings + k2[None]], -1).reshape(-1, 2)
        return b.astype(np.int32)

    def bin_of_lors(self, lors):
        """Sinogram bin of each LOR (N,>=6) tensor: transaxial angle and signed
        distance of the line, ring of the LOR midpoint (single-slice rebinning)."""
        p1, p2 = lors[:, 0:3].double(), lors[:, 3:6].double()
        d = p2[:, :2] - p1[:, :2]
        n = torch.stack([-d[:, 1], d[:, 0]], 1)
        n = n / torch.clamp(n.norm(dim=1, keepdim=True), min=1e-12)
        phi = torch.atan2(n[:, 1], n[:, 0])
        flip = (phi < 0) | (phi >= math.pi)
        n = torch.where(flip[:, None], -n, n)
        phi = torch.atan2(n[:, 1], n[:, 0]).clamp(0, math.pi - 1e-12)
        s = (p1[:, :2] * n).sum(1)
        a = torch.clamp((phi / math.pi * self.n_angles).long(), 0, self.n_angles - 1)
        rho = torch.clamp(((s + self.fov_radius) / (2 * self.fov_radius) * self.n_radial).floor().long(), 0,
                          self.n_radial - 1)
        zm = 0.5 * (p1[:, 2] + p2[:, 2])
        r = torch.clamp(((zm + self.axial_length / 2) / self.axial_length * self.n_rings).floor().long(), 0,
                        self.n_rings - 1)
        return (r * self.n_angles + a) * self.n_radial + rho

    def point_grid(self, shape, center, size, device, physical=False):
        """Scatter point grid: positions (ns, 3) float32, the flat (x, y, z)-
        order index of each point's voxel (ns,) int64, and the cell volume in
        mm^3.

        physical=False: preprocess.py:6-18 -- step = floor(spacing / voxel)
        voxels, so the points (and the sum) depend on the image grid.
        physical=True: a point every `spacing` mm from the image's lower
        corner, mu from the voxel holding it; the same points on every grid
        (identical to the first where spacing is a whole number of voxels)."""
        vox = [size[i] / shape[i] for i in range(3)]
        if physical:
            cnt = [max(1, int(math.floor(size[i] / self.spacing + 1e-9))) for i in range(3)]
            g = [torch.arange(cnt[i], dtype=torch.float64) for i in range(3)]
            X, Y, Z = torch.meshgrid(*g, indexing="ij")
            off = [(t + 0.5) * self.spacing for t in (X, Y, Z)]
            idx = [torch.clamp(torch.floor(off[i] / vox[i] + 1e-9), 0, shape[i] - 1).long().reshape(-1)
                   for i in range(3)]
            cell = self.spacing ** 3
        else:
            step = [max(1, int(math.floor(self.spacing / vox[i]))) for i in range(3)]
            cnt = [int(shape[i] // step[i]) for i in range(3)]
            g = [torch.arange(cnt[i], dtype=torch.float64) for i in range(3)]
            X, Y, Z = torch.meshgrid(*g, indexing="ij")
            off = [step[i] * (t + 0.5) * vox[i] for i, t in enumerate((X, Y, Z))]
            idx = [(step[i] * t + step[i] // 2).long().reshape(-1) for i, t in enumerate((X, Y, Z))]
            cell = step[0] * vox[0] * step[1] * vox[1] * step[2] * vox[2]
        pos = torch.stack([off[i] - size[i] / 2 + center[i] for i in range(3)], -1).reshape(-1, 3)
        flat = (idx[0] * shape[1] + idx[1]) * shape[2] + idx[2]
        return pos.to(torch.float32).to(device), flat.to(device), cell

    def scatter_points(self, mu_image, center, size, device):
        """(ns, 4): grid points with mu at the point's voxel; points outside
        the body carry mu = 0 and drop out of the sum."""
        mu = torch.as_tensor(mu_image).to(device, torch.float32)
        pos, flat, _ = self.point_grid(mu.shape, center, size, device)
        return torch.cat([pos, mu.reshape(-1)[flat][:, None]], 1).contiguous()


class SingleScatterSimulation:
    """SSS sinogram of (emission, mu) images on a fixed geometry.

    compact=True (default) keeps only the scatter points with mu > 0 for each
    call (the sum skips the others, so the sinogram is the same) and builds
    the (detector, point) LORs for those; compact=False projects the full grid
    and, with tiled=True, bins that fixed LOR set once into a projector plan.
    rings: evaluate only the bins of these rings (their detectors only);
    the output is then (len(rings), n_angles, n_radial)."""

    def __init__(self, geometry, shape, center, size, physical_kn=False, tiled=False, compact=True, rings=None,
                 cell_normalized=False, physical_grid=False, cell_ref=None):
        self.g = geometry
        self.shape = [int(v) for v in shape]
        self.center = [float(v) for v in center]
        self.size = [float(v) for v in size]
        self.geom = self.center + self.size
        self.dev = torch.device("cuda", torch.cuda.current_device())
        self.compact = bool(compact)
        det = geometry.detectors(self.dev)
        bins = geometry.bin_detectors().astype(np.int64)
        if rings is not None:
            rings = [int(r) for r in rings]
            if any(r < 0 or r >= geometry.n_rings for r in rings):
                raise ValueError(f"rings {rings} outside 0..{geometry.n_rings - 1}")
            det = det.reshape(geometry.n_rings, geometry.n_det, 4)[rings].reshape(-1, 4)
            per = geometry.n_angles * geometry.n_radial
            bins = np.concatenate([bins[r * per:(r + 1) * per] - r * geometry.n_det + k * geometry.n_det
                                   for k, r in enumerate(rings)])
        self.rings = rings
        self.det = det.contiguous()
        self.nd = self.det.shape[0]
        self.n_out_rings = len(rings) if rings is not None else geometry.n_rings
        self.dx_ = self.size[0] / self.shape[0]
        # the full scatter grid (mu decides which points count)
        self.pts_xyz, self.pts_flat, self.cell = geometry.point_grid(self.shape, self.center, self.size, self.dev,
                                                                     physical=physical_grid)
        self.ns = self.pts_xyz.shape[0]
        self.plan = None
        if not self.compact:
            # p1 = detector, p2 = scatter point: the projector's segment cut ends the
            # integral at the point (siddon.cu.cc:56-57; preprocess.py:46-54 order)
            self.lors, self.to_integral = self._lors(self.pts_xyz)
            self.plan = ops.Plan(self.lors, self.shape, self.geom, False) if tiled else None
        self.bins = torch.from_numpy(bins.astype(np.int32)).to(self.dev).contiguous()
        self.tables = torch.from_numpy(energy_tables(geometry.low_energy, geometry.high_energy,
                                                     geometry.resolution, physical_kn)).to(self.dev)
        self.scale = sss_scale(geometry.low_energy, geometry.high_energy, geometry.resolution)
        if cell_normalized:
            # weight each point by its cell volume in units of a reference cell
            # (default spacing^3): the sum then approximates the same volume
            # integral on every image grid and point spacing (the factor is 1 where
            # the point step times the voxel equals the reference edge)
            self.scale *= self.cell / (geometry.spacing ** 3 if cell_ref is None else float(cell_ref))

    def _lors(self, pts_xyz):
        ns = pts_xyz.shape[0]
        d = self.det[:, None, :3].expand(self.nd, ns, 3)
        sp = pts_xyz[None, :, :].expand(self.nd, ns, 3)
        lors = torch.cat([d, sp], -1).reshape(-1, 6).contiguous()
        L2 = ((lors[:, 3:6] - lors[:, 0:3]) ** 2).sum(1)
        return lors, (L2 / (self.dx_ * self.dx_)).contiguous()     # Project * L^2 / dx_^2 = integral

    def _integrals(self, img_xyz, lors, to_integral, plan=None):
        img_k = ops.to_kernel_layout(img_xyz.contiguous())
        if plan is not None:
            p = plan.project(img_k)
        else:
            p = ops.project(lors, img_k, self.shape, self.geom, False, 0.0, 0.0)
        return (p * to_integral).reshape(self.nd, -1).contiguous()

    def line_integrals(self, img_xyz, pts=None):
        """(nd, ns) integrals from each detector to each point of `pts`
        ((ns, 4) or (ns, 3); default: the full grid)."""
        img_xyz = img_xyz.to(self.dev, torch.float32)
        if pts is None and not self.compact:
            return self._integrals(img_xyz, self.lors, self.to_integral, self.plan)
        xyz = self.pts_xyz if pts is None else pts[:, :3].contiguous()
        lors, ti = self._lors(xyz)
        return self._integrals(img_xyz, lors, ti)

    def points(self, mu_xyz):
        mu = mu_xyz.detach().to(self.dev, torch.float32).reshape(-1)[self.pts_flat]
        m = torch.cat([self.pts_xyz, mu[:, None]], 1)
        if self.compact:
            m = m[mu != 0]
        return m.contiguous()

    def _rows_for(self, ns):
        # a fresh workspace per call (the caching allocator recycles it): an
        # estimator that keeps one simulation per image grid holds no memory
        need = int(lib().srf_sss_rows_bytes(self.nd, max(ns, 1))) // 4 + 4
        return torch.empty(need, dtype=torch.float32, device=self.dev)

    # largest kept set of attenuation integrals (floats): 2^28 (1 GiB) holds a
    # full set of 4096 detectors x 6.5e4 scatter points, so repeated estimates
    # for one mu-map (every OSEM iteration of a scan) skip half the projections
    MU_CACHE_MAX = 1 << 28

    def _mu_part(self, mu_xyz):
        """The scatter points and the attenuation line integrals of mu_xyz.
        They depend on the mu-map alone, so the last ones are kept and reused
        while the same tensor comes back unmodified (same object, same
        in-place version counter; the cache holds a reference, so the tensor
        cannot be freed and its address reused meanwhile)."""
        c = getattr(self, "_mu_cache", None)
        if c is not None and c[0] is mu_xyz and c[1] == mu_xyz._version:
            return c[2], c[3]
        pts = self.points(mu_xyz)
        mu = None
        if pts.shape[0] > 0:
            mu = self.line_integrals(mu_xyz, None if not self.compact else pts)
        # a full-size set (every detector to every point) is not kept
        keep = mu is None or mu.numel() <= self.MU_CACHE_MAX
        self._mu_cache = (mu_xyz, mu_xyz._version, pts, mu) if keep else None
        return pts, mu

    def __call__(self, emission_xyz, mu_xyz):
        """(n_rings or len(rings), n_angles, n_radial) float32 scatter sinogram."""
        pts, mu = self._mu_part(mu_xyz)
        out = torch.zeros(self.bins.shape[0], dtype=torch.float32, device=self.dev)
        shape = (self.n_out_rings, self.g.n_angles, self.g.n_radial)
        ns = pts.shape[0]
        if ns == 0:
            return out.reshape(shape)
        sel = None if not self.compact else pts
        em = self.line_integrals(emission_xyz, sel)
        rows = self._rows_for(ns)
        sref = stream_handle(self.dev)
        check(lib().srf_sss_prepare(self.det.data_ptr(), self.nd, pts.data_ptr(), ns, em.data_ptr(),
                                    mu.data_ptr(), rows.data_ptr(), rows.numel() * 4, sref), "srf_sss_prepare")
        check(lib().srf_sss_sinogram(rows.data_ptr(), self.nd, pts.data_ptr(), ns, self.bins.data_ptr(),
                                     self.bins.shape[0], self.tables.data_ptr(), float(self.scale), out.data_ptr(),
                                     sref), "srf_sss_sinogram")
        return out.reshape(shape)

    def inputs_for_oracle(self, emission_xyz, mu_xyz):
        """The tensors the SSS kernel consumes, on the host (tests)."""
        pts = self.points(mu_xyz)
        sel = None if not self.compact else pts
        em = self.line_integrals(emission_xyz, sel).cpu().numpy()
        mu = self.line_integrals(mu_xyz, sel).cpu().numpy()
        return self.det.cpu().numpy(), pts.cpu().numpy(), em, mu, self.bins.cpu().numpy()


@nef_class
class ScatterCorrect:
    """Scatter correction with srfnef's fields (scatter.py:18-55).

    __call__(emission_image, u_map, scanner, outdir, out_append=None) -> Listmode,
    by the type of the fourth argument:
      * a directory (str) holding scatter_preprocess's files, or the dict that
        srf_amd.scatter_pairs.scatter_preprocess returns: the reference's
        crystal-pair correction (scatter.py:25-55; srf_amd.scatter_pairs): the
        counted crystal pairs' sinogram minus the normalised per-pair SSS,
        times each pair's scale over its 1.3 mu attenuation, as a listmode of
        crystal-centre LORs;
      * a Listmode: the listmode values minus the SSS estimate at each event's
        bin of a coarse virtual-detector sinogram (the geometry field),
        normalised so the subtracted total is scatter_fraction of the counts
        (scatter.py:52-53); `scanner` is unused.  MlemFullTof passes its
        listmode there (mlem_full_tof.py:73), where the reference's np.load of
        outdir + '/lors.npy' would fail."""
    low_energy: float = 400.0
    high_energy: float = 650.0
    resolution: float = 0.15
    scatter_fraction: float = 0.3
    geometry: SssGeometry = None

    def estimate(self, emission_image, u_map):
        g = self.geometry or SssGeometry()
        g = g.update(low_energy=self.low_energy, high_energy=self.high_energy, resolution=self.resolution)
        sss = SingleScatterSimulation(g, emission_image.shape, emission_image.center, emission_image.size)
        em = torch.as_tensor(np.asarray(emission_image.data) if not isinstance(emission_image.data, torch.Tensor)
                             else emission_image.data, dtype=torch.float32)
        mu = torch.as_tensor(np.asarray(u_map.data) if not isinstance(u_map.data, torch.Tensor) else u_map.data,
                             dtype=torch.float32)
        return sss(em, mu), g

    def __call__(self, emission_image, u_map, scanner=None, outdir=None, out_append=None):
        if isinstance(outdir, (str, os.PathLike, dict)):
            from . import scatter_pairs as sp
            pre = outdir if isinstance(outdir, dict) else sp.load_preprocessed(outdir, out_append)
            frac, scale = sp.pair_scatter(pre, u_map, scanner, self.low_energy, self.high_energy, self.resolution)
            return sp.correct_pairs(pre, frac, scale, self.scatter_fraction)[1]
        listmode = outdir
        sino, g = self.estimate(emission_image, u_map)
        lors = torch.as_tensor(np.asarray(listmode.lors.data) if not isinstance(listmode.lors.data, torch.Tensor)
                               else listmode.lors.data).to(sino.device, torch.float32)
        y = torch.as_tensor(np.asarray(listmode.data) if not isinstance(listmode.data, torch.Tensor)
                            else listmode.data).to(sino.device, torch.float32).reshape(-1)
        bins = g.bin_of_lors(lors).to(torch.int32).contiguous()
        corr = scatter_correct(y, None, bins, sino, self.scatter_fraction)
        data = corr if isinstance(listmode.data, torch.Tensor) else corr.cpu().numpy()
        return Listmode(data, listmode.lors)


def scatter_correct(y, weight, bins, sino, fraction, out=None):
    """(y - sino[bins] / sum(sino[bins]) * sum(y) * fraction) * weight per event
    (scatter.py:51-53), by srf_scatter_correct.  y / weight None = ones."""
    n = bins.numel()
    dev = bins.device
    out = torch.empty(n, dtype=torch.float32, device=dev) if out is None else out
    work = torch.empty(2, dtype=torch.float64, device=dev)
    sino = sino.reshape(-1).contiguous()
    check(lib().srf_scatter_correct(y.data_ptr() if y is not None else None,
                                    weight.data_ptr() if weight is not None else None, bins.data_ptr(),
                                    sino.data_ptr(), sino.numel(), n, float(fraction), work.data_ptr(),
                                    out.data_ptr(), stream_handle(dev)), "srf_scatter_correct")
    return out


__all__ = ["SssGeometry", "SingleScatterSimulation", "ScatterCorrect", "energy_tables", "sss_scale",
           "scatter_correct"]
