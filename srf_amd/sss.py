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
    """5/sqrt(1022 pi)/resolution * eff(511) / (4 pi) (scatter_fraction.py:27-33)."""
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

    def scatter_points(self, mu_image, center, size, device):
        """(ns, 4): grid points of preprocess.py:6-18 (step = floor(spacing /
        voxel) voxels) with mu at the point's voxel; points outside the body
        carry mu = 0 and drop out of the sum."""
        shape = mu_image.shape
        vox = [size[i] / shape[i] for i in range(3)]
        step = [max(1, int(math.floor(self.spacing / vox[i]))) for i in range(3)]
        cnt = [int(shape[i] // step[i]) for i in range(3)]
        g = [torch.arange(cnt[i], dtype=torch.float64) for i in range(3)]
        X, Y, Z = torch.meshgrid(*g, indexing="ij")
        pos = torch.stack([step[0] * (X + 0.5) * vox[0] - size[0] / 2 + center[0],
                           step[1] * (Y + 0.5) * vox[1] - size[1] / 2 + center[1],
                           step[2] * (Z + 0.5) * vox[2] - size[2] / 2 + center[2]], -1).reshape(-1, 3)
        ix = (step[0] * X + step[0] // 2).long().reshape(-1)
        iy = (step[1] * Y + step[1] // 2).long().reshape(-1)
        iz = (step[2] * Z + step[2] // 2).long().reshape(-1)
        mu = torch.as_tensor(mu_image, device="cpu").to(torch.float64)[ix, iy, iz]
        out = torch.cat([pos, mu[:, None]], 1).to(torch.float32)
        return out.to(device)


class SingleScatterSimulation:
    """SSS sinogram of (emission, mu) images on a fixed geometry.

    compact=True (default) keeps only the scatter points with mu > 0 for each
    call (the sum skips the others, so the sinogram is the same) and builds
    the (detector, point) LORs for those; compact=False projects the full grid
    and, with tiled=True, bins that fixed LOR set once into a projector plan.
    rings: evaluate only the bins of these rings (their detectors only);
    the output is then (len(rings), n_angles, n_radial)."""

    def __init__(self, geometry, shape, center, size, physical_kn=False, tiled=False, compact=True, rings=None):
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
        grid = geometry.scatter_points(np.zeros(self.shape, np.float32), self.center, self.size, self.dev)
        self.pts_xyz = grid[:, :3].contiguous()
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
        self._rows = None

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
        m = self.g.scatter_points(mu_xyz.detach().cpu().numpy(), self.center, self.size, self.dev)
        if self.compact:
            m = m[m[:, 3] != 0]
        return m.contiguous()

    def _rows_for(self, ns):
        need = int(lib().srf_sss_rows_bytes(self.nd, max(ns, 1))) // 4 + 4
        if self._rows is None or self._rows.numel() < need:
            self._rows = torch.empty(need, dtype=torch.float32, device=self.dev)
        return self._rows

    def __call__(self, emission_xyz, mu_xyz):
        """(n_rings or len(rings), n_angles, n_radial) float32 scatter sinogram."""
        pts = self.points(mu_xyz)
        out = torch.zeros(self.bins.shape[0], dtype=torch.float32, device=self.dev)
        shape = (self.n_out_rings, self.g.n_angles, self.g.n_radial)
        ns = pts.shape[0]
        if ns == 0:
            return out.reshape(shape)
        sel = None if not self.compact else pts
        em = self.line_integrals(emission_xyz, sel)
        mu = self.line_integrals(mu_xyz, sel)
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

    __call__(emission_image, u_map, scanner, listmode) -> Listmode: the
    listmode values minus the SSS estimate at each event's sinogram bin,
    normalised so the subtracted total is scatter_fraction of the counts
    (scatter.py:52-53).  `scanner` is unused (the geometry field describes the
    virtual detectors); the fourth argument is the listmode, as
    MlemFull/MlemFullTof pass it (mlem_full_tof.py:118)."""
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

    def __call__(self, emission_image, u_map, scanner=None, listmode=None):
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
