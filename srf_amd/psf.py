"""Post-reconstruction PSF deconvolution (srfnef corrections/psf): a
separable, spatially varying blur and Richardson-Lucy iterations on the GPU.

Reference: corrections/psf/deconvolute.py:28-67 (Deconvolute.__call__),
:72-82 (kernel construction), :90-103 (the in-plane kernel), :163-213
(xy_main), :215-235 (z_main), :237-239 (the axial kernel);
psf_correct.py:16-30 (PsfCorrect); fitted_x.py / fitted_y.py / fitted_z.py
(the fitted parameter tables).

The forward model of image x (rows = in-plane pixels i * ny + j, columns =
slices) is A(x) = K_xy x K_z^T: K_z spreads each slice with a Gaussian whose
width is interpolated at the slice's |z|, K_xy spreads each pixel over its
11 x 11 neighbourhood with a Gaussian that is asymmetric along the radius
(widths sigx0 inward, sigx1 outward, sigy tangential, interpolated at the
pixel's radius).  Each iteration is x <- x * A^T(d / (A(x) + 1e-16)) from
x = 1.

Here K_xy is kept as a per-pixel table of its 121 patch weights (the sparse
matrix's columns, zero where the reference stores nothing), applied by
srf_psf_xy (a gather stencil, in both directions); K_z is a small dense matrix
applied by GEMM.  The kernels are built on the host from the fitted tables,
as the reference builds them with NumPy / SciPy; the iterations run on the
GPU.

Not on this path: PsfFit (fitting the tables to point-source images read from
HDF5 files) -- PsfCorrect takes fitted tables or a built Deconvolute.
"""
import ctypes

import numpy as np
import torch

from ._lib import check, lib, stream_handle
from .data import Image, nef_class
from .functions import device

HALF_PATCH = 5          # xy_main(image, half_patch_range=5), deconvolute.py:70


@nef_class
class FittedX:
    data: np.ndarray

    @property
    def ax(self):
        return self.data[:, 0]

    @property
    def sigx0(self):
        return self.data[:, 1]

    @property
    def sigx1(self):
        return self.data[:, 2]

    @property
    def ux(self):
        return self.data[:, 3]


@nef_class
class FittedY:
    data: np.ndarray

    @property
    def ay(self):
        return self.data[:, 0]

    @property
    def sigy(self):
        return self.data[:, 1]

    @property
    def uy(self):
        return self.data[:, 2]


@nef_class
class FittedZ:
    data: np.ndarray

    @property
    def az(self):
        return self.data[:, 0]

    @property
    def sigz(self):
        return self.data[:, 1]

    @property
    def uz(self):
        return self.data[:, 2]


def _interp(x, y):
    from scipy import interpolate
    return interpolate.interp1d(x, y, fill_value="extrapolate")


def xy_weights(image, fitted_x, fitted_y, half=HALF_PATCH):
    """(nx * ny, (2 half + 1)^2) float32: column p of the reference's K_xy as
    the weights of source pixel p on its patch, offset (a, b) at index
    (a + half) * (2 half + 1) + (b + half); 0 outside the image, below the
    1e-6 cut and for pixels beyond the largest fitted radius."""
    nx, ny = int(image.shape[0]), int(image.shape[1])
    ux, uy = float(image.unit_size[0]), float(image.unit_size[1])
    x = (np.arange(nx) + 0.5) * ux - image.size[0] / 2 + image.center[0]
    y = (np.arange(ny) + 0.5) * uy - image.size[1] / 2 + image.center[1]
    xm, ym = np.meshgrid(x, y, indexing="ij")
    r = np.sqrt(xm ** 2 + ym ** 2).reshape(-1)
    phi = np.arctan2(ym, xm).reshape(-1)
    s0 = _interp(fitted_x.ux, fitted_x.sigx0)(r)
    s1 = _interp(fitted_x.ux, fitted_x.sigx1)(r)
    sy = _interp(fitted_x.ux, fitted_y.sigy)(r)
    k = 2 * half + 1
    off = np.arange(-half, half + 1)
    ii = np.repeat(np.arange(nx), ny)
    jj = np.tile(np.arange(ny), nx)
    pi = ii[:, None, None] + off[None, :, None]            # (P, k, 1)
    pj = jj[:, None, None] + off[None, None, :]            # (P, 1, k)
    inside = (pi >= 0) & (pi < nx) & (pj >= 0) & (pj < ny)  # (P, k, k)
    xs0 = (pi + 0.5) * ux - image.size[0] / 2 + image.center[0]
    ys0 = (pj + 0.5) * uy - image.size[1] / 2 + image.center[1]
    c, s = np.cos(phi)[:, None, None], np.sin(phi)[:, None, None]
    xs = xs0 * c + ys0 * s - r[:, None, None]
    ys = -xs0 * s + ys0 * c
    s0, s1, sy = s0[:, None, None], s1[:, None, None], sy[:, None, None]
    g0 = 1.0 / np.sqrt(2 * np.pi * s0 ** 2) / np.sqrt(2 * np.pi * sy ** 2) * np.exp(-xs ** 2 / 2 / s0 ** 2) \
        * np.exp(-ys ** 2 / 2 / sy ** 2)
    g1 = 1.0 / np.sqrt(2 * np.pi * s1 ** 2) / np.sqrt(2 * np.pi * sy ** 2) * np.exp(-xs ** 2 / 2 / s1 ** 2) \
        * np.exp(-ys ** 2 / 2 / sy ** 2)
    w = np.where(xs > 0, g1, g0) * inside
    w = w / np.sum(w, axis=(1, 2), keepdims=True)
    w = np.where(w > 1e-6, w, 0.0)
    w[r > np.max(fitted_x.ux)] = 0.0
    return w.reshape(nx * ny, k * k).astype(np.float32)


def z_matrix(image, fitted_z):
    """(nz, nz) float32 K_z: column i is slice i's normalised Gaussian with the
    width interpolated at |z_i| (no image centre offset, as the reference),
    entries above 1e-16 kept."""
    nz = int(image.shape[2])
    uz = float(image.unit_size[2])
    f = _interp(fitted_z.uz, fitted_z.sigz)
    i = np.arange(nz)
    z = (i + 0.5) * uz - image.size[2] / 2
    sig = f(np.abs(z))
    mesh = (i[None, :] - i[:, None]) * uz                 # [i, col] = (col - i) * uz
    g = 1.0 / np.sqrt(2 * np.pi * sig[:, None] ** 2) * np.exp(-mesh ** 2 / 2 / sig[:, None] ** 2)
    g = g / np.sum(g, axis=1, keepdims=True)
    g = np.where(g > 1e-16, g, 0.0)
    return g.T.astype(np.float32)                         # K_z[col, i]


def weights_to_coo(w, nx, ny, half=HALF_PATCH):
    """The patch table as the reference's sparse K_xy (target, source)."""
    from scipy.sparse import coo_matrix
    k = 2 * half + 1
    p, o = np.nonzero(w)
    a, b = o // k - half, o % k - half
    tgt = (p // ny + a) * ny + (p % ny + b)
    return coo_matrix((w[p, o], (tgt, p)), shape=(nx * ny, nx * ny))


def psf_xy(x, w, nx, ny, adjoint, d=None, eps=0.0, half=HALF_PATCH):
    """K_xy x (adjoint 0) or K_xy^T x (adjoint 1) of x (nx * ny, nz) on the GPU
    (srf_psf_xy); with d: d / (K_xy x + eps)."""
    nz = int(x.shape[1])
    out = torch.empty_like(x)
    with torch.cuda.device(x.device):
        check(lib().srf_psf_xy(x.data_ptr(), int(nx), int(ny), nz, w.data_ptr(), int(half), int(adjoint),
                               None if d is None else d.data_ptr(), float(eps), out.data_ptr(),
                               stream_handle(x.device)), "srf_psf_xy")
    return out


@nef_class
class Deconvolute:
    """Richardson-Lucy deconvolution with the fitted separable PSF
    (deconvolute.py:20-67).  make_kernel_xy / make_kernel_z build the
    kernels for an image grid; __call__ runs n_iter iterations on the GPU."""
    n_iter: int
    fitted_x: FittedX
    fitted_y: FittedY
    fitted_z: FittedZ
    kernel_xy: object = None     # patch table (nx * ny, 121) float32
    kernel_z: object = None      # (nz, nz) float32

    def make_kernel_xy(self, image):
        if self.kernel_xy is None:
            object.__setattr__(self, "kernel_xy", xy_weights(image, self.fitted_x, self.fitted_y))
        return self.kernel_xy

    def make_kernel_z(self, image):
        if self.kernel_z is None:
            object.__setattr__(self, "kernel_z", z_matrix(image, self.fitted_z))
        return self.kernel_z

    def run(self, data):
        """n_iter iterations on a (nx, ny, nz) GPU tensor; returns the estimate."""
        if self.kernel_xy is None or self.kernel_z is None:
            raise ValueError("Please do make kernel first")
        dev = data.device
        nx, ny, nz = (int(v) for v in data.shape)
        d = data.reshape(nx * ny, nz).contiguous()
        w = torch.as_tensor(self.kernel_xy, device=dev).contiguous()
        kz = torch.as_tensor(self.kernel_z, device=dev).contiguous()
        if w.shape[0] != nx * ny or kz.shape != (nz, nz):
            raise ValueError("kernels were built for another grid")
        kzt = kz.t().contiguous()
        x = torch.ones_like(d)
        for _ in range(int(self.n_iter)):
            c = torch.matmul(x, kzt)                           # x K_z^T
            c = psf_xy(c, w, nx, ny, 0, d=d, eps=1e-16)        # d / (K_xy c + 1e-16)
            c = psf_xy(c, w, nx, ny, 1)                        # K_xy^T c
            x.mul_(torch.matmul(c, kz))                        # x * (c K_z)
        return x.reshape(nx, ny, nz)

    def __call__(self, image: Image) -> Image:
        out = self.run(torch.as_tensor(np.asarray(image.data, np.float32), device=device()))
        return image.update(data=out.cpu().numpy())


@nef_class
class PsfCorrect:
    """Deconvolution after reconstruction (psf_correct.py:16-30).  The fitted
    tables come with `deconv` (fitting them from point-source files,
    pnt_img_path, is not on this path)."""
    pnt_img_path: str
    deconv: Deconvolute

    def __call__(self, image: Image, n_iter: int = 10) -> Image:
        if self.deconv is None:
            raise NotImplementedError("PsfCorrect needs a Deconvolute with fitted tables")
        dec = self.deconv.update(n_iter=n_iter)
        dec.make_kernel_xy(image)
        dec.make_kernel_z(image)
        object.__setattr__(self, "deconv", dec)
        return dec(image)


__all__ = ["FittedX", "FittedY", "FittedZ", "Deconvolute", "PsfCorrect", "xy_weights", "z_matrix",
           "weights_to_coo", "psf_xy"]
