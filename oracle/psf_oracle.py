"""CPU oracle of srfnef's PSF deconvolution -- test infrastructure only
(the product path is srf_amd.psf on the GPU).

Restates corrections/psf/deconvolute.py with SciPy sparse matrices in float64:
- xy_kernel: xy_main (:163-213) pixel by pixel, as the reference loops, with
  locate_kernel_xy (:90-103) and scipy interp1d with extrapolation;
- z_kernel: z_main (:215-235) and locate_kernel_z (:237-239);
- deconvolute: the iterations of __call__ (:49-62),
  x <- x * K_xy^T (d / (K_xy x K_z^T + 1e-16)) K_z from x = 1.
Kernel entries are stored as float32 (the reference's lil_matrix dtype).

Pinned in tests/test_psf_cpu.py by closed forms (an isotropic kernel equals the
sampled 2-D Gaussian; columns sum to 1), the fixed point x = 1 of the
iteration for d = A(1), and agreement with srf_amd.psf's vectorised table.
"""
import numpy as np
from scipy import interpolate, sparse


def _f(x, y):
    return interpolate.interp1d(x, y, fill_value="extrapolate")


def kernel_xy_patch(x, y, sigx0, sigx1, sigy):
    ans0 = 1.0 / np.sqrt(2 * np.pi * sigx0 ** 2) / np.sqrt(2 * np.pi * sigy ** 2) * np.exp(-x ** 2 / 2 / sigx0 ** 2) \
        * np.exp(-y ** 2 / 2 / sigy ** 2)
    ans0[x > 0] = 0.0
    ans1 = 1.0 / np.sqrt(2 * np.pi * sigx1 ** 2) / np.sqrt(2 * np.pi * sigy ** 2) * np.exp(-x ** 2 / 2 / sigx1 ** 2) \
        * np.exp(-y ** 2 / 2 / sigy ** 2)
    ans1[x <= 0] = 0.0
    ans = ans0 + ans1
    return ans / np.sum(ans)


def xy_kernel(shape, center, size, ux, sigx0, sigx1, sigy, half=5):
    """Sparse K_xy (target, source) of an (nx, ny) grid."""
    nx, ny = shape[0], shape[1]
    dx, dy = size[0] / nx, size[1] / ny
    x = (np.arange(nx) + 0.5) * dx - size[0] / 2 + center[0]
    y = (np.arange(ny) + 0.5) * dy - size[1] / 2 + center[1]
    xm, ym = np.meshgrid(x, y, indexing="ij")
    rm, pm = np.sqrt(xm ** 2 + ym ** 2), np.arctan2(ym, xm)
    f0, f1, fy = _f(ux, sigx0), _f(ux, sigx1), _f(ux, sigy)
    out = sparse.lil_matrix((nx * ny, nx * ny), dtype=np.float32)
    for i in range(nx):
        for j in range(ny):
            irow = i * ny + j
            r, phi = rm[i, j], pm[i, j]
            if r > np.max(ux):
                continue
            px = np.arange(max(i - half, 0), min(i + half + 1, nx))
            py = np.arange(max(j - half, 0), min(j + half + 1, ny))
            xs0 = (px + 0.5) * dx - size[0] / 2 + center[0]
            ys0 = (py + 0.5) * dy - size[1] / 2 + center[1]
            xs0, ys0 = np.meshgrid(xs0, ys0, indexing="ij")
            xs = xs0 * np.cos(phi) + ys0 * np.sin(phi)
            ys = -xs0 * np.sin(phi) + ys0 * np.cos(phi)
            wei = kernel_xy_patch(xs - r, ys, f0(r), f1(r), fy(r))
            ipx, ipy = np.meshgrid(px, py, indexing="ij")
            keep = wei > 1e-6
            col = ipx[keep] * ny + ipy[keep]
            out[col, [irow] * int(keep.sum())] = wei[keep]
    return out.tocoo()


def z_kernel(nz, size_z, uz, sigz):
    dz = size_z / nz
    fz = _f(uz, sigz)
    out = sparse.lil_matrix((nz, nz), dtype=np.float32)
    for i in range(nz):
        z = (i + 0.5) * dz - size_z / 2
        mesh = (np.arange(nz) - i) * dz
        sig = fz(abs(z))
        k = 1.0 / np.sqrt(2 * np.pi * sig ** 2) * np.exp(-mesh ** 2 / 2 / sig ** 2)
        k = k / np.sum(k)
        keep = np.nonzero(k > 1e-16)[0]
        if keep.size:
            out[keep, [i] * keep.size] = k[keep]
    return out.tocoo()


def forward(x, kxy, kz):
    """A(x) = K_xy x K_z^T on (nx ny, nz) arrays."""
    kz = np.asarray(kz.todense() if sparse.issparse(kz) else kz, np.float64)
    return np.asarray(sparse.csr_matrix(kxy, dtype=np.float64) @ (np.asarray(x, np.float64) @ kz.T))


def deconvolute(d, kxy, kz, n_iter):
    """n_iter Richardson-Lucy iterations from x = 1 (float64)."""
    d = np.asarray(d, np.float64)
    kxy = sparse.csr_matrix(kxy, dtype=np.float64)
    kz = np.asarray(kz.todense() if sparse.issparse(kz) else kz, np.float64)
    x = np.ones_like(d)
    for _ in range(n_iter):
        c = kxy @ (x @ kz.T) + 1e-16
        c = d / c
        c = kxy.T @ c
        x = x * (c @ kz)
    return x
