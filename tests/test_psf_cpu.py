"""PSF deconvolution kernels on the CPU: the oracle (oracle/psf_oracle.py)
pinned by closed forms and the iteration's fixed point, and the product's
vectorised kernel tables (srf_amd.psf) against the oracle's pixel loop."""
import numpy as np
import pytest

from srf_amd import psf
from srf_amd.data import Image

import psf_oracle as po

UX = np.array([0.0, 50.0, 100.0, 200.0, 300.0])
SX0 = np.array([1.5, 1.6, 1.8, 2.2, 2.8]) * 2.0
SX1 = np.array([1.5, 1.7, 2.0, 2.6, 3.3]) * 2.0
SY = np.array([1.5, 1.5, 1.6, 1.7, 1.9]) * 2.0
UZ = np.array([0.0, 20.0, 40.0])
SZ = np.array([1.8, 2.0, 2.4]) * 2.0


def fitted(sx0=SX0, sx1=SX1, sy=SY, sz=SZ):
    fx = psf.FittedX(np.stack([np.ones_like(UX), sx0, sx1, UX], 1))
    fy = psf.FittedY(np.stack([np.ones_like(UX), sy, UX], 1))
    fz = psf.FittedZ(np.stack([np.ones_like(UZ), sz, UZ], 1))
    return fx, fy, fz


def image(shape=(24, 20, 8), voxel=4.0, data=None):
    size = [shape[k] * voxel for k in range(3)]
    return Image(np.zeros(shape, np.float32) if data is None else data, [1.0, -2.0, 0.0], size)


def test_isotropic_column_is_sampled_gaussian():
    s = 5.0
    img = image()
    k = po.xy_kernel((24, 20), img.center, img.size, UX, np.full(5, s), np.full(5, s), np.full(5, s)).tocsc()
    i, j = 12, 10
    col = k[:, i * 20 + j].toarray().reshape(24, 20)
    a, b = np.meshgrid(np.arange(24) - i, np.arange(20) - j, indexing="ij")
    g = np.exp(-((a * 4.0) ** 2 + (b * 4.0) ** 2) / 2 / s ** 2) * (np.abs(a) <= 5) * (np.abs(b) <= 5)
    g = g / g.sum()
    g[g <= 1e-6] = 0
    np.testing.assert_allclose(col, g, rtol=1e-6, atol=1e-9)


def test_columns_sum_to_one_and_fixed_point():
    img = image()
    fx, fy, fz = fitted()
    kxy = po.xy_kernel((24, 20), img.center, img.size, fx.ux, fx.sigx0, fx.sigx1, fy.sigy)
    kz = po.z_kernel(8, img.size[2], fz.uz, fz.sigz)
    cs = np.asarray(kxy.tocsc().sum(axis=0)).ravel()
    assert np.all(np.abs(cs - 1) < 200e-6)              # at most 121 entries cut below 1e-6
    assert np.allclose(np.asarray(kz.tocsc().sum(axis=0)).ravel(), 1, atol=1e-6)
    # d = A(1): x = 1 stays (up to the cut mass)
    d = po.forward(np.ones((480, 8)), kxy, kz.todense())
    x = po.deconvolute(d, kxy, kz, 5)
    assert np.max(np.abs(x - 1)) < 2e-3
    # the asymmetric radial widths make the blur of a point off-centre lean outwards
    e = np.zeros((480, 8))
    e[(20) * 20 + 10, 4] = 1.0                           # pixel far from the centre along +x
    out = po.forward(e, kxy, kz.todense())[:, 4].reshape(24, 20)
    assert out[21:, 10].sum() > out[:20, 10].sum()


def test_vectorised_tables_equal_oracle():
    img = image()
    fx, fy, fz = fitted()
    w = psf.xy_weights(img, fx, fy)
    coo = psf.weights_to_coo(w, 24, 20).tocsr()
    ref = po.xy_kernel((24, 20), img.center, img.size, fx.ux, fx.sigx0, fx.sigx1, fy.sigy).tocsr()
    diff = abs(coo - ref)
    assert diff.max() <= 1e-7 * ref.max()
    assert coo.nnz == ref.nnz
    kz = psf.z_matrix(img, fz)
    np.testing.assert_allclose(kz, np.asarray(po.z_kernel(8, img.size[2], fz.uz, fz.sigz).todense()), rtol=1e-6,
                               atol=0)


def test_pixels_beyond_the_fitted_radius_have_no_kernel():
    img = image((40, 40, 4), voxel=4.0)
    fx, fy, _ = fitted()
    fx = psf.FittedX(fx.data[:3])                        # fitted out to 100 mm only
    fy = psf.FittedY(fy.data[:3])                        # (sigy is read against fitted_x.ux)
    w = psf.xy_weights(img, fx, fy)
    x = (np.arange(40) + 0.5) * 4 - 80 + 1.0
    y = (np.arange(40) + 0.5) * 4 - 80 - 2.0
    r = np.hypot(*np.meshgrid(x, y, indexing="ij")).ravel()
    assert np.all(w[r > 100].sum(1) == 0) and np.all(w[r <= 100].sum(1) > 0.99)
