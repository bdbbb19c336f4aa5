"""PSF deconvolution on the GPU (srf_psf_xy stencil + GEMMs, srf_amd.psf)
against the CPU oracle (oracle/psf_oracle.py, SciPy sparse, float64)."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from test_psf_cpu import fitted, image  # noqa: E402


def rel_l2(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_psf_xy_both_directions():
    import torch
    import psf_oracle as po
    from srf_amd import psf
    img = image((40, 36, 6))
    fx, fy, _ = fitted()
    w = psf.xy_weights(img, fx, fy)
    k = po.xy_kernel((40, 36), img.center, img.size, fx.ux, fx.sigx0, fx.sigx1, fy.sigy).tocsr().astype(np.float64)
    x = np.random.default_rng(1).random((40 * 36, 6)).astype(np.float32) + 0.1
    X = torch.from_numpy(x).cuda()
    W = torch.from_numpy(w).cuda()
    fwd = psf.psf_xy(X, W, 40, 36, 0).cpu().numpy()
    adj = psf.psf_xy(X, W, 40, 36, 1).cpu().numpy()
    assert rel_l2(fwd, k @ x.astype(np.float64)) <= 1e-6
    assert rel_l2(adj, k.T @ x.astype(np.float64)) <= 1e-6
    d = np.random.default_rng(2).random((40 * 36, 6)).astype(np.float32)
    rat = psf.psf_xy(X, W, 40, 36, 0, d=torch.from_numpy(d).cuda(), eps=1e-16).cpu().numpy()
    assert rel_l2(rat, d / (k @ x.astype(np.float64) + 1e-16)) <= 1e-6


def test_deconvolute_matches_oracle():
    import psf_oracle as po
    from srf_amd import psf
    from srf_amd.data import Image
    shape = (48, 40, 12)
    rng = np.random.default_rng(3)
    truth = np.zeros(shape, np.float32)
    truth[10:30, 8:28, 3:9] = 1.0
    truth[30:36, 20:26, 5:8] = 4.0
    img = image(shape)
    fx, fy, fz = fitted()
    kxy = po.xy_kernel(shape[:2], img.center, img.size, fx.ux, fx.sigx0, fx.sigx1, fy.sigy)
    kz = po.z_kernel(shape[2], img.size[2], fz.uz, fz.sigz)
    blurred = po.forward(truth.reshape(-1, shape[2]).astype(np.float64), kxy, kz.todense())
    d = (blurred + 0.01 * rng.random(blurred.shape)).astype(np.float32)
    ref = po.deconvolute(d, kxy, kz, 8).reshape(shape)
    dec = psf.Deconvolute(8, fx, fy, fz)
    dimg = Image(d.reshape(shape), img.center, img.size)
    dec.make_kernel_xy(dimg)
    dec.make_kernel_z(dimg)
    out = dec(dimg).data
    err = rel_l2(out, ref)
    print(f"GPU vs oracle after 8 iterations: rel L2 {err:.2e}", flush=True)
    assert err <= 1e-5
    # the deconvolution sharpens: closer to the truth than the blurred data
    assert rel_l2(out, truth) < rel_l2(d.reshape(shape), truth)
    # PsfCorrect runs the same iterations
    pc = psf.PsfCorrect(None, psf.Deconvolute(8, fx, fy, fz))
    assert np.array_equal(pc(dimg, n_iter=8).data, out)


def test_deconvolute_timing_mct_grid():
    """10 iterations on the mCT image grid (200^3 x 2.05 mm)."""
    import torch
    from srf_amd import psf
    from srf_amd.data import Image
    shape, size = (200, 200, 200), [410.0, 410.0, 410.0]
    fx, fy, fz = fitted()
    img = Image(np.zeros(shape, np.float32), [0.0, 0.0, 0.0], size)
    dec = psf.Deconvolute(10, fx, fy, fz)
    dec.make_kernel_xy(img)
    dec.make_kernel_z(img)
    d = torch.rand(shape, device="cuda") + 0.5
    dec.run(d)
    torch.cuda.synchronize()
    t = time.perf_counter()
    x = dec.run(d)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t
    print(f"PSF deconvolution 200^3: {1e3 * dt / 10:.2f} ms per iteration", flush=True)
    assert torch.isfinite(x).all()
