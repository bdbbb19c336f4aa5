"""End-to-end OSEM with scatter inside the loop: the batched method (one
BatchOsem, HIP graph, surrogate with SSS check) against the plain per-scan
method, on a small batch of uneven scans."""
import math

import pytest
import torch

pytestmark = pytest.mark.gpu


def _small_scans():
    from srf_amd.pipeline import make_scans
    return make_scans(n_scans=3, seed=9, ev_range=(2e4, 1.2e5), sizes=(48, 64), slices=(16, 32), voxel=(4.0, 6.0),
                      n_emap=200_000, log=lambda m: print(m, flush=True))


def _estimator(tol):
    from srf_amd.pipeline import ScatterEstimator
    from srf_amd.surrogate import UNet
    net = UNet(seed=3, device="cuda")          # untrained: its estimate is far from SSS
    return ScatterEstimator(net=net, kappa=1.0, check_rings=(16,), tol=tol)


def test_untrained_surrogate_falls_back_to_sss_bit_for_bit():
    from srf_amd.batch import BatchConfig
    from srf_amd.pipeline import reconstruct_batched, reconstruct_plain
    scans = _small_scans()
    est = _estimator(0.02)
    imgs_b, t_b, info = reconstruct_batched(scans, n_sub=2, n_iter=3, estimator=est,
                                            config=BatchConfig(persistent=1), graph=True)
    print("checks", est.checks, "fallback", est.fallback, info, flush=True)
    # each scan's first estimate is checked, fails and switches that scan to SSS
    assert len(est.checks) == len(scans) and all(e > 0.02 for _, _, e in est.checks)
    assert sorted(i for i, _, _ in est.checks) == list(range(len(scans)))
    assert all(est.fallback.get(i) for i in range(len(scans)))
    imgs_p, t_p = reconstruct_plain(scans, n_sub=2, n_iter=3, estimator=_estimator(0.02), scatter="sss")
    for a, b in zip(imgs_b, imgs_p):
        assert torch.equal(a, b)


def test_batched_surrogate_loop_matches_per_scan_loop():
    from srf_amd.batch import BatchConfig
    from srf_amd.pipeline import reconstruct_batched, reconstruct_plain
    scans = _small_scans()
    est_b = _estimator(math.inf)
    imgs_b, _, _ = reconstruct_batched(scans, n_sub=2, n_iter=3, estimator=est_b, config=BatchConfig(), graph=True)
    assert not any(est_b.fallback.values())
    # every in-loop estimate (iterations 2 and 3 of every scan) is checked
    assert sorted((i, k) for i, k, _ in est_b.checks) == [(i, k) for i in range(len(scans)) for k in range(2)]
    est_p = _estimator(math.inf)
    imgs_p, _ = reconstruct_plain(scans, n_sub=2, n_iter=3, estimator=est_p, scatter="surrogate")
    for a, b in zip(imgs_b, imgs_p):
        err = float((a - b).norm() / b.norm())
        assert err <= 1e-4
        assert torch.equal(a, b)
    # an SSS estimate changes the values: the reconstruction differs from one without scatter
    # (the untrained network's sinogram may sum to <= 0, which subtracts nothing)
    imgs_s, _, _ = reconstruct_batched(scans, n_sub=2, n_iter=3, estimator=est_b, scatter="sss")
    imgs_n, _, _ = reconstruct_batched(scans, n_sub=2, n_iter=3, estimator=est_b, scatter=None)
    assert all(not torch.equal(a, b) for a, b in zip(imgs_s, imgs_n))


def test_scatter_correct_kernel_matches_formula():
    from srf_amd.sss import scatter_correct
    g = torch.Generator(device="cpu").manual_seed(1)
    n, nb = 10_000, 500
    sino = torch.rand(nb, generator=g).cuda()
    bins = torch.randint(-3, nb + 3, (n,), generator=g, dtype=torch.int32).cuda()
    y = torch.rand(n, generator=g).cuda() + 0.5
    w = torch.rand(n, generator=g).cuda() + 1.0
    out = scatter_correct(y, w, bins, sino, 0.3)
    b = bins.long()
    ok = (b >= 0) & (b < nb)
    s = torch.where(ok, sino.double()[b.clamp(0, nb - 1)], torch.zeros((), dtype=torch.float64, device="cuda"))
    want = (y.double() - s / s.sum() * y.double().sum() * 0.3) * w.double()
    assert float((out.double() - want).norm() / want.norm()) <= 1e-6
    # S = 0 leaves the values (times the weight) unchanged
    out0 = scatter_correct(y, w, bins, torch.zeros(nb, device="cuda"), 0.3)
    assert torch.equal(out0, y * w)


def test_smooth_matches_separable_gaussian():
    """srf_smooth against a float64 separable Gaussian with edge-renormalised
    weights; a constant stays constant; the mask is applied."""
    import numpy as np
    from srf_amd._lib import check, lib, stream_handle
    rng = np.random.default_rng(2)
    x = rng.random((20, 17, 23)).astype(np.float32)
    sig = (1.3, 2.1, 0.8)

    def ref(a):
        a = a.astype(np.float64)
        for ax in (2, 1, 0):
            s = sig[ax]
            r = int(np.ceil(3 * s))
            n = a.shape[ax]
            out = np.zeros_like(a)
            for p in range(n):
                ks = np.arange(max(-r, -p), min(r, n - 1 - p) + 1)
                w = np.exp(-0.5 * ks ** 2 / s ** 2)
                sl = np.take(a, p + ks, axis=ax)
                out_p = np.tensordot(w, np.moveaxis(sl, ax, 0), axes=1) / w.sum()
                idx = [slice(None)] * 3
                idx[ax] = p
                out[tuple(idx)] = out_p
            a = out
        return a

    def run(a, mask=None):
        t = torch.from_numpy(a).cuda()
        tmp, out = torch.empty_like(t), torch.empty_like(t)
        m = torch.from_numpy(mask).cuda() if mask is not None else None
        check(lib().srf_smooth(t.data_ptr(), *a.shape, *sig, m.data_ptr() if m is not None else None, tmp.data_ptr(),
                               out.data_ptr(), stream_handle()), "srf_smooth")
        return out.cpu().numpy()

    got = run(x)
    want = ref(x)
    assert np.abs(got - want).max() <= 1e-5 * np.abs(want).max()
    c = np.full((20, 17, 23), 3.5, np.float32)
    assert np.allclose(run(c), 3.5, rtol=1e-6)
    mask = (rng.random(x.shape) > 0.5).astype(np.float32)
    assert np.array_equal(run(x, mask) == 0, mask == 0)
