"""The full-size configurations the other GPU tests sample from:

- one subset of the TOF OSEM workload at its real size (5e6 events,
  256x256x128 x 2 mm) with the plans' TOF window 22, the setting the
  benchmark runs: projection and ratio against the oracle on a 2e5-event
  subsample, the windowed backprojection of that subsample against the
  oracle's, and one EM sub-iteration of the whole subset with window 22
  against window 0 (every reference term);
- the batched workload at its full size (32 scans, events log-uniform in
  1e5..2e7, in-plane 96..256, 64..128 slices, 2..4 mm voxels; 8 subsets x 2
  iterations, no scatter): every scan of the batched HIP-graph run
  bit-identical to its per-scan run, and the largest scan's OSEM against the
  oracle on a subsample of 2e5 events per subset (1.6e6 events: with fewer,
  most of its ~1e6 voxels get no term in a subset, fall to 0 after the first
  sub-iteration, and the EM ratio then amplifies float rounding).

Bars: 1e-5 relative L2 for projections / backprojections, 1e-4 for EM
images, bitwise for batched vs per-scan."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def rel_l2(a, b, mask=None):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    if mask is not None:
        a, b = a[mask], b[mask]
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_full_subset_tof_window_22_against_oracle(orc):
    from srf_amd import ops, phantom
    shape, voxel = (256, 256, 128), 2.0
    act, mu, center, size = phantom.random_ellipsoid_phantom(shape, voxel, 9, device="cuda")
    s2 = phantom.tof_sigma2_from_fwhm(60.0)
    n = 5_000_000
    lors = phantom.sample_events(act, center, size, n, 424.0, 300.0, 10, tof_fwhm_mm=60.0, device="cuda")
    geom = list(center) + list(size)
    print(f"events {n}: window-22 and exact plans", flush=True)
    win = ops.Plan(lors, shape, geom, True, 3.66, s2, tof_window=22.0)
    exact = ops.Plan(lors, shape, geom, True, 3.66, s2)
    img_k = ops.to_kernel_layout(act.contiguous())
    y = torch.rand(n, device="cuda", generator=torch.Generator("cuda").manual_seed(3)) + 0.5
    q = win.project_ratio(img_k, y, 1, 1e-5)
    p = win.project(img_k)
    # the oracle on a 2e5-event subsample (every reference term, float64 sums)
    sub = torch.randperm(n, generator=torch.Generator().manual_seed(4))[:200_000].sort().values
    ls = lors[sub.cuda()].cpu().numpy()
    act_np = act.cpu().numpy()
    p_o = orc.project(act_np, ls, center, size, tof=(3.66, s2), f64=True)
    ys = y[sub.cuda()].cpu().numpy()
    q_o = orc.ratio(ys, p_o.astype(np.float32), "full_tof")
    e_p = rel_l2(p[sub.cuda()].cpu().numpy(), p_o)
    e_q = rel_l2(q[sub.cuda()].cpu().numpy(), q_o)
    print(f"subsample vs oracle: projection {e_p:.3g}, ratio {e_q:.3g}", flush=True)
    assert e_p <= 1e-5 and e_q <= 1e-5
    # the windowed backprojection of the subsample against the oracle's
    sl = lors[sub.cuda()].contiguous()
    wsub = ops.Plan(sl, shape, geom, True, 3.66, s2, tof_window=22.0)
    vals = torch.from_numpy(q_o).cuda()
    bp = ops.from_kernel_layout(wsub.backproject(vals, 0)).cpu().numpy()
    bp_o = orc.backproject(q_o, ls, shape, center, size, tof=(3.66, s2))
    e_b = rel_l2(bp, bp_o)
    print(f"subsample backprojection vs oracle: {e_b:.3g}", flush=True)
    assert e_b <= 1e-5
    wsub.close()
    # one EM sub-iteration of the whole subset: window 22 against window 0
    emap = (torch.rand_like(img_k) + 0.5).contiguous()
    xs = []
    for plan in (win, exact):
        x = torch.ones_like(img_k)
        qq = plan.project_ratio(x, y, 1, 1e-5)
        plan.backproject_update(qq, x, emap, 1e8)
        xs.append(x.cpu().numpy())
    e_x = rel_l2(xs[0], xs[1])
    print(f"one sub-iteration, window 22 vs every term: {e_x:.3g}", flush=True)
    assert e_x <= 1e-5
    win.close()
    exact.close()


def test_batched_workload_full_size(orc):
    from srf_amd.batch import BatchConfig, scan_permutation
    from srf_amd.functions import OsemEngine
    from srf_amd.data import TofConfig
    from srf_amd.phantom import tof_sigma2_from_fwhm
    from srf_amd.pipeline import FWHM, TOF_BIN, TOF_WINDOW, ZERO_GUARD, make_scans, reconstruct_batched, \
        reconstruct_plain
    scans = make_scans(n_scans=32, seed=5, device="cuda")
    n_ev = sum(int(s.lors.shape[0]) for s in scans)
    print(f"32 scans, {n_ev} events, largest {max(int(s.lors.shape[0]) for s in scans)}", flush=True)
    assert 5e7 < n_ev < 3e8
    cfg = BatchConfig(item_records=8192, persistent=0, workgroups_per_cu=1, order=1)
    imgs_b, t_b, _ = reconstruct_batched(scans, 8, 2, None, config=cfg, graph=True, scatter=None)
    torch.cuda.empty_cache()
    imgs_p, t_p = reconstruct_plain(scans, 8, 2, None, scatter=None, item_records=8192)
    torch.cuda.empty_cache()
    same = [bool(torch.equal(a, b)) for a, b in zip(imgs_b, imgs_p)]
    print(f"batched {t_b:.2f} s, per-scan {t_p:.2f} s; bit-identical {sum(same)}/{len(same)}", flush=True)
    assert all(same)
    # the largest scan's OSEM on a 2e5-event subsample against the oracle
    big = max(scans, key=lambda s: int(s.lors.shape[0]))
    ns = 8 * 200_000
    idx = scan_permutation(int(big.lors.shape[0]), 99, big.lors.device)[:ns].sort().values
    lors = big.lors[idx].contiguous()
    vals = big.meta["inv_att"][idx].contiguous()
    s2 = tof_sigma2_from_fwhm(FWHM)
    perm = scan_permutation(ns, 7, lors.device)
    eng = OsemEngine(lors, vals, big.emap, big.shape, big.geom, 8, perm, tof=TofConfig(TOF_BIN, s2), ratio_kind=0,
                     eps=1e-8, zero_guard=ZERO_GUARD, tof_window=TOF_WINDOW)
    eng.run(16)
    x = eng.image().cpu().numpy()
    emap = big.emap.cpu().numpy()
    want = orc.osem(vals.cpu().numpy(), lors.cpu().numpy(), emap, list(big.center), list(big.size), 16, 8,
                    perm.cpu().numpy(), tof=(TOF_BIN, s2), driver="osem", guard_emap=True)
    err = rel_l2(x, want, emap > 0)
    print(f"largest scan {tuple(big.shape)}, {ns}-event subsample, 2 iterations vs oracle: {err:.3g}", flush=True)
    assert err <= 1e-4
