"""Failure handling on the GPU, with every failure forced from the host (no
kernel faults or hangs):

- OSEM checkpoints: a run interrupted after some sub-iterations and resumed
  from its checkpoint by a new engine ends bit-identical to an uninterrupted
  run (single scan, OsemEngine; many scans, BatchOsem with changed values);
- a HIP-graph capture that fails falls back to eager launches with the same
  result;
- a failed srf_plan_rebuild leaves an empty plan: backprojections write zeros,
  the EM update gives x * 0 / emap, projections write nothing, a batch refuses
  the plan, and a later rebuild works again.
"""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

FWHM, TOF_BIN = 60.0, 3.66


@pytest.fixture(scope="module")
def scan_data():
    from srf_amd import phantom
    act, mu, center, size = phantom.random_ellipsoid_phantom((48, 48, 24), 4.0, 7, device="cuda")
    lors = phantom.sample_events(act, center, size, 120000, 424.0, 300.0, 8, tof_fwhm_mm=FWHM, device="cuda")
    emap = torch.ones_like(act)
    return lors, emap, (48, 48, 24), list(center), list(size)


def _engine(scan_data):
    from srf_amd.data import TofConfig
    from srf_amd.functions import OsemEngine, subset_permutation
    from srf_amd.phantom import tof_sigma2_from_fwhm
    lors, emap, shape, center, size = scan_data
    n = lors.shape[0]
    return OsemEngine(lors, torch.ones(n, device="cuda"), emap, shape, center + size, 4,
                      subset_permutation(n, 3), tof=TofConfig(TOF_BIN, tof_sigma2_from_fwhm(FWHM)), ratio_kind=1,
                      eps=1e-5, zero_guard=1e8, tof_window=22.0)


def test_osem_resume_equals_uninterrupted(scan_data, tmp_path):
    path = str(tmp_path / "osem.ckpt")
    ref = _engine(scan_data)
    ref.run(12)
    a = _engine(scan_data)
    a.run(7, checkpoint=path, every=3)          # written at 3, 6 and 7; then "interrupted"
    del a
    b = _engine(scan_data)
    b.run_to(12, checkpoint=path)                # resumes at sub-iteration 7
    torch.cuda.synchronize()
    print(f"resumed from 7 of 12: max |diff| {float((b.x - ref.x).abs().max()):.3g}", flush=True)
    assert b.done == 12 and torch.equal(b.x, ref.x)
    # the reference subset order is another workload: its checkpoint key differs
    c = _engine(scan_data)
    with pytest.raises(ValueError, match="another workload"):
        c.run(1, reference_subset_order=True, checkpoint=path)


def _batch_scans(n_scans=3, seed0=40):
    from srf_amd import phantom
    from srf_amd.batch import Scan
    scans = []
    for j, (nxy, nz, n) in enumerate([(48, 24, 60000), (64, 16, 90000), (32, 32, 40000)][:n_scans]):
        act, mu, center, size = phantom.random_ellipsoid_phantom((nxy, nxy, nz), 4.0, seed0 + j, device="cuda")
        lors = phantom.sample_events(act, center, size, n, 424.0, 300.0, seed0 + 10 + j, tof_fwhm_mm=FWHM,
                                     device="cuda")
        scans.append(Scan(lors, (nxy, nxy, nz), tuple(center), tuple(size), torch.ones_like(act), seed=j))
    return scans


def _batch(scans):
    from srf_amd.batch import BatchConfig, BatchOsem
    from srf_amd.phantom import tof_sigma2_from_fwhm
    return BatchOsem(scans, 2, TOF_BIN, tof_sigma2_from_fwhm(FWHM), True, config=BatchConfig(), zero_guard=1e8,
                     tof_window=22.0)


def test_batch_resume_with_new_values_equals_uninterrupted(tmp_path):
    scans = _batch_scans()
    vals = torch.linspace(0.5, 1.5, scans[1].lors.shape[0], device="cuda")
    path = str(tmp_path / "batch.ckpt")
    ref = _batch(scans)
    ref.run(1, graph=True)
    ref.set_values(1, vals)
    ref.run(2, graph=True)
    a = _batch(scans)
    a.run(1, graph=True, checkpoint=path)
    a.set_values(1, vals)
    a.run(1, graph=True, checkpoint=path)        # 2 of 3 iterations, values in the checkpoint
    a.close()
    b = _batch(scans)
    b.run_to(3, graph=True, checkpoint=path)
    torch.cuda.synchronize()
    for i in range(len(scans)):
        assert torch.equal(b.x[i], ref.x[i]), i
    assert b.done == 3
    ref.close()
    b.close()


def test_capture_failure_runs_eager_with_same_result():
    scans = _batch_scans(2, seed0=60)
    eager = _batch(scans)
    forced = _batch(scans)

    def refuse():
        forced.graph_error = "forced from the host"
        raise RuntimeError("forced from the host")

    forced.capture = refuse
    eager.run(2, graph=False)
    with pytest.warns(UserWarning, match="capture failed"):
        forced.run(2, graph=True)
    torch.cuda.synchronize()
    for i in range(len(scans)):
        assert torch.equal(forced.x[i], eager.x[i])
    eager.close()
    forced.close()


def test_failed_rebuild_leaves_an_empty_plan(scan_data):
    from srf_amd import ops
    from srf_amd._lib import SrfBatchCfg, SrfBatchScan, SrfPlanOpts, lib, stream_handle
    from srf_amd.phantom import tof_sigma2_from_fwhm
    lors, emap, shape, center, size = scan_data
    s2 = tof_sigma2_from_fwhm(FWHM)
    geom = center + size
    nx, ny, nz = shape
    plan = ops.Plan(lors, list(shape), geom, True, TOF_BIN, s2)
    q = torch.rand(lors.shape[0], device="cuda") + 0.5
    bp0 = plan.backproject(q)
    assert float(bp0.abs().sum()) > 0
    # a rebuild with an invalid work-item size fails on the host
    opts = SrfPlanOpts(100, 0.0)
    grid = ops._grid(list(shape), geom)
    tof = ops._tof(True, TOF_BIN, s2)
    rc = lib().srf_plan_rebuild(plan.handle, lors.data_ptr(), lors.shape[0], 7, lors.data_ptr() + 24, 7, grid, tof,
                                ctypes.byref(opts), stream_handle(lors.device))
    assert rc == -1
    dev = lors.device
    img = torch.full((nz, ny, nx), 7.0, device=dev)
    assert lib().srf_plan_backproject(plan.handle, q.data_ptr(), 0, 0, img.data_ptr(), stream_handle(dev)) == 0
    acc = torch.full((nz, ny, nx), 7.0, device=dev)
    assert lib().srf_plan_backproject(plan.handle, q.data_ptr(), 0, 1, acc.data_ptr(), stream_handle(dev)) == 0
    x = torch.full((nz, ny, nx), 3.0, device=dev)
    e = torch.full((nz, ny, nx), 2.0, device=dev)
    e[0] = 0.0
    assert lib().srf_plan_backproject_update(plan.handle, q.data_ptr(), x.data_ptr(), e.data_ptr(), 0.0,
                                             stream_handle(dev)) == 0
    out = torch.full((lors.shape[0],), 5.0, device=dev)
    assert lib().srf_plan_project(plan.handle, ops.to_kernel_layout(torch.ones(shape, device=dev)).data_ptr(),
                                  out.data_ptr(), stream_handle(dev)) == 0
    torch.cuda.synchronize()
    assert torch.count_nonzero(img) == 0                     # written zeros
    assert torch.all(acc == 7.0)                             # nothing added
    assert torch.all(x[1:] == 0.0) and torch.isnan(x[0]).all()   # x * 0 / emap (0/0 where emap is 0)
    assert torch.all(out == 5.0)                             # an empty list projects nothing
    # a batch refuses the plan until a rebuild succeeds
    xs = torch.ones((nz, ny, nx), device=dev)
    P = (ctypes.c_void_p * 1)(plan.handle)
    S = (SrfBatchScan * 1)(SrfBatchScan(xs.data_ptr(), e.data_ptr(), q.data_ptr(), out.data_ptr()))
    h = ctypes.c_void_p()
    cfg = SrfBatchCfg(0, 1, 1)
    assert lib().srf_batch_create(P, S, 1, ctypes.byref(cfg), ctypes.byref(h), stream_handle(dev)) == -1
    plan.rebuild(lors)
    bp1 = plan.backproject(q)
    torch.cuda.synchronize()
    assert torch.equal(bp1, bp0)
    plan.close()
