"""Batched OSEM (srf_batch_*): per-scan equality with single-scan OsemEngine
runs, HIP-graph replays against eager steps, argument validation, and the
autotuned schedule's reproducibility."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

TOF_BIN, FWHM = 3.66, 60.0


def _scans(specs, seed0=0):
    from srf_amd import phantom, ops
    from srf_amd.batch import Scan
    from srf_amd.functions import uniform_cylinder_lors
    dev = torch.device("cuda", 0)
    out = []
    for j, (shape, vox, n) in enumerate(specs):
        act, mu, center, size = phantom.random_ellipsoid_phantom(shape, vox, seed0 + j,
                                                                 body_axes=(0.4 * shape[0] * vox,
                                                                            0.3 * shape[1] * vox,
                                                                            0.4 * shape[2] * vox))
        lors = phantom.sample_events(act, center, size, n, 424.0, 300.0, 100 + j, tof_fwhm_mm=FWHM, device=dev)
        u = uniform_cylinder_lors(200_000, 424.0, 300.0, 7 + j, dev)
        em = ops.backproject(u, None, list(shape), list(center) + list(size), False, 0.0, 0.0, 0)
        emap = ops.from_kernel_layout(em) / 200_000
        emap = emap + 1e-3 * emap.mean()          # strictly positive: no 0/0 in the update
        out.append(Scan(lors, shape, center, size, emap, seed=j))
    return out


SPECS = [((48, 48, 24), 4.0, 60_000), ((64, 64, 32), 3.0, 150_000), ((40, 40, 40), 3.5, 20_000),
         ((96, 96, 16), 2.5, 100_000), ((32, 32, 64), 4.0, 5_000)]


@pytest.mark.parametrize("cfg", [dict(), dict(persistent=1, workgroups_per_cu=2, order=0),
                                 dict(item_records=2048, persistent=1, workgroups_per_cu=1, order=1)])
def test_batch_matches_single_scan_runs(cfg):
    from srf_amd.batch import BatchConfig, BatchOsem, osem_single
    from srf_amd.phantom import tof_sigma2_from_fwhm
    print("batch cfg", cfg, flush=True)
    s2 = tof_sigma2_from_fwhm(FWHM)
    scans = _scans(SPECS)
    conf = BatchConfig(**cfg)
    b = BatchOsem(scans, 3, TOF_BIN, s2, True, config=conf)
    print(b.info(0), flush=True)
    b.run(2)
    torch.cuda.synchronize()
    for i, sc in enumerate(scans):
        eng = osem_single(sc, 3, 2, TOF_BIN, s2, True, item_records=conf.item_records)
        want = eng.image()
        got = b.image(i)
        err = float((got - want).norm() / want.norm())
        print(f"scan {i} {sc.shape}: rel L2 {err:.2e} bitwise {bool(torch.equal(got, want))}", flush=True)
        assert err <= 1e-4
        assert torch.equal(got, want)
    b.close()


def test_graph_replays_match_eager_steps():
    from srf_amd.batch import BatchConfig, BatchOsem
    from srf_amd.phantom import tof_sigma2_from_fwhm
    s2 = tof_sigma2_from_fwhm(FWHM)
    scans = _scans(SPECS[:3], seed0=20)
    conf = BatchConfig(persistent=1, workgroups_per_cu=1)
    eager = BatchOsem(scans, 2, TOF_BIN, s2, True, config=conf)
    graph = BatchOsem(scans, 2, TOF_BIN, s2, True, config=conf)
    graph.capture()
    for r in range(3):
        eager.iteration()
        graph.graph.replay()
        junk = torch.randn(1000, device="cuda").sum()      # unrelated eager work between replays
        torch.cuda.synchronize()
        for i in range(len(scans)):
            assert torch.equal(graph.image(i), eager.image(i)), (r, i)
        assert torch.isfinite(junk)
    # values changed in place are seen by the captured graph
    new_vals = torch.full((scans[0].lors.shape[0],), 2.0, device="cuda")
    eager.set_values(0, new_vals)
    graph.set_values(0, new_vals)
    eager.iteration()
    graph.graph.replay()
    torch.cuda.synchronize()
    assert torch.equal(graph.image(0), eager.image(0))
    eager.close()
    graph.close()


def test_batch_create_validation():
    from srf_amd import ops
    from srf_amd._lib import SrfBatchCfg, SrfBatchScan, lib
    dev = torch.device("cuda", 0)
    lors = torch.tensor([[-300.0, 5.0, 1.0, 300.0, -7.0, 2.0, 10.0]], device=dev).repeat(100, 1)
    p_tof = ops.Plan(lors, [32, 32, 16], [0, 0, 0, 128, 128, 64], True, 3.66, 600.0)
    p_non = ops.Plan(lors[:, :6].contiguous(), [32, 32, 16], [0, 0, 0, 128, 128, 64], False)
    x = torch.ones(16 * 32 * 32, device=dev)
    q = torch.empty(100, device=dev)
    y = torch.ones(100, device=dev)
    S = (SrfBatchScan * 2)(SrfBatchScan(x.data_ptr(), x.data_ptr(), y.data_ptr(), q.data_ptr()),
                           SrfBatchScan(x.data_ptr(), x.data_ptr(), y.data_ptr(), q.data_ptr()))
    h = ctypes.c_void_p()
    cfg = SrfBatchCfg(0, 1, 0)
    P = (ctypes.c_void_p * 2)(p_tof.handle, p_non.handle)
    # TOF and non-TOF plans share a batch (one launch per kernel variant)
    assert lib().srf_batch_create(P, S, 2, ctypes.byref(cfg), ctypes.byref(h), None) == 0
    assert lib().srf_batch_destroy(h) == 0
    h = ctypes.c_void_p()
    assert lib().srf_batch_create(P, S, 0, ctypes.byref(cfg), ctypes.byref(h), None) == -1
    bad = SrfBatchCfg(2, 1, 0)
    P1 = (ctypes.c_void_p * 1)(p_tof.handle)
    assert lib().srf_batch_create(P1, S, 1, ctypes.byref(bad), ctypes.byref(h), None) == -1
    S0 = (SrfBatchScan * 1)(SrfBatchScan(x.data_ptr(), x.data_ptr(), None, q.data_ptr()))
    assert lib().srf_batch_create(P1, S0, 1, ctypes.byref(cfg), ctypes.byref(h), None) == -1
    assert not h.value


def test_autotuned_schedule_reproducible(tmp_path):
    from srf_amd.autotune import Autotuner
    from srf_amd.batch import BatchConfig, BatchOsem
    from srf_amd.phantom import tof_sigma2_from_fwhm
    s2 = tof_sigma2_from_fwhm(FWHM)
    scans = _scans(SPECS, seed0=40)
    cache = str(tmp_path / "tune.json")
    tuner = Autotuner(cache, reps=2)
    best = tuner.tune(scans, 2, TOF_BIN, s2, True, log=lambda m: print(m, flush=True))
    print("chosen", best, flush=True)
    # a second tuner reads the cache and returns the same schedule without timing
    again = Autotuner(cache).tune(scans, 2, TOF_BIN, s2, True, timer=lambda c: 1 / 0)
    assert again == best
    # the tuned schedule reproduces itself bit for bit and agrees with the default schedule
    imgs = []
    for conf in (best, best, BatchConfig()):
        b = BatchOsem(scans, 2, TOF_BIN, s2, True, config=conf)
        b.run(2, graph=True)
        torch.cuda.synchronize()
        imgs.append([b.image(i).clone() for i in range(len(scans))])
        b.close()
    for i in range(len(scans)):
        assert torch.equal(imgs[0][i], imgs[1][i])
        err = float((imgs[0][i] - imgs[2][i]).norm() / imgs[2][i].norm())
        assert err <= 1e-5, (i, err)
