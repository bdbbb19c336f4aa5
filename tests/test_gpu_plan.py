"""GPU parity of the LDS-tiled plan kernels (srf_plan_*): the term digest of
the tiled walk equals the oracle's bit for bit (every term in exactly one tile
segment), and projections / backprojections match the golden fixtures and the
one-event-per-lane kernels within 1e-5 relative L2."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GRIDS = {
    "even": ((32, 28, 16), [0.0, 0.0, 0.0], [64.0, 56.0, 32.0]),
    "odd": ((31, 27, 15), [1.25, -0.75, 0.5], [62.0, 54.0, 30.0]),
    "aniso_z": ((24, 24, 20), [0.0, 0.0, 0.0], [48.0, 48.0, 30.0]),
}
TOF = (3.66, 649.2)


def rel_l2(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def golden():
    return np.load(os.path.join(HERE, "golden", "siddon_cases.npz"))


def _plan(golden, gname, tof):
    import torch
    from srf_amd import ops
    shape, center, size = GRIDS[gname]
    lors = golden["lors"] if tof else np.ascontiguousarray(golden["lors"][:, :6])
    tb, ts = TOF if tof else (0.0, 0.0)
    L = torch.from_numpy(np.ascontiguousarray(lors)).cuda()
    return ops.Plan(L, shape, list(center) + list(size), tof, tb, ts), L


@pytest.mark.parametrize("gname", list(GRIDS))
@pytest.mark.parametrize("tof", [False, True])
def test_plan_digest_and_values_vs_golden(golden, gname, tof):
    import torch
    from srf_amd import ops
    print(f"plan {gname} tof={tof}", flush=True)
    plan, L = _plan(golden, gname, tof)
    tag = f"{gname}_{'tof' if tof else 'nontof'}"
    for mode, vname in ((1, "fp"), (0, "bp")):
        h, c = plan.digest(mode)
        torch.cuda.synchronize()
        h = h.cpu().numpy().view(np.uint64)
        c = c.cpu().numpy()
        bad = np.nonzero((h != golden[f"{tag}_{vname}_hash"]) | (c != golden[f"{tag}_{vname}_count"]))[0]
        assert len(bad) == 0, f"{vname}: {len(bad)} events differ, first {bad[:10]}"
    img_k = ops.to_kernel_layout(torch.from_numpy(golden[f"{gname}_image"]).cuda())
    p = plan.project(img_k).cpu().numpy()
    assert rel_l2(p, golden[f"{tag}_proj"]) <= 1e-5
    vals = torch.from_numpy(golden[f"{gname}_values"]).cuda()
    for mode, key in ((0, "bp"), (1, "adj")):
        bp = ops.from_kernel_layout(plan.backproject(vals, mode)).cpu().numpy()
        err = rel_l2(bp, golden[f"{tag}_{key}"])
        print(f"  bp mode={mode} relL2={err:.2e}", flush=True)
        assert err <= 1e-5
    plan.close()


@pytest.mark.parametrize("tof", [False, True])
def test_plan_power_of_two_voxel_ratios_vs_oracle(golden, orc, tof):
    """dy = dz = 2 dx: the plan divides the line constants by the voxel ratio
    once at build time instead of dividing each slab's coordinates by it (an
    exact rewrite for powers of two).  Term digests equal the oracle's bit for
    bit; projection and both backprojections within 1e-5."""
    import torch
    from srf_amd import ops
    shape, center, size = (32, 16, 16), [0.5, -0.25, 0.0], [64.0, 64.0, 64.0]
    lors = golden["lors"] if tof else np.ascontiguousarray(golden["lors"][:, :6])
    t = TOF if tof else None
    tb, ts = TOF if tof else (0.0, 0.0)
    L = torch.from_numpy(np.ascontiguousarray(lors)).cuda()
    print(f"plan dy = dz = 2 dx tof={tof}", flush=True)
    plan = ops.Plan(L, shape, list(center) + list(size), tof, tb, ts)
    for mode, var in ((1, 0), (0, 1)):
        h, c = plan.digest(mode)
        torch.cuda.synchronize()
        wh, wc = orc.digest(lors, shape, center, size, tof=t, variant=var)
        bad = np.nonzero((h.cpu().numpy().view(np.uint64) != wh) | (c.cpu().numpy() != wc))[0]
        assert len(bad) == 0, f"mode {mode}: {len(bad)} events differ, first {bad[:10]}"
    rng = np.random.default_rng(3)
    img = rng.uniform(0.1, 2.0, shape).astype(np.float32)
    vals = rng.uniform(0.5, 1.5, len(lors)).astype(np.float32)
    p = plan.project(ops.to_kernel_layout(torch.from_numpy(img).cuda())).cpu().numpy()
    assert rel_l2(p, orc.project(img, lors, center, size, tof=t, f64=True)) <= 1e-5
    for mode, var in ((0, 1), (1, 0)):
        bp = ops.from_kernel_layout(plan.backproject(torch.from_numpy(vals).cuda(), mode)).cpu().numpy()
        assert rel_l2(bp, orc.backproject(vals, lors, shape, center, size, tof=t, variant=var)) <= 1e-5
    plan.close()


@pytest.mark.parametrize("tof", [False, True])
def test_plan_matches_event_kernels_scanner_size(tof):
    """cfg-2 geometry (256x256x128, 2 mm, R 424 mm): tiled vs one-event-per-lane."""
    import torch
    from srf_amd import ops, phantom
    shape, voxel = (256, 256, 128), 2.0
    act, mu, center, size = phantom.random_ellipsoid_phantom(shape, voxel, 3, device="cuda")
    lors = phantom.sample_events(act, center, size, 200000, 424.0, 300.0, 4,
                                 tof_fwhm_mm=60.0 if tof else None, device="cuda")
    geom = list(center) + list(size)
    tb, ts = TOF if tof else (0.0, 0.0)
    plan = ops.Plan(lors, shape, geom, tof, tb, ts)
    info = plan.info()
    print("plan info", info, flush=True)
    img_k = ops.to_kernel_layout(act.contiguous())
    p0 = ops.project(lors, img_k, list(shape), geom, tof, tb, ts)
    p1 = plan.project(img_k)
    assert rel_l2(p1.cpu().numpy(), p0.cpu().numpy()) <= 1e-5
    y = torch.rand(lors.shape[0], device="cuda") + 0.5
    for mode in (0, 1):
        b0 = ops.backproject_ratio(lors, y, p0, 1, 1e-5, list(shape), geom, tof, tb, ts, mode)
        b1 = plan.backproject_ratio(y, p0, 1, 1e-5, mode)
        assert rel_l2(b1.cpu().numpy(), b0.cpu().numpy()) <= 1e-5
    h0, c0 = ops.lor_digest(lors, list(shape), geom, tof, tb, ts, 1)
    h1, c1 = plan.digest(1)
    assert torch.equal(h0, h1) and torch.equal(c0, c1)
    plan.close()


def test_plan_empty_and_nan():
    import torch
    from srf_amd import ops
    L = torch.tensor([[np.nan, 0, 0, 100, 0, 0], [-100, 1, 1, 100, 1, 1]], dtype=torch.float32, device="cuda")
    plan = ops.Plan(L, (16, 16, 8), [0, 0, 0, 32, 32, 16], False)
    p = plan.project(torch.ones((8, 16, 16), device="cuda")).cpu().numpy()
    assert np.isnan(p[0]) and p[1] > 0
    plan.close()
    E = torch.zeros((0, 6), dtype=torch.float32, device="cuda")
    plan = ops.Plan(E, (16, 16, 8), [0, 0, 0, 32, 32, 16], False)
    assert plan.project(torch.ones((8, 16, 16), device="cuda")).numel() == 0
    bp = plan.backproject(None, 0)
    assert float(bp.abs().sum()) == 0.0
    plan.close()


def test_plan_backprojection_is_deterministic_and_propagates_nan():
    """Fixed-point LDS accumulation: two runs are bit-identical; a NaN value
    switches to the float path and poisons the voxels its event touches."""
    import torch
    from srf_amd import ops
    from siddon_ref import random_cylinder_lors
    lors = torch.from_numpy(random_cylinder_lors(100000, 150.0, 80.0, 9, tof_sigma=25.0)).cuda()
    shape, geom = (48, 48, 24), [0.0, 0.0, 0.0, 96.0, 96.0, 48.0]
    plan = ops.Plan(lors, shape, geom, True, 3.66, 649.2)
    v = torch.rand(lors.shape[0], device="cuda") * 3.0 - 1.0     # signed values
    b1 = plan.backproject(v, 0)
    b2 = plan.backproject(v, 0)
    assert torch.equal(b1, b2)
    ref = ops.backproject(lors, v, list(shape), geom, True, 3.66, 649.2, 0)
    assert rel_l2(b1.cpu().numpy(), ref.cpu().numpy()) <= 1e-5
    v[123] = float("nan")
    b3 = plan.backproject(v, 0)
    assert torch.isnan(b3).any() and not torch.isnan(b3).all()
    plan.close()


def test_plan_project_ratio_fused():
    """srf_plan_project_ratio == srf_ratio(y, srf_plan_project) bit for bit."""
    import torch
    from srf_amd import ops
    from siddon_ref import random_cylinder_lors
    lors = torch.from_numpy(random_cylinder_lors(50000, 150.0, 80.0, 19, tof_sigma=25.0)).cuda()
    shape, geom = (48, 48, 24), [0.0, 0.0, 0.0, 96.0, 96.0, 48.0]
    plan = ops.Plan(lors, shape, geom, True, 3.66, 649.2)
    img = torch.rand((24, 48, 48), device="cuda")
    y = torch.rand(lors.shape[0], device="cuda")
    p = plan.project(img)
    for kind, eps in ((0, 1e-8), (1, 1e-5), (1, 0.0)):
        q = plan.project_ratio(img, y, kind, eps)
        assert torch.equal(q, ops.ratio(y, p, kind, eps))
    plan.close()


def test_tof_window_drops_only_negligible_terms():
    """srf_plan_opts_t.tof_window = 22 at the OSEM workload's geometry: about a
    third fewer terms, projections and backprojections within 1e-6 of the
    exact plan (the dropped TOF weight is < 4e-8 of each event's)."""
    import torch
    from srf_amd import ops, phantom
    shape, voxel = (256, 256, 128), 2.0
    act, mu, center, size = phantom.random_ellipsoid_phantom(shape, voxel, 5, device="cuda")
    s2 = phantom.tof_sigma2_from_fwhm(60.0)
    lors = phantom.sample_events(act, center, size, 200000, 424.0, 300.0, 6, tof_fwhm_mm=60.0, device="cuda")
    geom = list(center) + list(size)
    exact = ops.Plan(lors, shape, geom, True, 3.66, s2)
    win = ops.Plan(lors, shape, geom, True, 3.66, s2, tof_window=22.0)
    for mode in (1, 0):
        _, c0 = exact.digest(mode)
        _, c1 = win.digest(mode)
        t0, t1 = float(c0.double().sum()), float(c1.double().sum())
        print(f"mode {mode}: terms exact {t0:.4e} windowed {t1:.4e} ({t1 / t0:.3f})", flush=True)
        assert bool((c1 <= c0).all()) and t1 < 0.8 * t0
    img_k = ops.to_kernel_layout(act.contiguous())
    p0, p1 = exact.project(img_k), win.project(img_k)
    assert rel_l2(p1.cpu().numpy(), p0.cpu().numpy()) <= 1e-6
    y = torch.rand(lors.shape[0], device="cuda") + 0.5
    b0, b1 = exact.backproject(y, 0), win.backproject(y, 0)
    assert rel_l2(b1.cpu().numpy(), b0.cpu().numpy()) <= 1e-6
    exact.close()
    win.close()


@pytest.mark.parametrize("tof", [False, True])
def test_plan_rebuild_equals_fresh_plan(tof):
    """srf_plan_rebuild reuses a plan's device memory: after rebuilding onto a
    larger list, a smaller one and an empty one, every result is bit-identical
    to a freshly created plan of the same events."""
    import torch
    from srf_amd import ops
    from srf_amd.functions import uniform_cylinder_lors
    shape, geom = (64, 64, 32), [0.0, 0.0, 0.0, 256.0, 256.0, 128.0]
    tb, ts = (3.66, 649.2) if tof else (0.0, 0.0)

    def lors_of(n, seed):
        L = uniform_cylinder_lors(n, 140.0, 120.0, seed, torch.device("cuda"))
        if tof:
            t = (torch.rand(n, 1, device="cuda", generator=torch.Generator("cuda").manual_seed(seed)) - 0.5) * 200
            L = torch.cat([L, t], 1).contiguous()
        return L

    img = torch.rand((32, 64, 64), device="cuda", generator=torch.Generator("cuda").manual_seed(3))
    plan = ops.Plan(lors_of(20000, 1), shape, geom, tof, tb, ts)
    for n, seed in ((60000, 2), (15000, 3), (0, 4), (30000, 5)):
        L = lors_of(n, seed) if n else torch.empty((0, 7 if tof else 6), device="cuda")
        plan.rebuild(L)
        fresh = ops.Plan(L, shape, geom, tof, tb, ts)
        assert plan.info()["segments"] == fresh.info()["segments"]
        print(f"n {n}: {plan.info()}", flush=True)
        if n:
            assert torch.equal(plan.project(img), fresh.project(img))
            y = torch.rand(n, device="cuda", generator=torch.Generator("cuda").manual_seed(seed)) + 0.5
            assert torch.equal(plan.backproject(y, 0), fresh.backproject(y, 0))
            assert torch.equal(plan.backproject(y, 1), fresh.backproject(y, 1))
        else:
            assert float(plan.backproject(None, 0).abs().sum()) == 0.0
        fresh.close()
    plan.close()


def test_backproject_update_equals_two_calls():
    """srf_plan_backproject_update == srf_plan_backproject then srf_em_update, bit for bit
    (emap zeros guarded)."""
    import torch
    from srf_amd import ops
    from srf_amd.functions import uniform_cylinder_lors
    shape, geom = (64, 64, 32), [0.0, 0.0, 0.0, 256.0, 256.0, 128.0]
    L = uniform_cylinder_lors(50000, 140.0, 120.0, 9, torch.device("cuda"))
    plan = ops.Plan(L, shape, geom, False)
    g = torch.Generator("cuda").manual_seed(5)
    q = torch.rand(50000, device="cuda", generator=g) + 0.5
    x = torch.rand((32, 64, 64), device="cuda", generator=g) + 0.1
    emap = torch.rand((32, 64, 64), device="cuda", generator=g)
    emap[::3, ::5, ::7] = 0.0
    ref = x.clone()
    ops.em_update(ref, plan.backproject(q, 0), emap, 1e8)
    fused = plan.backproject_update(q, x.clone(), emap, 1e8)
    assert torch.equal(fused, ref)
    plan.close()


def test_time_next_brackets_the_tiled_kernel():
    """srf_plan_time_next: the events bracket one launch (a positive interval no
    longer than the whole call), are dropped after it, and change no result."""
    import torch
    from srf_amd import ops
    from srf_amd.functions import uniform_cylinder_lors
    shape, geom = (64, 64, 32), [0.0, 0.0, 0.0, 256.0, 256.0, 128.0]
    L = uniform_cylinder_lors(200000, 140.0, 120.0, 11, torch.device("cuda"))
    plan = ops.Plan(L, shape, geom, False)
    x = torch.rand((32, 64, 64), device="cuda", generator=torch.Generator("cuda").manual_seed(3)) + 0.1
    ref = plan.project(x)
    a, b, c, d = (torch.cuda.Event(enable_timing=True) for _ in range(4))
    plan.time_next(a, b)
    c.record()
    out = plan.project(x)
    d.record()
    again = plan.project(x)   # no pending events: nothing recorded
    torch.cuda.synchronize()
    assert torch.equal(out, ref) and torch.equal(again, ref)
    k, whole = a.elapsed_time(b), c.elapsed_time(d)
    print(f"kernel {k:.4f} ms of {whole:.4f} ms", flush=True)
    assert 0.0 < k <= whole
    plan.close()


def test_full_subset_size_properties():
    """One OSEM subset of the headline workload at its full size (5e6 TOF events,
    256x256x128 x 2 mm): the plan's digest equals the event kernels' (every
    voxel index and branch code of every term, bit-exact), its projection is
    within 1e-5 of theirs, and its exact-adjoint backprojection satisfies
    <A x, y> = <x, A^T y> to 1e-5 with both sums in float64."""
    import torch
    from srf_amd import ops, phantom
    shape, voxel = (256, 256, 128), 2.0
    act, mu, center, size = phantom.random_ellipsoid_phantom(shape, voxel, 9, device="cuda")
    s2 = phantom.tof_sigma2_from_fwhm(60.0)
    n = 5_000_000
    lors = phantom.sample_events(act, center, size, n, 424.0, 300.0, 10, tof_fwhm_mm=60.0, device="cuda")
    geom = list(center) + list(size)
    print(f"events {n}, plan of the full subset", flush=True)
    plan = ops.Plan(lors, shape, geom, True, 3.66, s2)
    h0, c0 = ops.lor_digest(lors, list(shape), geom, True, 3.66, s2, 1)
    h1, c1 = plan.digest(1)
    assert torch.equal(h0, h1) and torch.equal(c0, c1)
    del h0, h1, c0, c1
    img_k = ops.to_kernel_layout(act.contiguous())
    p0 = ops.project(lors, img_k, list(shape), geom, True, 3.66, s2)
    p1 = plan.project(img_k)
    err = rel_l2(p1.cpu().numpy(), p0.cpu().numpy())
    print(f"projection vs event kernels: relative L2 {err:.3g}", flush=True)
    assert err <= 1e-5
    y = torch.rand(n, device="cuda", generator=torch.Generator("cuda").manual_seed(12)) + 0.5
    bt = plan.backproject(y, 1)
    lhs = float((p1.double() * y.double()).sum())
    rhs = float((img_k.double() * bt.double()).sum())
    print(f"<Ax,y> {lhs:.10g} <x,A^T y> {rhs:.10g}", flush=True)
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)
    plan.close()


@pytest.mark.parametrize("shape,voxel", [((96, 96, 106), 2.8), ((40, 36, 13), 3.0), ((33, 64, 9), 2.0),
                                         ((64, 64, 64), 2.5)],
                         ids=["96x96x106", "40x36x13", "33x64x9", "64x64x64"])
@pytest.mark.parametrize("window", [0.0, 22.0])
def test_plan_on_grids_not_aligned_to_tiles(orc, shape, voxel, window):
    """Tiled plans on grids whose sizes are not multiples of the 32 x 32 x 8
    tile (partial tiles at the high edges): the plan's digest equals the event
    kernels' (window 0) or holds a subset of their terms (window 22), and its
    projection and backprojection are within 1e-5 of the oracle's."""
    import torch
    from srf_amd import ops, phantom
    act, mu, center, size = phantom.random_ellipsoid_phantom(shape, voxel, 21, device="cuda")
    s2 = phantom.tof_sigma2_from_fwhm(60.0)
    lors = phantom.sample_events(act, center, size, 40000, 424.0, 300.0, 22, tof_fwhm_mm=60.0, device="cuda")
    geom = list(center) + list(size)
    plan = ops.Plan(lors, list(shape), geom, True, 3.66, s2, tof_window=window)
    for mode in (1, 0):
        h0, c0 = ops.lor_digest(lors, list(shape), geom, True, 3.66, s2, mode)
        h1, c1 = plan.digest(mode)
        if window == 0.0:
            bad = int((h0 != h1).sum() + (c0 != c1).sum())
            print(f"{shape} mode {mode}: {bad} events differ from the event kernels", flush=True)
            assert bad == 0
        else:
            assert bool((c1 <= c0).all())
    ln = lors.cpu().numpy()
    a = act.cpu().numpy()
    p = plan.project(ops.to_kernel_layout(act.contiguous())).cpu().numpy()
    p_o = orc.project(a, ln, center, size, tof=(3.66, s2), f64=True)
    y = np.random.default_rng(3).uniform(0.5, 1.5, len(ln)).astype(np.float32)
    bp = ops.from_kernel_layout(plan.backproject(torch.from_numpy(y).cuda(), 0)).cpu().numpy()
    bp_o = orc.backproject(y, ln, shape, center, size, tof=(3.66, s2))
    ep, eb = rel_l2(p, p_o), rel_l2(bp, bp_o)
    print(f"{shape} x {voxel} mm, window {window}: projection {ep:.3g}, backprojection {eb:.3g}", flush=True)
    assert ep <= 1e-5 and eb <= 1e-5
    plan.close()


def test_tof_window_keeps_events_centred_outside_the_image():
    """Events whose TOF centre lies 10-80 mm outside a 128 mm field of view
    (their in-image terms are all Gaussian tail) mixed with ordinary ones:
    the EM ratio y / sum(terms) spreads each of them over its tail, so a plan
    with TOF window 22 must keep that tail (the window is relative to the
    largest in-image factor).  One EM sub-iteration with window 22 equals
    the exact plan's to 1e-6."""
    import torch
    from srf_amd import ops
    g = torch.Generator().manual_seed(31)
    n = 60000
    shape, geom = [64, 64, 32], [0.0, 0.0, 0.0, 128.0, 128.0, 64.0]
    yz = (torch.rand(n, 2, generator=g) - 0.5) * torch.tensor([100.0, 50.0])
    slope = (torch.rand(n, 2, generator=g) - 0.5) * 0.3
    p1 = torch.stack([torch.full((n,), -300.0), yz[:, 0] - 300 * slope[:, 0], yz[:, 1] - 300 * slope[:, 1]], 1)
    p2 = torch.stack([torch.full((n,), 300.0), yz[:, 0] + 300 * slope[:, 0], yz[:, 1] + 300 * slope[:, 1]], 1)
    # TOF centre s mm from the midpoint towards p2 (tof = -2 s): half inside
    # the field of view, half 10-80 mm beyond its edge on either side
    s = (torch.rand(n, generator=g) - 0.5) * 100.0
    out = torch.rand(n, generator=g) < 0.5
    side = torch.where(torch.rand(n, generator=g) < 0.5, -1.0, 1.0)
    s = torch.where(out, side * (64.0 + 10.0 + 70.0 * torch.rand(n, generator=g)), s)
    lors = torch.cat([p1, p2, (-2.0 * s)[:, None]], 1).float().cuda().contiguous()
    exact = ops.Plan(lors, shape, geom, True, 3.66, 649.2)
    win = ops.Plan(lors, shape, geom, True, 3.66, 649.2, tof_window=22.0)
    x0 = torch.rand((32, 64, 64), device="cuda", generator=torch.Generator("cuda").manual_seed(2)) + 0.5
    emap = torch.ones_like(x0)
    y = torch.ones(n, device="cuda")
    res = []
    for plan in (exact, win):
        x = x0.clone()
        q = plan.project_ratio(x, y, 0, 1e-8)
        plan.backproject_update(q, x, emap, 0.0)
        res.append(x.cpu().numpy())
    _, c0 = exact.digest(0)
    _, c1 = win.digest(0)
    err = rel_l2(res[1], res[0])
    print(f"terms kept {float(c1.double().sum() / c0.double().sum()):.3f}; one sub-iteration, window 22 vs "
          f"exact: {err:.3g}", flush=True)
    assert err <= 1e-6
    exact.close()
    win.close()


def test_backprojection_keeps_small_values_beside_huge_ones(orc):
    """Work items whose event values span 15 decades (the EM ratio of an event
    with a projection ~0 is 1e8 times the others'): 1% of the events carry 1e9,
    the rest ~1e-6.  One fixed-point scale per item resolves 2^-50 of the
    largest term, ~1e-6 here, and would round the small events' terms away;
    with exponent bands every voxel whose value comes at least half from the
    small events is within 1e-5 of the oracle's float64 sum, relative to that
    voxel's own value.  Non-TOF, so that every term of a 1e9 event is large
    (with TOF its far tail can be as small as the small events' terms, and
    sits in the 1e9 band at that band's absolute resolution)."""
    import torch
    from siddon_ref import random_cylinder_lors
    from srf_amd import ops
    lors = np.ascontiguousarray(random_cylinder_lors(40000, 150.0, 80.0, 23)[:, :6])
    shape, center, size = (48, 48, 24), [0.0, 0.0, 0.0], [96.0, 96.0, 48.0]
    rng = np.random.default_rng(4)
    v = rng.uniform(0.5e-6, 1.5e-6, len(lors)).astype(np.float32)
    v[rng.random(len(lors)) < 0.01] = 1e9
    plan = ops.Plan(torch.from_numpy(lors).cuda(), list(shape), center + size, False)
    small_v = np.where(v > 1.0, 0, v).astype(np.float32)
    for mode in (0, 1):
        var = orc.BP if mode == 0 else orc.FP
        bp = ops.from_kernel_layout(plan.backproject(torch.from_numpy(v).cuda(), mode)).cpu().numpy()
        want = orc.backproject(v, lors, shape, center, size, variant=var)
        small = orc.backproject(small_v, lors, shape, center, size, variant=var)
        m = (want > 0) & (small >= 0.5 * want)
        rel = np.abs(bp[m] - want[m]) / want[m]
        print(f"mode {mode}: {int(m.sum())} voxels mostly from the 1e-6 events, max relative error {rel.max():.3g}; "
              f"relative L2 over all {rel_l2(bp, want):.3g}", flush=True)
        assert m.sum() > 1000 and rel.max() <= 1e-5 and rel_l2(bp, want) <= 1e-6
    plan.close()
