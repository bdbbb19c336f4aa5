"""GPU parity: libsrf_hip.so (through the C-ABI, via the srf:: torch ops)
against the CPU oracle and the golden fixtures.

Bar: term digests (voxel index + branch of every term) bit-identical;
projections and backprojections within 1e-5 relative L2 in FP32.
"""
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
RTOL_L2 = 1e-5


def rel_l2(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def golden():
    return np.load(os.path.join(HERE, "golden", "siddon_cases.npz"))


@pytest.fixture(scope="module")
def T():
    import torch
    from srf_amd import ops  # noqa: F401
    return torch


def _gpu(T, a):
    return T.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.parametrize("gname", list(GRIDS))
@pytest.mark.parametrize("tof", [False, True])
@pytest.mark.parametrize("mode", [1, 0])   # 1: projector predicates, 0: backprojector predicates
def test_digest_bit_exact_vs_golden(T, golden, gname, tof, mode):
    from srf_amd import ops
    shape, center, size = GRIDS[gname]
    lors = golden["lors"] if tof else np.ascontiguousarray(golden["lors"][:, :6])
    tag = f"{gname}_{'tof' if tof else 'nontof'}_{'fp' if mode == 1 else 'bp'}"
    print(f"digest {tag} n={len(lors)}", flush=True)
    tb, ts = TOF if tof else (0.0, 0.0)
    h, c = ops.lor_digest(_gpu(T, lors), list(shape), list(center) + list(size), tof, tb, ts, mode)
    T.cuda.synchronize()
    h = h.cpu().numpy().view(np.uint64)
    c = c.cpu().numpy()
    want_h, want_c = golden[tag + "_hash"], golden[tag + "_count"]
    bad = np.nonzero((h != want_h) | (c != want_c))[0]
    assert len(bad) == 0, f"{len(bad)} events differ, first {bad[:10]}"


@pytest.mark.parametrize("gname", list(GRIDS))
@pytest.mark.parametrize("tof", [False, True])
def test_projection_vs_golden(T, golden, gname, tof):
    from srf_amd import ops
    shape, center, size = GRIDS[gname]
    lors = golden["lors"] if tof else np.ascontiguousarray(golden["lors"][:, :6])
    img = golden[f"{gname}_image"]
    tb, ts = TOF if tof else (0.0, 0.0)
    img_k = ops.to_kernel_layout(_gpu(T, img))
    p = ops.project(_gpu(T, lors), img_k, list(shape), list(center) + list(size), tof, tb, ts)
    want = golden[f"{gname}_{'tof' if tof else 'nontof'}_proj"]
    err = rel_l2(p.cpu().numpy(), want)
    print(f"proj {gname} tof={tof} relL2={err:.3e}", flush=True)
    assert err <= RTOL_L2


@pytest.mark.parametrize("gname", list(GRIDS))
@pytest.mark.parametrize("tof", [False, True])
@pytest.mark.parametrize("mode", [0, 1])
def test_backprojection_vs_golden(T, golden, gname, tof, mode):
    from srf_amd import ops
    shape, center, size = GRIDS[gname]
    lors = golden["lors"] if tof else np.ascontiguousarray(golden["lors"][:, :6])
    vals = golden[f"{gname}_values"]
    tb, ts = TOF if tof else (0.0, 0.0)
    bp_k = ops.backproject(_gpu(T, lors), _gpu(T, vals), list(shape), list(center) + list(size), tof, tb, ts, mode)
    bp = ops.from_kernel_layout(bp_k).cpu().numpy()
    want = golden[f"{gname}_{'tof' if tof else 'nontof'}_{'adj' if mode == 1 else 'bp'}"]
    err = rel_l2(bp, want)
    print(f"bp {gname} tof={tof} mode={mode} relL2={err:.3e}", flush=True)
    assert err <= RTOL_L2


@pytest.mark.parametrize("tof", [False, True])
def test_live_oracle_parity_random_cylinder(T, orc, tof):
    """Seeded LORs on a scanner-sized cylinder, 64x64x32 image: GPU vs live oracle."""
    from siddon_ref import random_cylinder_lors
    from srf_amd import ops
    shape, center, size = (64, 64, 32), [0.0, 0.0, 0.0], [217.6, 217.6, 108.8]
    rng = np.random.default_rng(11)
    lors = random_cylinder_lors(20000, 300.0, 110.0, 12, tof_sigma=25.48 if tof else None)
    img = rng.uniform(0, 1, shape).astype(np.float32)
    t = TOF if tof else None
    tb, ts = TOF if tof else (0.0, 0.0)
    geom = list(center) + list(size)
    p = ops.project(_gpu(T, lors), ops.to_kernel_layout(_gpu(T, img)), list(shape), geom, tof, tb, ts)
    want = orc.project(img, lors, center, size, tof=t, f64=True)
    assert rel_l2(p.cpu().numpy(), want) <= RTOL_L2
    h, c = ops.lor_digest(_gpu(T, lors), list(shape), geom, tof, tb, ts, 1)
    wh, wc = orc.digest(lors, shape, center, size, tof=t, variant=0)
    assert np.array_equal(h.cpu().numpy().view(np.uint64), wh) and np.array_equal(c.cpu().numpy(), wc)
    vals = rng.uniform(0.5, 1.5, len(lors)).astype(np.float32)
    bp = ops.from_kernel_layout(ops.backproject(_gpu(T, lors), _gpu(T, vals), list(shape), geom, tof, tb, ts, 0))
    wbp = orc.backproject(vals, lors, shape, center, size, tof=t, variant=1)
    assert rel_l2(bp.cpu().numpy(), wbp) <= RTOL_L2


def test_adjoint_dot_product_fp64(T):
    """<A x, y> == <x, A^T y> with A^T the adjoint-mode backprojector (FP64 sums)."""
    from siddon_ref import random_cylinder_lors
    from srf_amd import ops
    shape, center, size = (48, 48, 24), [0.0, 0.0, 0.0], [96.0, 96.0, 48.0]
    rng = np.random.default_rng(3)
    lors = random_cylinder_lors(50000, 150.0, 80.0, 4, tof_sigma=25.48)
    x = rng.uniform(0, 1, shape).astype(np.float32)
    y = rng.uniform(0, 1, len(lors)).astype(np.float32)
    geom = list(center) + list(size)
    for tof in (False, True):
        tb, ts = TOF if tof else (0.0, 0.0)
        L = _gpu(T, lors if tof else np.ascontiguousarray(lors[:, :6]))
        Ax = ops.project(L, ops.to_kernel_layout(_gpu(T, x)), list(shape), geom, tof, tb, ts)
        Aty = ops.from_kernel_layout(ops.backproject(L, _gpu(T, y), list(shape), geom, tof, tb, ts, 1))
        lhs = float(np.dot(Ax.cpu().numpy().astype(np.float64), y.astype(np.float64)))
        rhs = float(np.sum(x.astype(np.float64) * Aty.cpu().numpy().astype(np.float64)))
        assert abs(lhs - rhs) <= 1e-5 * abs(lhs)


def test_elementwise_ops(T):
    from srf_amd import ops
    rng = np.random.default_rng(0)
    y = rng.uniform(0, 2, 1001).astype(np.float32)
    p = rng.uniform(0, 1, 1001).astype(np.float32)
    p[::7] = 0.0
    q = ops.ratio(_gpu(T, y), _gpu(T, p), 1, 0.0).cpu().numpy()
    with np.errstate(divide="ignore", invalid="ignore"):
        want = np.where(p == 0, 0, y / p).astype(np.float32)
    np.testing.assert_array_equal(q, want)
    q2 = ops.ratio(_gpu(T, y), _gpu(T, p), 0, 1e-8).cpu().numpy()
    np.testing.assert_array_equal(q2, y / (p + np.float32(1e-8)))
    x = rng.uniform(0, 1, (7, 9, 13)).astype(np.float32)
    bp = rng.uniform(0, 1, x.shape).astype(np.float32)
    em = rng.uniform(0, 1, x.shape).astype(np.float32)
    em[0, 0, :5] = 0
    xt = _gpu(T, x)
    ops.em_update(xt, _gpu(T, bp), _gpu(T, em), 1e8)
    want = x * bp / np.where(em == 0, np.float32(1e8), em)
    np.testing.assert_array_equal(xt.cpu().numpy(), want)
    # layout round trip
    img = rng.uniform(0, 1, (33, 17, 9)).astype(np.float32)
    k = ops.to_kernel_layout(_gpu(T, img))
    np.testing.assert_array_equal(k.cpu().numpy(), np.transpose(img, (2, 1, 0)))
    np.testing.assert_array_equal(ops.from_kernel_layout(k).cpu().numpy(), img)


def test_nan_lors_and_empty(T):
    from srf_amd import ops
    lors = np.array([[np.nan, 0, 0, 100, 0, 0], [-100, 1, 1, 100, 1, 1]], np.float32)
    img = np.ones((16, 16, 8), np.float32)
    p = ops.project(_gpu(T, lors), ops.to_kernel_layout(_gpu(T, img)), [16, 16, 8], [0, 0, 0, 32, 32, 16],
                    False, 0.0, 0.0).cpu().numpy()
    assert np.isnan(p[0]) and p[1] > 0
    e = ops.project(_gpu(T, np.zeros((0, 6), np.float32)), ops.to_kernel_layout(_gpu(T, img)), [16, 16, 8],
                    [0, 0, 0, 32, 32, 16], False, 0.0, 0.0)
    assert e.numel() == 0


@pytest.mark.parametrize("vox", [1.0, 2.0])
def test_tof_width_closed_form(T, vox):
    """The GPU's TOF Gaussian has variance tof_sigma2/dx_^2 + (tof_bin/dx_)^2/12
    in voxels (siddon_tof.cu.cc:56-59) and integral tof_bin_: the moments of
    its per-slab weights along an axis-aligned LOR, from projections of the
    images 1, k and k^2 (test_oracle.tof_moments), against the closed form --
    not against the oracle, so a slip shared by both would fail here."""
    from srf_amd import ops
    from test_oracle import tof_moments
    tof_bin, sig2 = 12.0 * vox, 9.0 * vox * vox

    def project(img, lor, center, size):
        p = ops.project(_gpu(T, lor), ops.to_kernel_layout(_gpu(T, img)), list(img.shape),
                        list(center) + list(size), True, tof_bin, sig2)
        return p.double().cpu().numpy()

    for tofv in (0.0, 30.0 * vox):
        s0, mean, var = tof_moments(project, vox, tof_bin, sig2, tofv)
        print(f"vox {vox} tof {tofv}: integral {s0 * 800.0 ** 2 / vox:.6f} mean {mean:.4f} variance {var:.5f}",
              flush=True)
        assert s0 * 800.0 * 800.0 / vox == pytest.approx(tof_bin / vox, rel=1e-5)
        assert var == pytest.approx(21.0, rel=1e-4)
        assert mean == pytest.approx(-0.5 - tofv / vox / 2.0, abs=1e-3)
