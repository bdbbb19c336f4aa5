"""GPU single-scatter simulation against the float64 oracle (formula-level
parity: the kernel evaluates exp / acos in fp32 and the integer-degree
quantisation may flip on a boundary), plus the ScatterCorrect wrapper."""
import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _small_geometry():
    from srf_amd.sss import SssGeometry
    return SssGeometry(radius=424.0, axial_length=300.0, n_det=32, n_rings=4, n_angles=8, n_radial=8,
                       fov_radius=128.0, spacing=32.0)


def _phantom():
    from srf_amd import phantom
    act, mu, center, size = phantom.random_ellipsoid_phantom((32, 32, 16), 8.0, 11, body_axes=(110, 80, 50))
    return act, mu, center, size


@pytest.mark.parametrize("physical_kn", [False, True])
def test_sss_kernel_matches_oracle(orc, physical_kn):
    import sss_oracle as so
    from srf_amd.sss import SingleScatterSimulation
    g = _small_geometry()
    act, mu, center, size = _phantom()
    sss = SingleScatterSimulation(g, act.shape, center, size, physical_kn=physical_kn)
    out = sss(act, mu).cpu().numpy().ravel()
    det, pts, em, mui, bins = sss.inputs_for_oracle(act, mu)
    want = so.sss_sinogram(det[:, :3], det[:, 3], pts[:, :3], pts[:, 3], em, mui, bins, g.low_energy,
                           g.high_energy, g.resolution, physical_kn=physical_kn)
    err = np.linalg.norm(out - want) / np.linalg.norm(want)
    print(f"SSS kernel vs oracle relL2 {err:.2e} (physical_kn={physical_kn})", flush=True)
    # measured 8e-8: the fp32 exp / acos; an integer-degree flip would move
    # one (bin, point) term by a table step and show well above this bar
    assert err <= 1e-5
    # the line integrals (detector -> scatter point, the reference's projector
    # times L^2 / dx^2, preprocess.py:46-54) against the C oracle's projector
    nd, ns = em.shape
    lors = np.concatenate([np.repeat(det[:, None, :3], ns, 1), np.repeat(pts[None, :, :3], nd, 0)], 2)
    lors = lors.reshape(-1, 6).astype(np.float32)
    to_int = ((lors[:, 3:6] - lors[:, 0:3]).astype(np.float64) ** 2).sum(1) / (size[0] / act.shape[0]) ** 2
    for name, img, got in (("emission", act.numpy(), em), ("mu", mu.numpy(), mui)):
        want_li = orc.project(img, lors, center, size).astype(np.float64) * to_int
        e_li = np.linalg.norm(got.ravel() - want_li) / np.linalg.norm(want_li)
        print(f"{name} line integrals ({nd} x {ns}) vs oracle projector relL2 {e_li:.2e}", flush=True)
        assert e_li <= 1e-5


def test_sss_zero_mu_gives_zero():
    from srf_amd.sss import SingleScatterSimulation
    g = _small_geometry()
    act, mu, center, size = _phantom()
    sss = SingleScatterSimulation(g, act.shape, center, size)
    assert float(sss(act, torch.zeros_like(mu)).abs().sum()) == 0.0


def test_scatter_correct_subtracts_fraction():
    from srf_amd import Listmode, Lors, phantom
    from srf_amd.data import Image, Umap
    from srf_amd.sss import ScatterCorrect
    act, mu, center, size = _phantom()
    lors = phantom.sample_events(act, center, size, 20000, 424.0, 300.0, 3)
    lm = Listmode(np.ones(20000, np.float32), Lors(lors.numpy()))
    sc = ScatterCorrect(400.0, 650.0, 0.15, 0.3, _small_geometry())
    out = sc(Image(act.numpy(), center, size), Umap(mu.numpy(), center, size), None, lm)
    assert out.data.shape == (20000,)
    assert float(np.sum(lm.data) - np.sum(out.data)) == pytest.approx(0.3 * 20000, rel=1e-4)


def test_mu_integrals_kept_for_the_same_mu_tensor():
    """The scatter points and attenuation integrals are reused while the same
    mu tensor comes back unmodified; an in-place change or another tensor
    recomputes them (results equal to a fresh simulation each time)."""
    from srf_amd.sss import SingleScatterSimulation
    g = _small_geometry()
    act, mu, center, size = _phantom()
    act, mu = act.cuda(), mu.cuda()
    sss = SingleScatterSimulation(g, act.shape, center, size)
    first = sss(act, mu)
    assert sss._mu_cache is not None and sss._mu_cache[0] is mu
    kept = sss._mu_cache[3]
    again = sss(act * 2.0, mu)
    assert sss._mu_cache[3] is kept     # reused
    assert torch.equal(again, SingleScatterSimulation(g, act.shape, center, size)(act * 2.0, mu))
    assert torch.equal(first, SingleScatterSimulation(g, act.shape, center, size)(act, mu))
    mu.mul_(1.5)                         # in place: the version counter moves
    changed = sss(act, mu)
    assert sss._mu_cache[3] is not kept
    assert torch.equal(changed, SingleScatterSimulation(g, act.shape, center, size)(act, mu))
    other = mu.clone()
    assert torch.equal(sss(act, other), changed) and sss._mu_cache[0] is other
