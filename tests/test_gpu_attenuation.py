"""Attenuation correction (AttenuationCorrect / UmapProject, srfnef
corrections/attenuation/attenuation_correct.py:16-40) on the GPU against the
oracle's restatement: exp(-Project(u) * L^2 / unit) (reference scaling) and
/ unit^2 (physical), values divided by the factor.  Bar: 1e-5 relative L2."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("physical", [False, True])
def test_attenuation_correct_matches_oracle(orc, physical):
    from siddon_ref import random_cylinder_lors
    from srf_amd import Listmode, Lors, phantom
    from srf_amd.corrections import AttenuationCorrect
    from srf_amd.data import Umap
    shape, voxel = (64, 64, 32), 4.0
    act, mu, center, size = phantom.random_ellipsoid_phantom(shape, voxel, 13)
    mu = np.asarray(mu, np.float32)
    lors = random_cylinder_lors(100000, 300.0, 120.0, 14, tof_sigma=25.0)   # (N, 7): TOF column ignored
    y = np.random.default_rng(15).uniform(0.5, 2.0, len(lors)).astype(np.float32)
    out = AttenuationCorrect(Umap(mu, list(center), list(size)), physical)(Listmode(y, Lors(lors)))
    a = orc.attenuation_factor(mu, lors[:, :6], list(center), list(size), physical=physical)
    want = y / a
    got = out.data.cpu().numpy() if isinstance(out.data, torch.Tensor) else np.asarray(out.data)
    err = rel_l2(got, want)
    print(f"attenuation (physical={physical}): factors {a.min():.3g}..{a.max():.3g}, relative L2 {err:.3g}",
          flush=True)
    assert a.min() < 0.5 and err <= 1e-5
