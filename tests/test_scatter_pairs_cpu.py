"""Pins the oracle's restatement of the reference's crystal-pair scatter
correction (oracle/sss_oracle.py, ref_*; srfnef/corrections/scattering/
scatter.py, preprocess.py, scatter_fraction/scatter_fraction.py) and the
product's host-side pieces (srf_amd.scatter_pairs) without a GPU:

- block normals (get_block_theta, scatter_fraction.py:122-131) equal the
  block rotation on the central plane, and off it follow the reference's 3-D
  angle (quirk Q23, a hand-worked case);
- project_area is 1 along a block normal; get_scale of two facing block
  centres is |AB|^2 times the constants;
- the per-pair sum equals the independently pinned virtual-detector oracle
  (sss_sinogram) given crystals as detectors, times the full constant;
- mu = 0 gives no scatter; the sum is symmetric in the pair;
- the bin -> pair decoding equals get_all_lors_id's order, on the host and in
  the product's torch decoder;
- scatter positions follow preprocess.py:6-18 (closed form on a grid).
"""
import math

import numpy as np
import pytest
import torch

import sss_oracle as so


def _ecat():
    from srf_amd.scanner import Block, PetEcatScanner
    return PetEcatScanner(100.0, 120.0, 4, 12, 0.0, Block([20.0, 40.0, 40.0], [1, 4, 4]))


@pytest.mark.parametrize("nb_blocks", [12, 48])
def test_block_normals_on_the_central_plane_are_block_rotations(nb_blocks):
    """One ring, one crystal row (z = 0): the normal is the block's rotation."""
    from srf_amd import scanner as S
    from srf_amd.scatter_pairs import block_normals, block_params
    sc = S.PetEcatScanner(400.0, 420.0, 1, nb_blocks, 0.0, S.Block([20.0, 50.0, 10.0], [1, 10, 1]))
    pos = S.crystal_positions(sc)
    assert np.all(pos[:, 2] == 0)
    ind = np.arange(sc.nb_crystals)
    want = (ind // 10) % nb_blocks * 2 * math.pi / nb_blocks
    nbr, ncr, grid, _ = block_params(sc)
    got = so.ref_block_theta(pos, ncr, nbr, grid[1])
    assert np.allclose(got, want, atol=1e-12)
    assert np.array_equal(block_normals(pos, ncr, nbr, grid[1]), got)


def test_block_normal_off_plane_quirk():
    """get_block_theta (scatter_fraction.py:124-126) takes acos(x / |p|) with the
    3-D norm, so off the central plane the angle grows (quirk Q23).  Worked by
    hand: p = (110, -15, -75): |p| = 134.0, acos(110 / 134.0) = 34.8 deg, y < 0
    so 325.2 deg; plus half a 30-degree block, 340.2 deg, is block 11 (330
    deg), where the azimuth (-7.8 deg) lies in block 0."""
    p = np.array([[110.0, -15.0, -75.0], [110.0, -15.0, 0.0]])
    got = so.ref_block_theta(p, 48, 12, 4)
    assert got[0] == pytest.approx(11 * math.pi / 6) and got[1] == 0.0


def test_project_area_and_scale_closed_forms():
    # a point on block 0's normal: obliquity 1; the reverse direction: -1
    assert so.ref_project_area([10.0, 0.0, 5.0], [110.0, 0.0, 5.0], 0.0) == pytest.approx(1.0)
    assert so.ref_project_area([210.0, 0.0, 5.0], [110.0, 0.0, 5.0], 0.0) == pytest.approx(-1.0)
    # 45 degrees off the normal
    assert so.ref_project_area([0.0, -100.0, 0.0], [100.0, 0.0, 0.0], 0.0) == pytest.approx(math.sqrt(0.5))
    # two facing block centres: scale = |AB|^2 * 4 pi / e_ws^2 / area^2
    pos = np.array([[110.0, 0.0, 0.0], [-110.0, 0.0, 0.0]], np.float32)
    pts = np.array([[0.0, 0.0, 0.0]], np.float32)
    em = np.ones(2, np.float32)
    at = np.ones(2, np.float32)
    mu = np.zeros((4, 4, 4), np.float32)
    scat, scale = so.ref_scatter_fraction([(0, 1)], pos, pts, em, at, mu, [40.0] * 3, 400, 650, 0.15,
                                          2, 2, [1, 1, 1], [20.0, 10.0, 10.0])
    sigma = math.sqrt(511.0) * 0.15
    e_ws = so.ref_eff_without_scatter(400, 650, sigma) * 5 / math.sqrt(1022 * math.pi) / 0.15
    assert scale[0] == pytest.approx(220.0 ** 2 * 4 * math.pi / e_ws ** 2 / 100.0 ** 2, rel=1e-12)
    assert scat[0] == 0.0                      # mu = 0 at the point: no scatter


def _problem(seed=1):
    from srf_amd.scanner import crystal_positions
    sc = _ecat()
    pos = crystal_positions(sc)
    pts = so.ref_scatter_positions((96, 96, 80), (192.0, 192.0, 160.0))
    rng = np.random.default_rng(seed)
    n, ns = pos.shape[0], pts.shape[0]
    em = rng.uniform(0.0, 50.0, n * ns).astype(np.float32)
    at = np.exp(-rng.uniform(0.0, 2.0, n * ns)).astype(np.float32)
    mu = rng.choice([0.0, 0.0096, 0.0172], (96, 96, 80)).astype(np.float32)
    return sc, pos, pts, em, at, mu


def test_pair_sum_equals_virtual_detector_oracle():
    from srf_amd.scatter_pairs import block_params
    sc, pos, pts, em, at, mu = _problem()
    nbr, ncr, grid, size_block = block_params(sc)
    rng = np.random.default_rng(2)
    a = rng.integers(0, sc.nb_crystals, 300)
    b = (a + rng.integers(1, sc.nb_crystals, 300)) % sc.nb_crystals
    pairs = np.stack([np.maximum(a, b), np.minimum(a, b)], 1)
    scat, _ = so.ref_scatter_fraction(pairs, pos, pts, em, at, mu, [192.0, 192.0, 160.0], 400, 650, 0.15,
                                      nbr, ncr, grid, size_block)
    ns = pts.shape[0]
    th = so.ref_block_theta(pos, ncr, nbr, grid[1])
    smu = so.ref_mu_at(pts, mu, [192.0, 192.0, 160.0])
    base = so.sss_sinogram(pos, th, pts, smu, em.reshape(-1, ns), -np.log(at.astype(np.float64)).reshape(-1, ns),
                           pairs, 400, 650, 0.15)
    sigma = math.sqrt(511.0) * 0.15
    e_ws = so.ref_eff_without_scatter(400, 650, sigma) * 5 / math.sqrt(1022 * math.pi) / 0.15
    area = size_block[1] * size_block[2] / grid[1] / grid[2]
    full = e_ws * area ** 2 * 5 / math.sqrt(1022 * math.pi) / 0.15 / 4 / math.pi
    C = so.sss_constant(400, 650, 0.15)
    # sigma = sqrt(511) * 0.15 keV leaves little but small-angle scatter in the
    # window: many pairs are orders of magnitude below the largest
    assert np.count_nonzero(scat > 1e-6 * scat.max()) > 50
    np.testing.assert_allclose(scat, base / C * full, rtol=1e-9, atol=1e-12 * scat.max())
    # symmetric in the pair
    scat_r, _ = so.ref_scatter_fraction(pairs[:, ::-1], pos, pts, em, at, mu, [192.0, 192.0, 160.0], 400, 650,
                                        0.15, nbr, ncr, grid, size_block)
    np.testing.assert_allclose(scat_r, scat, rtol=1e-12)


def test_scatter_positions_closed_form():
    from srf_amd.scatter_pairs import scatter_positions
    p = so.ref_scatter_positions((96, 96, 80), (192.0, 192.0, 160.0))
    # step floor(50 / 2) = 25 voxels, floor(96 / 25) = 3 points, first at 25 mm - 96 mm
    assert p.shape == (27, 3)
    assert sorted(set(p[:, 0].tolist())) == [-71.0, -21.0, 29.0]
    assert sorted(set(p[:, 2].tolist())) == [-55.0, -5.0, 45.0]
    assert np.array_equal(scatter_positions((96, 96, 80), (192.0, 192.0, 160.0)), p)
    # a 50 mm voxel edge: one point per voxel
    q = so.ref_scatter_positions((4, 4, 2), (200.0, 200.0, 100.0))
    assert q.shape == (32, 3) and q[0].tolist() == [-75.0, -75.0, -25.0]


def test_bin_pairs_in_get_all_lors_id_order():
    from srf_amd.data_transform import get_all_lors_id
    from srf_amd.scatter_pairs import pairs_of_bins
    n = 300
    want = get_all_lors_id(n)
    pos = np.arange(3 * n, dtype=np.float64).reshape(n, 3)
    out = so.ref_sino2lm(np.ones(n * (n - 1) // 2), pos)
    assert np.array_equal(out[:, 0], pos[want[:, 0], 0]) and np.array_equal(out[:, 3], pos[want[:, 1], 0])
    i, j = pairs_of_bins(torch.arange(n * (n - 1) // 2))
    assert np.array_equal(i.numpy(), want[:, 0]) and np.array_equal(j.numpy(), want[:, 1])
    # large bin indices (the mCT's 5.3e8 bins) decode exactly
    big = torch.tensor([0, 1, 2, 526_417_127, 526_417_128 - 1], dtype=torch.int64)
    i, j = pairs_of_bins(big)
    assert torch.all(i * (i - 1) // 2 + j == big) and torch.all((j >= 0) & (j < i))


def test_correct_closed_form():
    sino = np.array([0, 4, 0, 10, 6], np.float64)
    index = np.array([1, 3, 4])
    frac = np.array([1.0, 3.0, 0.0])
    scale = np.array([2.0, 1.0, 1.0])
    att = np.array([0.5, 1.0, 0.25], np.float32)
    out = so.ref_correct(sino, index, frac, scale, att, 0.3)
    # sum(sino[index]) = 20; subtract frac / 4 * 20 * 0.3 = 1.5 frac
    assert out.tolist() == [0.0, (4 - 1.5) * 2 / 0.5, 0.0, (10 - 4.5), 6 / 0.25]
