"""Cross-check of the reference's unwired TF scatter op
(srfnef/cc/scatter/scatter.cu.cc:18-58, 121-135; nothing in srfnef calls it)
against the formulas this build follows, the numba ScatterCorrect path
(scatter_fraction/scatter_fraction.py, kn_formula.py) as restated by
oracle/sss_oracle.py and the SSS kernel's tables (srf_amd.sss.energy_tables).
The op's device functions are restated here in float64 numpy, only to compare.

What agrees:
  * window efficiency: the op's sum over the 5 keV grid divided by
    sqrt(2 pi) sigma is the numba sum times 5/sqrt(1022 pi)/resolution over 5
    (the same Gaussian density, without the 5 keV step), sigma = sqrt(511) res;
  * the Klein-Nishina polynomial 1/(2-c)^2 ((1+c^2)/2 + (1-c)^2/2/(2-c)) is
    the same function;
  * a crystal's block normal: round(atan2(y, x) / block angle) equals
    get_block_theta on the central plane; the obliquity factor is the same up
    to the op's abs() and crystal area.
What differs:
  * the op's sigma_KN carries Z = 10: ten times the numba one;
  * the op evaluates KN and E' at c = cos(angle A-S-B) = -cos(theta): E' =
    511/(2 + cos theta), 511/3 keV for unscattered photons, where the numba
    path uses cos(theta_q) for E' and theta_q itself for KN (quirk Q16) and
    the physical tables cos(theta_q) for both;
  * the op attenuates both legs by exp(-mu_a - mu_b) with no 511/E' scaling of
    the scattered leg;
  * off the central plane get_block_theta's 3-D angle (quirk Q23) picks other
    blocks than the op's atan2;
  * the op's pair scale is 4 pi L^2 / eps^2 with the obliquities commented out
    (:131-134); the numba one is (L / area)^2 4 pi / eff_ws^2 / crystal_area^2.
"""
import math

import numpy as np
import pytest

import sss_oracle as so

LOW, HIGH, RES = 400, 650, 0.15
SIGMA = math.sqrt(511.0) * RES


def tf_efficiency(low, high, res_eng, energy):
    """scatter.cu.cc:18-27."""
    en = np.arange(low, high, 5.0)
    return np.sum(np.exp(-(en - energy) ** 2 / 2 / res_eng / res_eng)) / math.sqrt(2 * math.pi) / res_eng


def tf_differential(c):
    """scatter.cu.cc:55 (with its sigma at :53-54)."""
    c = np.asarray(c, np.float64)
    return 1 / (2 - c) / (2 - c) * ((1.0 + c * c) / 2 + (1.0 - c) * (1.0 - c) / 2 / (2 - c))


TF_SIGMA = math.pi * 10 * (40.0 / 9.0 - 3.0 * math.log(3.0))


def tf_block_normal(p, angle_per_block):
    """scatter.cu.cc:34."""
    return np.round(np.arctan2(p[..., 1], p[..., 0]) / angle_per_block) * angle_per_block


def tf_project_area(angle_per_block, crystal_area, inner, crystal):
    """scatter.cu.cc:30-37: |(inner - crystal) . n| / |inner - crystal| * area."""
    th = tf_block_normal(crystal, angle_per_block)
    v = inner - crystal
    return np.abs(v[..., 0] * np.cos(th) + v[..., 1] * np.sin(th)) / np.linalg.norm(v, axis=-1) * crystal_area


def tf_cos_and_energy(A, P, B):
    """scatter.cu.cc:95-96, 102: cos of the angle A-P-B and the op's E'."""
    u, v = A - P, B - P
    c = np.sum(u * v, -1) / np.linalg.norm(u, axis=-1) / np.linalg.norm(v, axis=-1)
    return c, 511 / (2 - c)


def test_window_efficiency_agrees_up_to_the_5_kev_step():
    r, e, _ = so.energy_tables(LOW, HIGH, RES)
    Ep = 511.0 / r
    tf = np.array([tf_efficiency(LOW, HIGH, SIGMA, E) for E in Ep])
    numba_normalised = e * 5 / math.sqrt(1022 * math.pi) / RES
    np.testing.assert_allclose(tf * 5, numba_normalised, rtol=1e-12, atol=1e-300)


def test_kn_polynomial_same_sigma_ten_times():
    c = np.linspace(-1, 1, 401)
    np.testing.assert_allclose(tf_differential(c), so.kn(c), rtol=1e-14)
    assert TF_SIGMA == pytest.approx(10 * so.SIGMA_KN, rel=1e-15)


def test_op_uses_the_supplementary_cosine():
    """Scatter point S on a circle, A and B fixed: the op's cosine is -cos(theta)
    of the scattering angle theta; its E' is 511/(2 + cos theta), 511/3 keV
    at theta = 0, where the tables give 511 keV."""
    A = np.array([-300.0, 0.0, 0.0])
    B = np.array([300.0, 40.0, 0.0])
    phi = np.linspace(0.05, 2 * np.pi - 0.05, 50)
    S = np.stack([60 * np.cos(phi), 60 * np.sin(phi), np.zeros_like(phi)], 1)
    c_op, e_op = tf_cos_and_energy(A[None], S, B[None])
    u, v = S - A[None], B[None] - S
    cos_theta = np.sum(u * v, 1) / np.linalg.norm(u, axis=1) / np.linalg.norm(v, axis=1)
    np.testing.assert_allclose(c_op, -cos_theta, atol=1e-13)
    np.testing.assert_allclose(e_op, 511 / (2 + cos_theta), rtol=1e-13)
    # unscattered (S on the segment): the op 511/3, the numba path and the tables 511
    c0, e0 = tf_cos_and_energy(A, np.array([0.0, 20.0, 0.0]), B)
    deg = so.theta_quantised_deg(A[None], np.array([[0.0, 20.0, 0.0]]), B[None])
    r, _, _ = so.energy_tables(LOW, HIGH, RES)
    assert c0 == pytest.approx(-1.0) and e0 == pytest.approx(511 / 3)
    assert deg[0] == 0 and 511 / r[deg[0]] == pytest.approx(511.0)


def test_block_normals_and_obliquity():
    from srf_amd import scanner as S
    from srf_amd.scatter_pairs import block_normals, block_params
    flat = S.PetEcatScanner(424.5, 444.5, 1, 48, 0.0, S.Block([20.0, 53.3, 10.0], [1, 13, 1]))
    for sc, same in ((flat, True), (S.scanner_mct, False)):
        pos = S.crystal_positions(sc).astype(np.float64)
        nbr, ncr, grid, size_block = block_params(sc)
        numba = block_normals(pos, ncr, nbr, grid[1])
        op = np.mod(tf_block_normal(pos, sc.angle_per_block), 2 * np.pi)
        agree = np.isclose(np.mod(numba - op + np.pi, 2 * np.pi) - np.pi, 0.0, atol=1e-9)
        assert agree.all() == same
        if not same:
            # off the central plane only (quirk Q23)
            assert np.all(agree[np.abs(pos[:, 2]) < 1e-3]) and not agree.all()
    # obliquity: the op's factor is |numba's| times the crystal area
    pos = S.crystal_positions(flat).astype(np.float64)
    pts = np.random.default_rng(1).uniform(-150, 150, (len(pos), 3))
    nbr, ncr, grid, size_block = block_params(flat)
    th = block_normals(pos, ncr, nbr, grid[1])
    area = size_block[1] * size_block[2] / grid[1] / grid[2]
    numba = so.ref_project_area(pts, pos, th)
    np.testing.assert_allclose(tf_project_area(flat.angle_per_block, area, pts, pos), np.abs(numba) * area,
                               rtol=1e-12)


def test_attenuation_and_scale_differ():
    """Both legs exp(-mu_a - mu_b) in the op against att_a att_b^(511/E') in the
    numba path: equal at theta = 0 only.  The op's scale 4 pi L^2 / eps^2 equals
    the numba one where the obliquity is 1 and eps = eff_ws * crystal area."""
    r, _, _ = so.energy_tables(LOW, HIGH, RES)
    mu_a, mu_b = 0.7, 1.3
    numba = np.exp(-mu_a) * np.exp(-mu_b * r)
    op = np.exp(-mu_a - mu_b) * np.ones_like(r)
    assert numba[0] == pytest.approx(op[0]) and np.all(numba[1:] < op[1:])
    # two facing block centres (obliquity 1), L = 880 mm, crystals 10 x 10 mm
    pos = np.array([[440.0, 0.0, 0.0], [-440.0, 0.0, 0.0]], np.float32)
    _, scale = so.ref_scatter_fraction([(0, 1)], pos, np.zeros((1, 3), np.float32), np.ones(2), np.ones(2),
                                       np.zeros((2, 2, 2), np.float32), [10.0] * 3, LOW, HIGH, RES, 2, 2, [1, 1, 1],
                                       [20.0, 10.0, 10.0])
    e_ws = so.ref_eff_without_scatter(LOW, HIGH, SIGMA) * 5 / math.sqrt(1022 * math.pi) / RES
    eps = e_ws * 100.0
    op_scale = 4.0 * math.pi * 880.0 * 880.0 / eps / eps          # scatter.cu.cc:134
    assert op_scale == pytest.approx(scale[0], rel=1e-12)
