"""Pins the SSS oracle (oracle/sss_oracle.py) with closed forms and invariants
of the reference formula (scatter_fraction.py:70-131, kn_formula.py:6-46)."""
import math

import numpy as np
import pytest

import sss_oracle as so


def test_theta_quantisation_closed_forms():
    # right angle at S: scattering angle 90 deg; collinear A-S-B: 0 deg
    assert so.theta_quantised_deg([-100, 0, 0], [0, 0, 0], [0, 100, 0]) == 90
    assert so.theta_quantised_deg([-100, 0, 0], [10, 0, 0], [100, 0, 0]) == 0
    # truncation to whole degrees (int() in kn_formula.py:41): 30.7 deg -> 30
    t = math.radians(180 - 30.7)
    B = [100 * math.cos(t), 100 * math.sin(t), 0]
    assert so.theta_quantised_deg([100, 0, 0], [0, 0, 0], B) == 30


def test_energy_tables_closed_forms():
    r, eff, kn = so.energy_tables(400, 650, 0.15)
    # E' = 511 / (2 - cos theta): theta 0 -> 511 keV, 60 -> 340.67, 180 -> 170.33
    assert r[0] == pytest.approx(1.0) and r[60] == pytest.approx(1.5) and r[180] == pytest.approx(3.0)
    # eff at 511 keV is the unscattered efficiency (eff_without_scatter, :94-100)
    assert eff[0] == pytest.approx(float(so.eff(511.0, 400, 650, 0.15)))
    # reference KN is evaluated at the angle in radians as if it were a cosine (Q16):
    # theta 0 -> c = 0 -> (1/4)(1/2 + 1/4) = 0.1875; physical c = cos 0 = 1 -> 1
    assert kn[0] * so.SIGMA_KN == pytest.approx(0.1875)
    _, _, knp = so.energy_tables(400, 650, 0.15, physical_kn=True)
    assert knp[0] * so.SIGMA_KN == pytest.approx(1.0)
    # eff falls off as E' leaves the 400-650 keV window
    assert eff[0] > eff[60] > eff[120]


def _toy(seed=0):
    rng = np.random.default_rng(seed)
    nd = 16
    phi = 2 * math.pi * np.arange(nd) / nd
    det = np.stack([300 * np.cos(phi), 300 * np.sin(phi), np.zeros(nd)], 1)
    spos = rng.uniform(-80, 80, (40, 3)) * np.array([1, 1, 0.3])
    smu = rng.uniform(0.005, 0.02, 40)
    em = rng.uniform(1, 10, (nd, 40))
    mu = rng.uniform(0.1, 2.0, (nd, 40))
    bins = np.array([[i, j] for i in range(nd) for j in range(nd) if i != j])
    return det, phi, spos, smu, em, mu, bins


def test_sss_invariants():
    det, phi, spos, smu, em, mu, bins = _toy()
    s = so.sss_sinogram(det, phi, spos, smu, em, mu, bins, 400, 650, 0.15)
    # A <-> B symmetry
    sw = so.sss_sinogram(det, phi, spos, smu, em, mu, bins[:, ::-1], 400, 650, 0.15)
    np.testing.assert_allclose(s, sw, rtol=1e-12)
    # linear in the emission integrals, proportional to mu at the scatter points
    np.testing.assert_allclose(so.sss_sinogram(det, phi, spos, smu, 2 * em, mu, bins, 400, 650, 0.15), 2 * s,
                               rtol=1e-12)
    np.testing.assert_allclose(so.sss_sinogram(det, phi, spos, 3 * smu, em, mu, bins, 400, 650, 0.15), 3 * s,
                               rtol=1e-12)
    # no attenuating medium -> no scatter
    assert not so.sss_sinogram(det, phi, spos, 0 * smu, em, mu, bins, 400, 650, 0.15).any()
    # more attenuation on the paths -> less scatter reaches the detectors (with the
    # physical KN factor; the reference's angle-as-cosine factor (Q16) turns
    # negative past 2 rad, so its sum is not monotone)
    sp = so.sss_sinogram(det, phi, spos, smu, em, mu, bins, 400, 650, 0.15, physical_kn=True)
    assert np.all(sp >= 0) and np.mean(sp > 0) > 0.5
    more = so.sss_sinogram(det, phi, spos, smu, em, mu + 1.0, bins, 400, 650, 0.15, physical_kn=True)
    assert np.all(more[sp > 0] < sp[sp > 0])


def test_single_point_analytic():
    # one scatter point at the origin, detectors on the x and y axes: 90 deg
    det = np.array([[300.0, 0, 0], [0, 300.0, 0]])
    phi = np.array([0.0, math.pi / 2])
    spos = np.zeros((1, 3))
    em = np.array([[2.0], [3.0]])
    mu = np.array([[0.4], [0.7]])
    s = so.sss_sinogram(det, phi, spos, np.array([0.01]), em, mu, np.array([[0, 1]]), 400, 650, 0.15)[0]
    th = math.pi / 2
    Ep = 511 / (2 - math.cos(th))
    rr = 511 / Ep
    eff = sum(math.exp(-(E - Ep) ** 2 / 2 / (math.sqrt(511) * 0.15) ** 2) for E in range(400, 650, 5))
    c = th
    kn = 1 / (2 - c) ** 2 * ((1 + c * c) / 2 + (1 - c) ** 2 / 2 / (2 - c))
    g = 1.0 / 300.0 ** 2    # cos = 1 for radial incidence
    Ia = math.exp(-(0.4 + 0.7 * rr)) * 2.0
    Ib = math.exp(-(0.4 * rr + 0.7)) * 3.0
    want = so.sss_constant(400, 650, 0.15) * g * g * eff * kn / so.SIGMA_KN * 0.01 * (Ia + Ib)
    assert s == pytest.approx(want, rel=1e-12)
