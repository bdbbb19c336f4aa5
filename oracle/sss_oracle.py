"""Single-scatter simulation (SSS) oracle -- TEST INFRASTRUCTURE ONLY.

Float64 numpy restatement of the reference's per-LOR single-scatter sum,
srfnef/corrections/scattering/scatter_fraction/scatter_fraction.py:70-131 and
kn_formula.py:6-46, evaluated on the coarse virtual-detector sinogram the
product kernel (srf_amd/csrc/srf_sss.hip) computes:

  scatter(A,B) = C * sum_S  g(S,A) g(S,B) eff(E') kn(theta_q) mu(S) / sigma_KN
                            * (Ia + Ib)
  g(S,D)  = ((D - S) . n_D) / |D - S|   /  |D - S|^2      (project_area / distance^2, :112-131, :89-90)
  theta_q = int(acos(-(|AS|^2 + |SB|^2 - |AB|^2) / (2|AS||SB|)) / pi * 180) / 180 * pi   (kn_formula.py:36-41)
  E'      = 511 / (2 - cos(theta_q))                        (:78-79)
  Ia      = att(S,A) * att(S,B)^(511/E') * em(S,A)          (:80-81)
  Ib      = att(S,A)^(511/E') * att(S,B) * em(S,B)          (:82-83)
  eff(E)  = sum_{E_i = low, low+5, .., < high} exp(-(E_i - E)^2 / (2 s^2)),  s = sqrt(511)*resolution  (:103-109)
  kn(c)   = 1/(2-c)^2 * ((1+c^2)/2 + (1-c)^2/2/(2-c)),  sigma_KN = pi (40/9 - 3 ln 3)          (kn_formula.py:11-20)
With reference_kn (default) kn is evaluated at c = theta_q, the quantised ANGLE
used as if it were a cosine (kn_formula.py:12,16: quirk Q16); physical_kn uses
cos(theta_q).  C = 5 / (sqrt(1022 pi) * resolution) * eff(511) / (4 pi)
(scatter_fraction.py:27-33, crystal-area factors dropped: the sinogram is
normalised before use).
"""
import math

import numpy as np

SIGMA_KN = math.pi * (40.0 / 9.0 - 3.0 * math.log(3.0))


def theta_quantised_deg(A, S, B):
    """Integer-degree scattering angle at S (kn_formula.py:36-41), float64."""
    A, S, B = (np.asarray(v, np.float64) for v in (A, S, B))
    d_ab = np.linalg.norm(A - B, axis=-1)
    d_as = np.linalg.norm(A - S, axis=-1)
    d_sb = np.linalg.norm(S - B, axis=-1)
    x = -(d_as ** 2 + d_sb ** 2 - d_ab ** 2) / (2 * d_as * d_sb)
    x = np.clip(x, -1.0, 1.0)
    return np.trunc(np.arccos(x) / math.pi * 180.0).astype(np.int64)


def kn(c):
    c = np.asarray(c, np.float64)
    return 1.0 / (2.0 - c) ** 2 * ((1.0 + c ** 2) / 2.0 + (1.0 - c) ** 2 / 2.0 / (2.0 - c))


def eff(E, low, high, resolution):
    s = math.sqrt(511.0) * resolution
    Ei = np.arange(int(low), int(high), 5, dtype=np.float64)
    E = np.asarray(E, np.float64)
    return np.sum(np.exp(-(Ei[None, :] - E.reshape(-1, 1)) ** 2 / 2.0 / s ** 2), axis=1).reshape(E.shape)


def energy_tables(low, high, resolution, physical_kn=False):
    """Per integer scattering angle k = 0..180: (511/E', eff(E'), kn/sigma_KN)."""
    k = np.arange(181, dtype=np.float64)
    th = k / 180.0 * math.pi
    c = np.cos(th)
    Ep = 511.0 / (2.0 - c)
    knv = kn(c if physical_kn else th) / SIGMA_KN
    return 511.0 / Ep, eff(Ep, low, high, resolution), knv


def sss_constant(low, high, resolution):
    return 5.0 / math.sqrt(1022.0 * math.pi) / resolution * float(eff(511.0, low, high, resolution)) / (4 * math.pi)


def sss_sinogram(det_pos, det_nrm, spos, smu, em, mu, bins, low, high, resolution, physical_kn=False):
    """Oracle sinogram.

    det_pos (nd,3) mm, det_nrm (nd,) normal angle (rad); spos (ns,3) scatter
    points, smu (ns,) attenuation at the point (1/mm); em, mu (nd,ns): emission
    and attenuation line integrals from each point to each detector; bins
    (nb,2) detector pairs.  Returns (nb,) float64."""
    det_pos = np.asarray(det_pos, np.float64)
    spos = np.asarray(spos, np.float64)
    em = np.asarray(em, np.float64)
    mu = np.asarray(mu, np.float64)
    smu = np.asarray(smu, np.float64)
    r_tab, e_tab, k_tab = energy_tables(low, high, resolution, physical_kn)
    C = sss_constant(low, high, resolution)
    out = np.zeros(len(bins), np.float64)
    keep = smu > 0
    S = spos[keep]
    sm = smu[keep]
    emk, muk = em[:, keep], mu[:, keep]
    for i, (a, b) in enumerate(np.asarray(bins)):
        A, B = det_pos[a], det_pos[b]
        deg = theta_quantised_deg(A[None, :], S, B[None, :])
        rr = r_tab[deg]

        def g(D, nrm):
            v = D[None, :] - S
            d = np.linalg.norm(v, axis=1)
            return (v[:, 0] * math.cos(nrm) + v[:, 1] * math.sin(nrm)) / d / d ** 2

        Ia = np.exp(-(muk[a] + muk[b] * rr)) * emk[a]
        Ib = np.exp(-(muk[a] * rr + muk[b])) * emk[b]
        out[i] = C * np.sum(g(A, det_nrm[a]) * g(B, det_nrm[b]) * e_tab[deg] * k_tab[deg] * sm * (Ia + Ib))
    return out
