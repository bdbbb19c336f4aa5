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
cos(theta_q).  C = 5 / (sqrt(1022 pi) * resolution) * eff(511) / (4 pi): the
reference's constant (scatter_fraction.py:27-33) is eff_ws * area^2 * 5 /
(sqrt(1022 pi) * resolution) / (4 pi) with eff_ws = eff(511) * 5 /
(sqrt(1022 pi) * resolution), so C drops one 5 / (sqrt(1022 pi) * resolution)
and the crystal area squared, which cancel where the sinogram is normalised
(ref_scatter_fraction below applies the full constant).
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


# ------------------------------------------------- crystal-pair ScatterCorrect --
# The reference's own scatter correction over crystal pairs, restated in
# float64 numpy (projections by the C oracle, oracle.project):
#   scatter_preprocess   srfnef/corrections/scattering/scatter.py:57-73
#   scatter positions    preprocess.py:6-18, 129-133
#   line integrals       preprocess.py:21-43 (Project * |p2 - p1|^2; UmapProject;
#                        UmapProject of 1.3 mu along the sinogram's crystal pairs)
#   per-LOR scatter      scatter_fraction.py:12-37 (constants), :40-61 (per LOR),
#                        :64-67 get_scale, :70-91 loop_all_s, :94-109 efficiencies,
#                        :112-131 project_area / get_block_theta, kn_formula.py:6-20
#   correction           scatter.py:51-55, listmode2sinogram.py:17-35, 88-99

def ref_scatter_positions(shape, size):
    """preprocess.py:6-18: step = floor(50 / voxel), floor(shape / step) points
    per axis at step (k + 1/2) voxel - size / 2 (the image centre is not added)."""
    vox = np.asarray(size, np.float64) / np.asarray(shape, np.float64)
    step = np.floor(50.0 / vox)
    cnt = np.floor(np.asarray(shape, np.float64) / step).astype(int)
    g = np.meshgrid(*[np.arange(c, dtype=np.float64) for c in cnt], indexing="ij")
    pos = np.stack([step[i] * (g[i] + 0.5) * vox[i] for i in range(3)], -1).reshape(-1, 3) \
        - np.asarray(size, np.float64) / 2
    return pos.astype(np.float32)


def ref_block_theta(p, nb_crystals_per_ring, nb_blocks_per_ring, grid_block_y):
    """scatter_fraction.py:122-131: normal angle of the block holding point p."""
    p = np.asarray(p, np.float64)
    norm = np.sqrt(np.sum(p ** 2, axis=-1))
    th = np.arccos(p[..., 0] / norm) / math.pi * 180.0
    th = np.where(p[..., 1] < 0, 360.0 - th, th)
    fixed = np.mod(th + 180.0 / nb_crystals_per_ring * grid_block_y, 360.0)
    return np.floor(fixed / 360.0 * nb_blocks_per_ring) / nb_blocks_per_ring * 2 * math.pi


def ref_project_area(pa, pb, theta_b):
    """scatter_fraction.py:112-119: ((pb - pa) . n_b) / |pb - pa| (no abs)."""
    pa = np.asarray(pa, np.float64)
    pb = np.asarray(pb, np.float64)
    v = pb - pa
    return (v[..., 0] * np.cos(theta_b) + v[..., 1] * np.sin(theta_b)) / np.sqrt(np.sum(v ** 2, axis=-1))


def ref_mu_at(S, u_map, size):
    """kn_formula.py:12-14,20: u_map[int((S / size + 0.5) * shape)] (a zero-centred
    image, quirk Q16)."""
    u = np.asarray(u_map)
    S = np.asarray(S, np.float64)
    idx = [np.trunc((S[:, k] / size[k] + 0.5) * u.shape[k]).astype(np.int64) for k in range(3)]
    return u[idx[0], idx[1], idx[2]].astype(np.float64)


def ref_eff_without_scatter(low, high, sigma):
    return float(eff(511.0, low, high, sigma / math.sqrt(511.0)))


def ref_scatter_fraction(pairs, crystal_pos, scatter_pos, sumup_emission, atten, u_map, size, low, high,
                         resolution, nb_blocks_per_ring, nb_crystals_per_ring, grid_block, size_block):
    """scatter_fraction.py:12-37 (with loop_all_lors / loop_all_s / get_scale):
    (scatter, scale) per crystal pair (a, b) of `pairs`, float64."""
    crystal_pos = np.asarray(crystal_pos, np.float64)
    S = np.asarray(scatter_pos, np.float64)
    ns = S.shape[0]
    em = np.asarray(sumup_emission, np.float64).reshape(-1, ns)
    at = np.asarray(atten, np.float64).reshape(-1, ns)
    th = ref_block_theta(crystal_pos, nb_crystals_per_ring, nb_blocks_per_ring, grid_block[1])
    muS = ref_mu_at(S, u_map, size)
    r_tab, e_tab, k_tab = energy_tables(low, high, resolution, physical_kn=False)
    scatter = np.zeros(len(pairs))
    scale = np.zeros(len(pairs))
    for i, (a, b) in enumerate(np.asarray(pairs)):
        A, B = crystal_pos[a], crystal_pos[b]
        deg = theta_quantised_deg(A[None, :], S, B[None, :])
        rr = r_tab[deg]      # 511 / E'
        Ia = at[a] * np.exp(np.log(at[b]) * rr) * em[a]
        Ib = np.exp(np.log(at[a]) * rr) * at[b] * em[b]
        gA = ref_project_area(S, A[None, :], th[a]) / np.sum((A[None, :] - S) ** 2, axis=1)
        gB = ref_project_area(S, B[None, :], th[b]) / np.sum((B[None, :] - S) ** 2, axis=1)
        with np.errstate(invalid="ignore"):
            term = gA * e_tab[deg] * gB * k_tab[deg] * muS * (Ia + Ib)
        scatter[i] = np.sum(np.where(muS != 0, term, 0.0))
        area_ab = ref_project_area(A, B, th[b])
        scale[i] = (np.sqrt(np.sum((A - B) ** 2)) / area_ab) ** 2
    sigma = math.sqrt(511.0) * resolution
    e_ws = ref_eff_without_scatter(low, high, sigma) * 5.0 / math.sqrt(1022.0 * math.pi) / resolution
    area = size_block[1] * size_block[2] / grid_block[1] / grid_block[2]
    scatter = scatter * e_ws * area ** 2 * 5.0 / math.sqrt(1022.0 * math.pi) / resolution / 4 / math.pi
    scale = scale * 4 * math.pi / e_ws ** 2 / area ** 2
    return scatter, scale


def ref_correct(sinogram, index, fraction, scale, atten_lors, scatter_fraction):
    """scatter.py:51-53 on the bins of `index` (a float32 atten_lors plus
    sys.float_info.min stays atten_lors in float32)."""
    s = np.asarray(sinogram, np.float64).reshape(-1)
    f = np.asarray(fraction, np.float64)
    out = np.zeros_like(s)
    with np.errstate(divide="ignore", invalid="ignore"):
        out[index] = (s[index] - f / np.sum(f) * np.sum(s[index]) * scatter_fraction) * np.asarray(
            scale, np.float64) / np.asarray(atten_lors, np.float32).astype(np.float64)
    return out


def ref_sino2lm(corrected, crystal_pos):
    """listmode2sinogram.py:88-99 with get_all_lors_id's pair of a bin:
    (M, 7) float64 [crystal i, crystal j, value] for 0 < value < 1e8."""
    c = np.asarray(corrected, np.float64).reshape(-1)
    idx = np.where((c > 0) & (c < 1e8))[0]
    i = np.floor((1 + np.sqrt(1 + 8 * idx.astype(np.float64))) / 2).astype(np.int64)
    i = np.where(i * (i - 1) // 2 > idx, i - 1, i)
    i = np.where((i + 1) * i // 2 <= idx, i + 1, i)
    j = idx - i * (i - 1) // 2
    pos = np.asarray(crystal_pos, np.float64)
    return np.hstack((pos[i], pos[j], c[idx][:, None]))
