"""Pins the CPU oracle (oracle/srf_oracle.c) against mathematics that does not
depend on it: exact Siddon line integrals, closed forms, invariants and the
reference's quirks (file:line of srfnef/cc/siddon*/ per test).  A plausible
slip in the restatement -- a dropped 1/L^2, a wrong branch index, a swapped
predicate, an off-by-half TOF centre -- fails at least one of these."""
import numpy as np
import pytest

from siddon_ref import chord_in_box, random_cylinder_lors, siddon_line_integral

FP, BP = 0, 1


def _exact_rays(lors, shape, center, size, same_sign=True):
    """Rays on which the reference slab marcher equals exact Siddon: at most one
    secondary-plane crossing per slab, endpoints outside the box, non-TOF cut
    (transaxial >= |zd|) inactive, dominant axis unambiguous, and the secondary
    and axial slopes of one sign (see test_quirk_q22_opposite_slopes)."""
    d = np.asarray(size) / np.asarray(shape)
    keep = []
    for l in lors:
        D = (l[3:6] - l[0:3]).astype(np.float64) / d[0]
        ax, ay, az = abs(D[0]), abs(D[1]), abs(D[2])
        prim = max(ax, ay)
        if abs(ax - ay) < 1e-3 * prim:
            continue
        dz = d[2] / d[0]
        if az / prim >= 0.95 * dz:
            continue
        if np.hypot(D[0], D[1]) < az or np.hypot(D[0], D[1]) < 10:
            continue
        sec = D[1] if ax > ay else D[0]
        if same_sign is not None and (sec * D[2] >= 0) != same_sign:
            continue
        keep.append(l)
    return np.array(keep, np.float32)


@pytest.mark.parametrize("tof", [False, True])
def test_projector_equals_exact_siddon(orc, tof):
    # siddon.cu.cc:86-189 weights r*sqrt(1+k^2+kz^2)*dx_/L/L: for shallow rays this
    # is the exact chord length, so p = integral(img) * dx_^2 / L_mm^2.
    rng = np.random.default_rng(1)
    shape, vox = (24, 20, 12), 2.5
    size = [shape[0] * vox, shape[1] * vox, shape[2] * vox]
    center = [1.5, -2.0, 0.75]
    img = rng.uniform(0.5, 2.0, shape).astype(np.float32)
    lors = _exact_rays(random_cylinder_lors(600, 80.0, 40.0, 7), shape, center, size)[:150]
    assert len(lors) > 100
    if tof:
        # a TOF kernel with a Gaussian far wider than the box is the non-TOF
        # integral times tof_bin_ * N(0; sigma_e) (siddon_tof.cu.cc:56-59)
        tof_bin, sig2 = 3.66, 1e9
        l7 = np.concatenate([lors, np.zeros((len(lors), 1), np.float32)], 1)
        p = orc.project(img, l7, center, size, tof=(tof_bin, sig2), f64=True)
        dx_ = size[0] / shape[0]
        tb_, ts_ = tof_bin / dx_, sig2 / dx_ / dx_
        tse = ts_ + tb_ * tb_ / 12
        scale = tb_ / np.sqrt(2 * np.pi * tse)
    else:
        p = orc.project(img, lors, center, size, f64=True)
        scale = 1.0
    dx_ = size[0] / shape[0]
    for l, pe in zip(lors, p):
        L = np.linalg.norm(l[3:6].astype(np.float64) - l[0:3])
        want = siddon_line_integral(img, center, size, l[0:3], l[3:6]) * dx_ * dx_ / (L * L) * scale
        assert pe == pytest.approx(want, rel=2e-5, abs=1e-12)


def test_backprojector_conservation_and_missing_inv_L2(orc):
    # BP weight has no 1/L^2 (Q1, siddon.cu.cc:418): sum_v BP(1)[v] = sum_e chord_e / dx_ * dx_ ...
    # i.e. the total is the summed chord length in mm on exact rays.
    shape, vox = (20, 20, 10), 3.0
    size = [shape[0] * vox, shape[1] * vox, shape[2] * vox]
    center = [0.0, 0.0, 0.0]
    lors = _exact_rays(random_cylinder_lors(400, 70.0, 30.0, 3), shape, center, size)[:200]
    img = orc.backproject(np.ones(len(lors), np.float32), lors, shape, center, size, variant=BP)
    want = sum(chord_in_box(center, size, l[0:3], l[3:6]) for l in lors)
    assert img.sum() == pytest.approx(want, rel=1e-6)


def test_adjoint_relation_between_reference_fp_and_bp(orc):
    # <A x, y> = <x, B(y / L_vox^2)> where the FP cuts are inactive and the
    # dominant-axis tests agree (Q1, Q3): ties the two reference code paths.
    rng = np.random.default_rng(5)
    shape, vox = (16, 18, 10), 2.0
    size = [shape[0] * vox, shape[1] * vox, shape[2] * vox]
    center = [0.0, 1.0, -0.5]
    lors = _exact_rays(random_cylinder_lors(600, 60.0, 25.0, 11), shape, center, size)
    x = rng.uniform(0, 1, shape).astype(np.float32)
    y = rng.uniform(0, 1, len(lors)).astype(np.float32)
    Ax = orc.project(x, lors, center, size, f64=True)
    Lv2 = np.sum(((lors[:, 3:6] - lors[:, 0:3]).astype(np.float64) / (size[0] / shape[0])) ** 2, 1)
    Bty = orc.backproject((y / Lv2).astype(np.float32), lors, shape, center, size, variant=BP)
    lhs = float(np.dot(Ax, y))
    rhs = float(np.sum(x.astype(np.float64) * Bty))
    assert lhs == pytest.approx(rhs, rel=1e-5)


def test_axis_aligned_lor_terms(orc):
    # A LOR along x through voxel centres: one term per slab, weight dx_/L^2,
    # voxel index ix + iy*nx + iz*nx*ny (siddon.cu.cc:94-95).
    shape, vox = (10, 8, 6), 2.0
    size = [20.0, 16.0, 12.0]
    center = [0.0, 0.0, 0.0]
    iy, iz = 3, 4
    y = -8.0 + (iy + 0.5) * 2.0
    z = -6.0 + (iz + 0.5) * 2.0
    lor = np.array([-50.0, y, z, 50.0, y, z], np.float32)
    vox_idx, slot, w, _ = orc.terms(lor, shape, center, size, variant=FP)
    assert list(vox_idx) == [ix + iy * 10 + iz * 80 for ix in range(10)]
    assert set(slot) == {0}
    L = 100.0 / 2.0
    np.testing.assert_allclose(w, 2.0 / L / L, rtol=1e-7)


def test_tof_closed_forms(orc):
    # uniform image, Gaussian well inside the box: sum_slabs w_tof * sq -> tof_bin_
    # (Riemann sum of a normalised Gaussian), so p_tof / p_nontof = tof_bin / chord_mm.
    shape, vox = (120, 120, 10), 2.0
    size = [240.0, 240.0, 20.0]
    center = [0.0, 0.0, 0.0]
    img = np.ones(shape, np.float32)
    lor = np.array([[-300.0, 10.0, 0.3, 300.0, -20.0, -0.4, 0.0]], np.float32)
    tof_bin, sig2 = 3.66, 400.0
    p_tof = orc.project(img, lor, center, size, tof=(tof_bin, sig2), f64=True)[0]
    p_non = orc.project(img, lor[:, :6], center, size, f64=True)[0]
    chord = chord_in_box(center, size, lor[0, 0:3], lor[0, 3:6])
    assert p_tof / p_non == pytest.approx(tof_bin / chord, rel=1e-5)
    # the peak sits at p1 + ratio1*(p2-p1): move it with the TOF value (Q21)
    for tofv, expect_x in ((0.0, 0.0), (120.0, -60.0), (-120.0, 60.0)):
        l = lor.copy()
        l[0, 6] = tofv
        vox_idx, _, w, wt = orc.terms(l[0], shape, center, size, tof=(tof_bin, sig2), variant=FP)
        ix = vox_idx % 120
        xc = -120.0 + (ix + 0.5) * 2.0
        peak = np.sum(xc * wt * w) / np.sum(wt * w)
        # expected x of the TOF centre; LOR x-extent 600 mm, tof shifts by tof/2 along the LOR
        assert peak == pytest.approx(expect_x * (600.0 / np.linalg.norm(l[0, 3:6] - l[0, 0:3])), abs=0.25)


def test_tof_gaussian_is_symmetric_in_d(orc):
    shape, vox = (64, 64, 8), 2.0
    size = [128.0, 128.0, 16.0]
    lor = np.array([-200.0, 0.0, 0.0, 200.0, 0.0, 0.0, 0.0], np.float32)
    _, _, _, wt = orc.terms(lor, shape, [0, 0, 0], size, tof=(3.66, 300.0), variant=BP)
    np.testing.assert_allclose(wt, wt[::-1], rtol=1e-5)  # float d rounding only


def test_quirk_q8_odd_nx_shifts_tof_centre(orc):
    # siddon_tof.cu.cc:55: xc = ix - nx/2 + 0.5 with integer nx/2; for odd nx
    # the Gaussian is evaluated half a voxel off the voxel centres.
    size = [5.0 * 2, 8.0, 8.0]
    lor = np.array([-100.0, 0.0, 0.0, 100.0, 0.0, 0.0, 0.0], np.float32)
    _, _, _, wt = orc.terms(lor, (5, 4, 4), [0, 0, 0], size, tof=(3.66, 4.0), variant=BP)
    # xc = -1.5, -0.5, 0.5, 1.5, 2.5 (not -2..2): symmetric about ix = 1.5
    assert wt[1] == pytest.approx(wt[2], rel=1e-6)
    assert wt[0] == pytest.approx(wt[3], rel=1e-6)
    _, _, _, wt6 = orc.terms(lor, (6, 4, 4), [0, 0, 0], [12.0, 8, 8], tof=(3.66, 4.0), variant=BP)
    assert wt6[2] == pytest.approx(wt6[3], rel=1e-6)


def test_quirk_q5_short_transaxial_skipped(orc):
    # siddon.cu.cc:34 / :364: sqrt(xd^2+yd^2) < 10 voxels -> no terms, FP and BP
    lor = np.array([0.0, 0.0, -50.0, 9.0, 0.0, 50.0], np.float32)
    for v in (FP, BP):
        assert len(orc.terms(lor, (16, 16, 16), [0, 0, 0], [16, 16, 16], variant=v)[0]) == 0


def test_quirk_q2_nontof_fp_axial_cut(orc):
    # siddon.cu.cc:38: non-TOF projector returns when transaxial < |zd|; BP and TOF do not.
    lor = np.array([-12.0, 0.0, -20.0, 12.0, 0.0, 20.0, 0.0], np.float32)
    g = ((16, 16, 16), [0, 0, 0], [16.0, 16.0, 16.0])
    assert len(orc.terms(lor[:6], *g, variant=FP)[0]) == 0
    assert len(orc.terms(lor[:6], *g, variant=BP)[0]) > 0
    assert len(orc.terms(lor, *g, tof=(3.66, 100.0), variant=FP)[0]) > 0


def test_quirk_q3_tie_breaks(orc):
    # |xd| == |yd|: non-TOF FP takes the x branch (>=, siddon.cu.cc:46), non-TOF BP
    # (siddon.cu.cc:372) and TOF (siddon_tof.cu.cc:48) take the y branch.  On the
    # diagonal y = x + 0.25 both branches credit the same cells, but the x branch
    # reports the diagonal cells as its first (r = 0.75) term and the y branch as
    # its second: the branch ids tell them apart.
    lor = np.array([-40.0, -39.75, 0.5, 40.0, 40.25, 0.5, 0.0], np.float32)
    g = ((16, 16, 4), [0, 0, 0], [16.0, 16.0, 4.0])

    def diagonal_branch(t):
        vox, slot, w, _ = t
        x, y = vox % 16, (vox // 16) % 16
        d = slot[x == y]
        assert len(set(d.tolist())) == 1
        return int(d[0])

    assert diagonal_branch(orc.terms(lor[:6], *g, variant=FP)) == 3
    assert diagonal_branch(orc.terms(lor[:6], *g, variant=BP)) == 4
    assert diagonal_branch(orc.terms(lor, *g, tof=(3.66, 100.0), variant=FP)) == 4
    assert diagonal_branch(orc.terms(lor, *g, tof=(3.66, 100.0), variant=BP)) == 4


def test_quirk_q4_segment_cuts(orc):
    # endpoints inside the image: non-TOF FP drops slabs beyond p2
    # (siddon.cu.cc:56-57); TOF FP drops slabs at/before p1 (siddon_tof.cu.cc:67-74);
    # BP keeps every slab.
    g = ((20, 8, 8), [0, 0, 0], [20.0, 8.0, 8.0])
    lor = np.array([-4.6, 0.5, 0.5, 5.4, 0.5, 0.5, 0.0], np.float32)    # x1=-4.6, x2=5.4 (inside)
    ix = lambda v: sorted(set((v % 20).tolist()))
    fp = ix(orc.terms(lor[:6], *g, variant=FP)[0])
    bp = ix(orc.terms(lor[:6], *g, variant=BP)[0])
    tfp = ix(orc.terms(lor, *g, tof=(3.66, 100.0), variant=FP)[0])
    assert bp == list(range(20))
    # non-TOF: keep slabs with xx2 = ix-10+1 <= 5.4  -> ix <= 14
    assert fp == list(range(15))
    # TOF: drop slabs with xx1 = ix-10 <= -4.6  -> keep ix >= 6
    assert tfp == list(range(6, 20))


def test_quirk_q6_steep_rays_skip_middle_cells(orc):
    # at most one crossing per secondary axis per slab (siddon.cu.cc:86-189): a ray
    # crossing several z planes in one slab only credits the first and last cell.
    lor = np.array([-30.0, 0.2, -25.0, 30.0, 0.2, 25.0], np.float32)   # kz = 5/6 per mm ... dz 0.25
    g = ((30, 4, 200), [0, 0, 0], [30.0, 4.0, 50.0])                    # dz = 0.25 voxel units
    vox, slot, w, _ = orc.terms(lor, *g, variant=BP)
    per_slab = np.bincount(vox % 30, minlength=30)
    assert per_slab.max() <= 3
    # weights of a slab sum to sqrt(1+k^2+kz^2)*dx_ only when the ray stays in two z cells
    assert np.all(np.isin(slot, [1, 2, 0]))


def test_bp_of_zero_is_zero(orc):
    # tests/ops/test_backproject_mixin.py:23-24 of the reference
    lors = random_cylinder_lors(100, 50.0, 20.0, 1)
    img = orc.backproject(np.zeros(100, np.float32), lors, (16, 16, 8), [0, 0, 0], [32.0, 32.0, 16.0])
    assert img.shape == (16, 16, 8) and not img.any()


def test_projection_length_matches_lors(orc):
    # tests/ops/test_project_mixin.py:21 of the reference
    lors = random_cylinder_lors(77, 50.0, 20.0, 2)
    p = orc.project(np.ones((16, 16, 8), np.float32), lors, [0, 0, 0], [32.0, 32.0, 16.0])
    assert p.shape == (77,)


def test_digest_is_order_independent_and_sensitive(orc):
    lors = random_cylinder_lors(50, 60.0, 30.0, 4)
    g = ((16, 16, 8), [0, 0, 0], [40.0, 40.0, 20.0])
    h, c = orc.digest(lors, *g, variant=FP)
    h2, c2 = orc.digest(lors[::-1].copy(), *g, variant=FP)
    assert np.array_equal(h, h2[::-1]) and np.array_equal(c, c2[::-1])
    # nudging the image centre by one voxel changes the hashes of rays that hit the box
    h3, c3 = orc.digest(lors, (16, 16, 8), [2.5, 0, 0], [40.0, 40.0, 20.0], variant=FP)
    hit = c > 0
    assert np.mean(h3[hit] != h[hit]) > 0.9


def test_em_drivers_closed_forms(orc):
    # one MLEM iteration from x=1 is B(y / A1) / emap (mlem.py:39-43);
    # OSEM with one subset equals MLEM (osem_full.py:67-77 with n_sub=1).
    rng = np.random.default_rng(9)
    shape, vox = (12, 12, 6), 4.0
    size = [48.0, 48.0, 24.0]
    center = [0.0, 0.0, 0.0]
    lors = random_cylinder_lors(3000, 60.0, 30.0, 6)
    y = rng.poisson(3.0, len(lors)).astype(np.float32)
    emap = orc.backproject(np.ones(len(lors), np.float32), lors, shape, center, size).astype(np.float32)
    x1 = orc.mlem(y, lors, emap, center, size, 1, driver="osem")
    A1 = orc.project(np.ones(shape, np.float32), lors, center, size)
    q = y / (A1 + np.float32(1e-8))
    want = orc.backproject(q, lors, shape, center, size).astype(np.float32) / emap
    np.testing.assert_allclose(x1[emap > 0], want[emap > 0], rtol=1e-6)
    xo = orc.osem(y, lors, emap, center, size, 3, 1, np.arange(len(lors)), driver="osem")
    xm = orc.mlem(y, lors, emap, center, size, 3, driver="osem")
    np.testing.assert_allclose(xo[emap > 0], xm[emap > 0], rtol=1e-6)


def test_attenuation_uniform_cylinder(orc):
    # uniform mu along a diameter: exponent = dx_ * mu * 2r (reference scaling, Q15)
    # or mu * 2r (physical).
    shape, vox = (50, 50, 10), 2.0
    size = [100.0, 100.0, 20.0]
    mu = 0.0096
    img = np.full(shape, mu, np.float32)
    lor = np.array([[-200.0, 1.0, 0.5, 200.0, 1.0, 0.5]], np.float32)
    a_ref = orc.attenuation_factor(img, lor, [0, 0, 0], size)
    a_phy = orc.attenuation_factor(img, lor, [0, 0, 0], size, physical=True)
    assert a_ref[0] == pytest.approx(np.exp(-2.0 * mu * 100.0), rel=1e-5)
    assert a_phy[0] == pytest.approx(np.exp(-mu * 100.0), rel=1e-5)


def test_quirk_q22_opposite_slopes(orc):
    # siddon.cu.cc:59-84 orders each secondary coordinate low->high on its own
    # (yy1/zz1 come from xx1 or xx2 by the sign of ky / kz), and :141-185 then
    # combines ry and rz as if both were measured from the same slab end.  When
    # ky and kz differ in sign and both cross in one slab the cells are credited
    # wrongly; the per-slab total (r1 + (r2-r1) + (1-r2) = 1) is still right.
    rng = np.random.default_rng(1)
    shape, vox = (24, 20, 12), 2.5
    size = [shape[0] * vox, shape[1] * vox, shape[2] * vox]
    center = [1.5, -2.0, 0.75]
    lors = _exact_rays(random_cylinder_lors(600, 80.0, 40.0, 7), shape, center, size, same_sign=False)[:100]
    img = rng.uniform(0.5, 2.0, shape).astype(np.float32)
    dx_ = size[0] / shape[0]
    p = orc.project(img, lors, center, size, f64=True)
    p1 = orc.project(np.ones(shape, np.float32), lors, center, size, f64=True)
    off = 0
    for l, pe, pu in zip(lors, p, p1):
        L2 = float(np.sum((l[3:6].astype(np.float64) - l[0:3]) ** 2))
        chord = chord_in_box(center, size, l[0:3], l[3:6])
        assert pu == pytest.approx(chord * dx_ * dx_ / L2, rel=1e-5)
        want = siddon_line_integral(img, center, size, l[0:3], l[3:6]) * dx_ * dx_ / L2
        off += abs(pe - want) > 1e-5 * want
    assert off >= 5  # every ray with a double-crossing slab of opposite slopes


def tof_moments(project, vox, tof_bin, sig2, tofv):
    """Mean and variance, in voxels, of the TOF weights along an axis-aligned
    LOR (x through voxel centres, one term per slab of geometric weight
    dx_/L^2): projections of the images 1, k - 128 and (k - 128)^2, k the x
    index.  `project(image, lors)` returns float64 projections."""
    shape = (256, 4, 4)
    size = [256.0 * vox, 4.0 * vox, 4.0 * vox]
    y = z = 0.5 * vox
    lor = np.array([[-400.0 * vox, y, z, 400.0 * vox, y, z, tofv]], np.float32)
    k = np.arange(256, dtype=np.float64) - 128.0
    s = []
    for g in (np.ones_like(k), k, k * k):
        img = np.broadcast_to(g[:, None, None], shape).astype(np.float32)
        s.append(float(project(img, lor, [0.0, 0.0, 0.0], size)[0]))
    mean = s[1] / s[0]
    return s[0], mean, s[2] / s[0] - mean * mean


@pytest.mark.parametrize("vox", [1.0, 2.0])
def test_tof_width_includes_bin_term(orc, vox):
    # siddon_tof.cu.cc:56-59: the Gaussian's variance is
    # tof_sigma2/dx_^2 + (tof_bin/dx_)^2/12 (the bin's uniform spread added to
    # the timing spread).  A Gaussian sampled at unit spacing keeps its
    # variance to within exp(-2 pi^2 var) and its integral to tof_bin_, so the
    # moments of the per-slab TOF weights fix both terms: with the bin term
    # dropped the variance would be 9, not 21.
    tof_bin, sig2 = 12.0 * vox, 9.0 * vox * vox
    for tofv in (0.0, 30.0 * vox):
        s0, mean, var = tof_moments(lambda im, l, c, s: orc.project(im, l, c, s, tof=(tof_bin, sig2), f64=True),
                                    vox, tof_bin, sig2, tofv)
        L = 800.0   # LOR length in voxels; geometric weight dx_ / L^2 per slab
        assert s0 * L * L / vox == pytest.approx(tof_bin / vox, rel=1e-5)   # integral of w_tof = tof_bin_
        assert var == pytest.approx(9.0 + 12.0 * 12.0 / 12.0, rel=2e-5)   # sigma^2 + bin^2 / 12
        # the centre sits tof/2 from the LOR midpoint towards p1 (tof = d(p2) - d(p1), Q21)
        assert mean == pytest.approx(-0.5 - tofv / vox / 2.0, abs=1e-3)
