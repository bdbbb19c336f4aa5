"""Crystal geometry (srf_amd.scanner) pinned without the reference: closed-form
crystal centres, the inverse maps against the forward ones on every crystal,
rotation of blocks / sectors, and the reference's module-stride quirk."""
import math

import numpy as np
import pytest

from srf_amd import scanner as S


def test_ecat_presets_sizes():
    # geometry/existing_geometry.py:15-20
    assert S.scanner_mct.nb_crystals == 4 * 48 * 13 * 13
    assert S.scanner_1_4m.nb_crystals == 26 * 48 * 15 * 15
    assert S.scanner_mct.axial_length == pytest.approx(4 * 53.3)
    assert S.scanner_1_4m.axial_length == pytest.approx(26 * 51.3 + 25 * 3.42)
    assert S.scanner_mct.blocks.unit_size == pytest.approx([20.0, 53.3 / 13, 53.3 / 13])


def test_ecat_crystal_zero_closed_form():
    sc = S.scanner_mct
    p = S.EcatIndexToCrystalPos(sc)(np.array([0]))[0]
    u = 53.3 / 13
    assert p[0] == pytest.approx((424.5 + 444.5) / 2, abs=1e-4)        # block 0 on +x
    assert p[1] == pytest.approx(0.5 * u - 53.3 / 2, abs=1e-4)
    assert p[2] == pytest.approx(0.5 * u - 4 * 53.3 / 2, abs=1e-4)


def test_ecat_block_rotation_and_radius():
    sc = S.scanner_mct
    pos = S.crystal_positions(sc).astype(np.float64)
    npb = sc.nb_crystals_per_block
    b0 = pos[:npb]
    for k in (1, 7, 12, 47):
        t = k * 2 * math.pi / 48
        rot = np.stack([b0[:, 0] * math.cos(t) - b0[:, 1] * math.sin(t),
                        b0[:, 0] * math.sin(t) + b0[:, 1] * math.cos(t), b0[:, 2]], 1)
        np.testing.assert_allclose(pos[k * npb:(k + 1) * npb], rot, atol=1e-3)
    # distance from the axis: sqrt(R^2 + y^2) with y the in-block offset
    u = 53.3 / 13
    iy = np.arange(sc.nb_crystals) % 13
    y = (iy + 0.5) * u - 53.3 / 2
    np.testing.assert_allclose(np.hypot(pos[:, 0], pos[:, 1]), np.hypot(434.5, y), rtol=1e-6)
    # neighbouring crystals of a block are one unit apart
    d = np.linalg.norm(pos[1:13] - pos[0:12], axis=1)
    np.testing.assert_allclose(d, u, rtol=1e-5)


@pytest.mark.parametrize("name", ["scanner_mct", "scanner_1_4m"])
def test_ecat_round_trip(name):
    sc = getattr(S, name)
    ind = np.arange(sc.nb_crystals)
    pos = S.EcatIndexToCrystalPos(sc)(ind)
    assert np.array_equal(S.EcatCrystalPosToIndex(sc)(pos), ind)
    iy, iz, ib, ir = S.EcatCrystalPosToIndex(sc)(pos, verbose=True)
    np.testing.assert_array_equal(S.EcatIndexToCrystalPos(sc)((iy, iz, ib, ir), verbose=True), pos)


def test_cylindrical_sizes_closed_form():
    sc = S.scanner_8panel
    assert sc.sz_submodule == pytest.approx([3 + 3.2 * 5, 3 + 3.2 * 5])
    assert sc.sz_module == pytest.approx([19.0 + 20 * 2, 19.0 + 20 * 3])
    assert sc.sz_rsector == pytest.approx([59.0 + 60, 79.0 + 80])
    assert sc.nb_crystals == 8 * 4 * 12 * 36
    p = S.CylindricalIndexToCrystalPos(sc)(np.array([0]))[0]
    # crystal 0: first crystal of the first submodule of the first module, rsector 0 on +x
    y = 0.5 * 3 - 19.0 / 2 + (-59.0 / 2 + 19.0 / 2) + (-119.0 / 2 + 59.0 / 2)
    z = 0.5 * 3 - 19.0 / 2 + (-79.0 / 2 + 19.0 / 2) + (-159.0 / 2 + 79.0 / 2)
    assert p == pytest.approx([(146.7 + 166.7) / 2, y, z], abs=1e-4)


@pytest.mark.parametrize("name", ["scanner_8panel", "scanner_20panel"])
def test_cylindrical_round_trip(name):
    sc = getattr(S, name)
    ind = np.arange(sc.nb_crystals)
    pos = S.CylindricalIndexToCrystalPos(sc)(ind)
    assert np.array_equal(S.CylindricalCrystalPosToIndex(sc)(pos), ind)
    # rsector k is rsector 0 rotated by k * 2 pi / nb_rsector
    n = sc.nb_crystal_per_rsector
    t = 3 * 2 * math.pi / sc.nb_rsector
    b0 = pos[:n].astype(np.float64)
    rot = np.stack([b0[:, 0] * math.cos(t) - b0[:, 1] * math.sin(t),
                    b0[:, 0] * math.sin(t) + b0[:, 1] * math.cos(t), b0[:, 2]], 1)
    np.testing.assert_allclose(pos[3 * n:4 * n], rot, atol=1e-3)


def test_cylindrical_module_stride_quirk():
    """With nb_module[0] != nb_module[1] the reference's position -> index
    strides the module z index by nb_module[1] (pet_cylindrical_scanner.py:37-38)
    where index -> position uses nb_module[0] (:132-136): module (0, 1) comes
    back as the index of module 3, not 2."""
    sc = S.scanner_8panel.update(nb_module=[2, 3])
    v = S.CylindricalIndexToCrystalPos(sc)
    cy, cz, sy, sz, my, mz, rs = [np.array([k]) for k in (0, 0, 0, 0, 0, 1, 0)]
    p = v((cy, cz, sy, sz, my, mz, rs), verbose=True)
    back = S.CylindricalCrystalPosToIndex(sc)(p, verbose=True)
    assert (back[4][0], back[5][0]) == (0, 1)                     # the decomposition itself is right
    got = S.CylindricalCrystalPosToIndex(sc)(p)
    assert got[0] == 1 * 3 * sc.nb_crystal_per_module             # mz * nb_module[1]
    assert got[0] != 1 * 2 * sc.nb_crystal_per_module             # what the forward map uses


def test_all_lors_id_order():
    """Pairs (i, j), i > j, row by row: bin (i - 1) i / 2 + j, N (N - 1) / 2 bins
    (listmode2sinogram.py:27-35)."""
    from srf_amd.data_transform import get_all_lors_id
    n = 37
    g = get_all_lors_id(n)
    assert g.shape == (n * (n - 1) // 2, 2) and g.dtype == np.int32
    assert (g[:, 0] > g[:, 1]).all() and (g[:, 1] >= 0).all()
    b = (g[:, 0].astype(np.int64) - 1) * g[:, 0] // 2 + g[:, 1]
    assert np.array_equal(b, np.arange(g.shape[0]))
    assert g[:6].tolist() == [[1, 0], [2, 0], [2, 1], [3, 0], [3, 1], [3, 2]]
