"""The srfnef reconstruction drivers with their corrections, on the GPU, against
the oracle's restatement of the reference drivers.

- OsemFull (srfnef/functions/osem_full.py:49-87): the listmode is snapped to
  crystal centres (PetListmodeTrans, :54), attenuation corrected along the
  snapped LORs (:56-58) and reconstructed along the original LORs (:70-74)
  with q = y / (p + 1e-8) and an unguarded emap (:74-77).  Both subset orders:
  i % n_sub and the reference's i // n_sub (:68, quirk Q14).  The emap (all
  crystal pairs plus the listmode's LORs) is zero outside the ring, where no
  LOR passes, so those voxels become 0/0 = NaN (quirk Q13) in both.
- MlemFullTof with AttenuationCorrect (mlem_full_tof.py:43-67): the
  attenuation-corrected values, div_no_nan(y, p + 1e-5), emap zeros -> 1e8.
- MlemFull with AttenuationCorrect (mlem_full.py:47-72): div_no_nan(y, p); its
  scatter block is commented out in the reference (:74-93), so a scatter
  corrector passed to it is never called.

Oracle side: host crystal maps (srf_amd.scanner, pinned by
test_scanner_cpu.py) for the snapping, oracle.attenuation_factor for the
factors, oracle.osem / oracle.mlem for the loop.  Bar: NaN voxels identical,
finite voxels within 1e-4 relative L2."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

SHAPE, CENTER, SIZE = (32, 32, 8), [0.0, 0.0, 0.0], [128.0, 128.0, 32.0]


def rel_l2(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def small_ecat():
    from srf_amd.scanner import Block, PetEcatScanner
    return PetEcatScanner(60.0, 70.0, 2, 8, 2.0, Block([10.0, 24.0, 12.0], [1, 4, 2]))


def water_umap():
    ix = (np.arange(SHAPE[0]) + 0.5) * SIZE[0] / SHAPE[0] - SIZE[0] / 2
    r = np.sqrt(ix[:, None] ** 2 + ix[None, :] ** 2)
    mu = np.where(r < 40.0, 0.0096, 0.0).astype(np.float32)
    return np.repeat(mu[:, :, None], SHAPE[2], axis=2)


def crystal_listmode(scanner, n, seed):
    """n events between random distinct crystal pairs, endpoints jittered by up
    to 1 mm (well inside a 6 mm crystal)."""
    from srf_amd.scanner import crystal_positions
    rng = np.random.default_rng(seed)
    pos = crystal_positions(scanner).astype(np.float64)
    nc = pos.shape[0]
    a = rng.integers(0, nc, n)
    b = (a + rng.integers(1, nc, n)) % nc
    p1 = pos[a] + rng.uniform(-1.0, 1.0, (n, 3))
    p2 = pos[b] + rng.uniform(-1.0, 1.0, (n, 3))
    return np.hstack((p1, p2)).astype(np.float32)


def host_snap(scanner, lors):
    from srf_amd.scanner import crystal_pos_to_index, crystal_positions
    table = crystal_positions(scanner)
    i1 = np.asarray(crystal_pos_to_index(scanner, lors[:, 0:3]))
    i2 = np.asarray(crystal_pos_to_index(scanner, lors[:, 3:6]))
    assert i1.min() >= 0 and i2.min() >= 0 and max(i1.max(), i2.max()) < len(table)
    return np.hstack((table[i1], table[i2])).astype(np.float32)


def pair_emap(orc, scanner):
    """The all-pairs sensitivity image by the oracle's backprojection."""
    from srf_amd.scanner import crystal_positions
    pos = crystal_positions(scanner)
    n = pos.shape[0]
    pairs = np.hstack((np.repeat(pos, n, axis=0), np.tile(pos, (n, 1)))).astype(np.float32)
    return orc.backproject(np.ones(len(pairs), np.float32), pairs, SHAPE, CENTER, SIZE).astype(np.float32)


def compare_with_nan(got, want, tag):
    got = np.asarray(got, np.float64)
    want = np.asarray(want, np.float64)
    nan_g, nan_w = np.isnan(got), np.isnan(want)
    fin = np.isfinite(want)
    err = rel_l2(got[fin], want[fin])
    print(f"{tag}: NaN voxels {int(nan_w.sum())} (GPU {int(nan_g.sum())}), finite relative L2 {err:.3g}",
          flush=True)
    assert np.array_equal(nan_g, nan_w)
    assert np.array_equal(np.isinf(got), np.isinf(want))
    assert err <= 1e-4


@pytest.mark.parametrize("reference_order", [False, True], ids=["i_mod_nsub", "i_div_nsub"])
def test_osem_full_snapped_attenuated_matches_oracle(orc, reference_order):
    from srf_amd import Emap, ImageConfig, Listmode, Lors
    from srf_amd.corrections import AttenuationCorrect
    from srf_amd.data import Umap
    from srf_amd.functions import OsemFull, subset_permutation
    sc = small_ecat()
    lors = crystal_listmode(sc, 40000, 3)
    y = np.random.default_rng(4).uniform(0.5, 1.5, len(lors)).astype(np.float32)
    mu = water_umap()
    # all crystal pairs plus the listmode's own LORs: zero exactly where no LOR
    # passes (outside the ring), so 0/0 = NaN stays in those voxels (Q13)
    emap = pair_emap(orc, sc) + orc.backproject(np.ones(len(lors), np.float32), lors, SHAPE, CENTER,
                                                SIZE).astype(np.float32)
    assert (emap == 0).sum() > 0
    n_iter, n_sub, seed = 6, 3, 11
    x = OsemFull(n_iter=n_iter, n_sub=n_sub, image_config=ImageConfig(list(SHAPE), CENTER, SIZE), scanner=sc,
                 listmode=Listmode(y, Lors(lors)), atten_corr=AttenuationCorrect(Umap(mu, CENTER, SIZE)),
                 scatter_corr=None, norm_corr=False, psf_corr=None, emap=Emap(emap, CENTER, SIZE), image=None,
                 seed=seed, reference_subset_order=reference_order)().data
    snapped = host_snap(sc, lors)
    a = orc.attenuation_factor(mu, snapped, CENTER, SIZE)
    assert a.min() < 0.6
    vals = (y / a).astype(np.float32)
    perm = subset_permutation(len(lors), seed)
    want = orc.osem(vals, lors, emap, CENTER, SIZE, n_iter, n_sub, perm, driver="osem",
                    reference_subset_order=reference_order, guard_emap=False)
    compare_with_nan(x, want, f"OsemFull reference_subset_order={reference_order}")
    assert np.isnan(want).sum() > 0


def test_osem_full_default_emap_is_crystal_pairs(orc):
    """With emap None a crystal scanner gets the all-pairs sensitivity image
    (the reference's 'ring' mode exists in no generator); it equals the
    oracle's backprojection of every crystal pair."""
    from srf_amd import ImageConfig, Listmode, Lors
    from srf_amd.functions import OsemFull
    sc = small_ecat()
    lors = crystal_listmode(sc, 5000, 5)
    osem = OsemFull(n_iter=1, n_sub=1, image_config=ImageConfig(list(SHAPE), CENTER, SIZE), scanner=sc,
                    listmode=Listmode(np.ones(len(lors), np.float32), Lors(lors)), atten_corr=None,
                    scatter_corr=None, norm_corr=False, psf_corr=None, emap=None, image=None)
    osem()
    want = pair_emap(orc, sc)
    err = rel_l2(osem.emap.data, want)
    print(f"OsemFull default emap vs oracle all-pairs: relative L2 {err:.3g}", flush=True)
    assert err <= 1e-5


@pytest.fixture(scope="module")
def cylinder_problem(orc):
    from srf_amd import phantom
    act, center, size = phantom.cylinder_spheres_phantom((32, 32, 16), 6.8)
    lors = phantom.sample_events(act, center, size, 50000, 300.0, 110.0, seed=31, tof_fwhm_mm=60.0).numpy()
    elors = phantom.sample_events(torch.ones_like(act), center, size, 50000, 300.0, 110.0, seed=32).numpy()
    emap = orc.backproject(np.ones(len(elors), np.float32), elors[:, :6], act.shape, center, size)
    mu = np.where(act.numpy() > 0, 0.0096, 0.0).astype(np.float32)
    return act.numpy(), center, size, lors, emap.astype(np.float32), mu


def test_mlem_full_tof_attenuation_matches_oracle(orc, cylinder_problem):
    from srf_amd import Emap, ImageConfig, Listmode, Lors, TofConfig
    from srf_amd.corrections import AttenuationCorrect
    from srf_amd.data import Umap
    from srf_amd.functions import CylinderScanner, MlemFullTof
    act, center, size, lors, emap, mu = cylinder_problem
    tof = TofConfig(3.66, (60.0 / 2.3548200450309493) ** 2)
    y = np.ones(len(lors), np.float32)
    x = MlemFullTof(n_iter=4, image_config=ImageConfig(list(act.shape), center, size),
                    scanner=CylinderScanner(300.0, 110.0, tof), listmode=Listmode(y, Lors(lors)),
                    atten_corr=AttenuationCorrect(Umap(mu, center, size)), scatter_corr=None, psf_corr=None,
                    emap=Emap(emap, center, size))().data
    a = orc.attenuation_factor(mu, lors[:, :6], center, size)
    want = orc.mlem((y / a).astype(np.float32), lors, emap, center, size, 4, tof=(tof.tof_bin, tof.tof_sigma2),
                    driver="full_tof")
    err = rel_l2(x[emap > 0], want[emap > 0])
    print(f"MlemFullTof + attenuation: factors {a.min():.3g}..{a.max():.3g}, relative L2 {err:.3g}", flush=True)
    assert a.min() < 0.3 and err <= 1e-4


class _NeverCalled:
    def __call__(self, *args, **kwargs):
        raise AssertionError("MlemFull must not run a scatter correction (mlem_full.py:74-93 is commented out)")


def test_mlem_full_attenuation_matches_oracle_and_ignores_scatter(orc, cylinder_problem):
    from srf_amd import Emap, ImageConfig, Listmode, Lors
    from srf_amd.corrections import AttenuationCorrect
    from srf_amd.data import Umap
    from srf_amd.functions import CylinderScanner, MlemFull
    act, center, size, lors, emap, mu = cylinder_problem
    l6 = np.ascontiguousarray(lors[:, :6])
    y = np.ones(len(l6), np.float32)
    x = MlemFull(n_iter=4, image_config=ImageConfig(list(act.shape), center, size),
                 scanner=CylinderScanner(300.0, 110.0, None), listmode=Listmode(y, Lors(l6)),
                 atten_corr=AttenuationCorrect(Umap(mu, center, size)), scatter_corr=_NeverCalled(),
                 psf_corr=None, emap=Emap(emap, center, size))().data
    a = orc.attenuation_factor(mu, l6, center, size)
    want = orc.mlem((y / a).astype(np.float32), l6, emap, center, size, 4, driver="full")
    err = rel_l2(x[emap > 0], want[emap > 0])
    print(f"MlemFull + attenuation: relative L2 {err:.3g}", flush=True)
    assert err <= 1e-4
