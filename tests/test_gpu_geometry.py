"""Crystal-pair sensitivity image on the GPU (srf_pair_lors + tiled plan
backprojection, functions.EmapGenerator / crystal_pair_emap) against the CPU
oracle's backprojection of the same pair LORs, plus chunking invariance and the
rotational symmetry of a ring whose block count is a multiple of four."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPE, CENTER, SIZE = (32, 32, 8), [0.0, 0.0, 0.0], [128.0, 128.0, 32.0]


def rel_l2(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def small_ecat():
    from srf_amd.scanner import Block, PetEcatScanner
    return PetEcatScanner(60.0, 70.0, 2, 8, 2.0, Block([10.0, 24.0, 12.0], [1, 4, 2]))


def small_cyl():
    from srf_amd.scanner import PetCylindricalScanner
    return PetCylindricalScanner(60.0, 70.0, 6, [1, 1], [20, 20], [1, 1], [20, 20], [3, 2], [4, 4], [3.5, 3.5])


def config():
    from srf_amd.data import ImageConfig
    return ImageConfig(list(SHAPE), CENTER, SIZE)


def host_pairs(pos):
    n = pos.shape[0]
    return np.hstack((np.repeat(pos, n, axis=0), np.tile(pos, (n, 1)))).astype(np.float32)


def test_pair_lors_exact():
    import torch
    from srf_amd._lib import check, lib, stream_handle
    from srf_amd.scanner import crystal_positions
    pos = crystal_positions(small_ecat())
    n = pos.shape[0]
    P = torch.from_numpy(pos).cuda()
    out = torch.empty((n * n - 100, 6), dtype=torch.float32, device="cuda")
    check(lib().srf_pair_lors(P.data_ptr(), n, 37, out.shape[0], out.data_ptr(), stream_handle(P.device)),
          "srf_pair_lors")
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), host_pairs(pos)[37:37 + out.shape[0]])
    # a range past the last pair is rejected
    rc = lib().srf_pair_lors(P.data_ptr(), n, n * n - 10, 11, out.data_ptr(), stream_handle(P.device))
    assert rc != 0


@pytest.mark.parametrize("make", [small_ecat, small_cyl], ids=["ecat", "cylindrical"])
def test_emap_matches_oracle(make):
    import oracle
    from srf_amd.functions import EmapGenerator
    from srf_amd.scanner import crystal_positions
    sc = make()
    lors = host_pairs(crystal_positions(sc))
    print(f"{type(sc).__name__}: {sc.nb_crystals} crystals, {lors.shape[0]} pairs", flush=True)
    ref = oracle.backproject(np.ones(lors.shape[0], np.float32), lors, SHAPE, CENTER, SIZE)
    em = EmapGenerator("full", sc)(config())
    assert em.data.shape == SHAPE
    assert float(ref.sum()) > 0
    err = rel_l2(em.data, ref)
    print(f"rel L2 vs oracle {err:.2e}", flush=True)
    assert err <= 1e-5


def test_emap_chunking_and_modes():
    from srf_amd.functions import crystal_pair_emap, EmapGenerator
    sc = small_ecat()
    stats = {}
    a = crystal_pair_emap(sc, config(), stats=stats).cpu().numpy()
    b = crystal_pair_emap(sc, config(), chunk_pairs=1000).cpu().numpy()
    assert stats["pairs"] == sc.nb_crystals ** 2 and stats["chunks"] == 1
    assert rel_l2(b, a) <= 1e-6
    for mode in ("block", "block-full", "rsector-full"):
        assert np.array_equal(EmapGenerator(mode, sc)(config()).data, EmapGenerator("full", sc)(config()).data)
    with pytest.raises(NotImplementedError):
        EmapGenerator("nope", sc)(config())


def test_emap_quarter_turn_symmetry():
    """8 blocks per ring: a quarter turn maps the crystal set onto itself, and the
    centred square grid onto itself, so the sensitivity image is symmetric up to
    the marcher's x/y tie-breaking and position round-off."""
    from srf_amd.functions import EmapGenerator
    em = EmapGenerator("full", small_ecat())(config()).data
    rot = np.rot90(em, k=1, axes=(0, 1))
    err = rel_l2(rot, em)
    print(f"quarter-turn asymmetry {err:.2e}", flush=True)
    assert err <= 2e-2
