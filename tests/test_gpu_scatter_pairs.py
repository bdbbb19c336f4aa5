"""The reference's crystal-pair scatter correction on the GPU
(srf_amd.scatter_pairs; sss.ScatterCorrect with a preprocessing directory or
dict) against the oracle's restatement (oracle/sss_oracle.py ref_*, the C
oracle's projector for the line integrals) on a 768-crystal ECAT scanner:

  step (reference)                          bar
  sinogram, counted bins, scatter points    identical
  emission / attenuation integrals          1e-5 relative L2 (projector)
  1.3 mu pair attenuation (atten_lors)      1e-5
  per-pair scatter sum (fraction)           1e-3 relative L2 (integer-degree
                                            angles decided in float32 vs float64)
  per-pair scale                            1e-6
  corrected counted bins                    1e-3 relative L2 (the fraction's
                                            difference carried through the
                                            subtraction)
  corrected listmode (sino2lm)              same crystal pairs except bins whose
                                            value is within 1e-3 of the 0 cut

and the file round trip: scatter_preprocess's .npy files (the reference's
names, dtypes and shapes) read back by ScatterCorrect give the same listmode as
the in-memory path."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

SHAPE, SIZE, CENTER = (96, 96, 80), [192.0, 192.0, 160.0], [0.0, 0.0, 0.0]
LOW, HIGH, RES, SF = 400.0, 650.0, 0.15, 0.3


def rel_l2(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def ecat():
    from srf_amd.scanner import Block, PetEcatScanner
    return PetEcatScanner(100.0, 120.0, 4, 12, 0.0, Block([20.0, 40.0, 40.0], [1, 4, 4]))


@pytest.fixture(scope="module")
def case():
    return make_case()


def make_case():
    """The scanner, a cylinder emission with a hot sphere, a water mu-map with
    a denser sphere, and 30000 events between random crystal pairs (endpoints
    within 1 mm of the crystal centres)."""
    from srf_amd.scanner import crystal_positions
    sc = ecat()
    c = (np.arange(SHAPE[0]) + 0.5) * 2.0 - 96.0
    z = (np.arange(SHAPE[2]) + 0.5) * 2.0 - 80.0
    X, Y, Z = np.meshgrid(c, c, z, indexing="ij")
    r = np.sqrt(X ** 2 + Y ** 2)
    em = np.where(r < 60.0, 1.0, 0.0) + np.where(X ** 2 + (Y - 20) ** 2 + Z ** 2 < 15.0 ** 2, 4.0, 0.0)
    mu = np.where(r < 70.0, 0.0096, 0.0) + np.where((X + 25) ** 2 + Y ** 2 + (Z - 10) ** 2 < 20.0 ** 2, 0.0076, 0.0)
    rng = np.random.default_rng(8)
    pos = crystal_positions(sc).astype(np.float64)
    n = 30000
    a = rng.integers(0, sc.nb_crystals, n)
    b = (a + rng.integers(1, sc.nb_crystals, n)) % sc.nb_crystals
    lors = np.hstack((pos[a] + rng.uniform(-1, 1, (n, 3)), pos[b] + rng.uniform(-1, 1, (n, 3)))).astype(np.float32)
    return sc, em.astype(np.float32), mu.astype(np.float32), lors


def oracle_side(orc, sc, em, mu, lors):
    import sss_oracle as so
    from srf_amd.scanner import crystal_pos_to_index, crystal_positions
    from srf_amd.scatter_pairs import block_params
    pos = crystal_positions(sc)
    i1 = np.asarray(crystal_pos_to_index(sc, lors[:, 0:3]))
    i2 = np.asarray(crystal_pos_to_index(sc, lors[:, 3:6]))
    hi, lo = np.maximum(i1, i2), np.minimum(i1, i2)
    ncr = sc.nb_crystals
    sino = np.bincount(hi * (hi - 1) // 2 + lo, minlength=ncr * (ncr - 1) // 2).astype(np.float64)
    index = np.where(sino > 0)[0].astype(np.int32)
    pts = so.ref_scatter_positions(SHAPE, SIZE)
    ns = pts.shape[0]
    lcp = np.hstack((np.repeat(pos, ns, axis=0), np.tile(pts, (len(pos), 1)))).astype(np.float32)
    d = lcp[:, 3:6] - lcp[:, 0:3]
    len2 = np.power(d[:, 0], 2) + np.power(d[:, 1], 2) + np.power(d[:, 2], 2)
    sumup = orc.project(em, lcp, CENTER, SIZE) * len2
    atten = orc.attenuation_factor(mu, lcp, CENTER, SIZE)
    pairs = np.stack(_pairs(index), 1)
    plors = np.hstack((pos[pairs[:, 0]], pos[pairs[:, 1]])).astype(np.float32)
    atten_lors = orc.attenuation_factor((mu * np.float32(1.3)).astype(np.float32), plors, CENTER, SIZE)
    nbr, ncrr, grid, size_block = block_params(sc)
    frac, scale = so.ref_scatter_fraction(pairs, pos, pts, sumup, atten, mu, SIZE, int(LOW), int(HIGH), RES,
                                          nbr, ncrr, grid, size_block)
    corr = so.ref_correct(sino, index, frac, scale, atten_lors, SF)
    lm = so.ref_sino2lm(corr, pos)
    return dict(sino=sino, index=index, pts=pts, sumup=sumup, atten=atten, atten_lors=atten_lors, frac=frac,
                scale=scale, corr=corr[index], lm=lm)


def _pairs(index):
    k = np.asarray(index, np.int64)
    i = np.floor((1 + np.sqrt(1 + 8 * k.astype(np.float64))) / 2).astype(np.int64)
    i = np.where(i * (i - 1) // 2 > k, i - 1, i)
    i = np.where((i + 1) * i // 2 <= k, i + 1, i)
    return i, k - i * (i - 1) // 2


def test_crystal_pair_scatter_correct_matches_oracle(orc, case, tmp_path):
    from srf_amd import Listmode, Lors
    from srf_amd.data import Image, Umap
    from srf_amd.scanner import crystal_positions
    from srf_amd.scatter_pairs import correct_pairs, pair_scatter, scatter_preprocess
    from srf_amd.sss import ScatterCorrect
    sc, em, mu, lors = case
    emission = Image(em, CENTER, SIZE)
    umap = Umap(mu, CENTER, SIZE)
    lm_in = Listmode(np.ones(len(lors), np.float32), Lors(lors))
    pre = scatter_preprocess(sc, lm_in, emission, umap)
    frac, scale = pair_scatter(pre, umap, sc, LOW, HIGH, RES)
    corr, lm = correct_pairs(pre, frac, scale, SF)
    torch.cuda.synchronize()
    want = oracle_side(orc, sc, em, mu, lors)
    g = {k: v.cpu().numpy() for k, v in pre.items()}
    assert np.array_equal(g["sinogram"], want["sino"])
    assert np.array_equal(g["index"], want["index"])
    assert np.array_equal(g["scatter_position"], want["pts"])
    # pairs on one block face: (B - A) . n_B = 0, so the reference's scale is
    # infinite or, after rounding, astronomically large; such bins are
    # dropped by sino2lm's value < 1e8 on both sides (checked below)
    deg = ~np.isfinite(want["scale"]) | (np.abs(want["scale"]) > 1e12)
    ok = ~deg
    gs, gc = scale.cpu().numpy(), corr.cpu().numpy()
    errs = {"sumup": rel_l2(g["sumup_emission"], want["sumup"]), "atten": rel_l2(g["atten"], want["atten"]),
            "atten_lors": rel_l2(g["atten_lors"], want["atten_lors"]),
            "fraction": rel_l2(frac.cpu().numpy(), want["frac"]), "scale": rel_l2(gs[ok], want["scale"][ok]),
            "corrected": rel_l2(gc[ok], want["corr"][ok])}
    print(f"{len(want['index'])} counted pairs ({int(deg.sum())} on one block face), {pts_count(want)} scatter "
          f"points; relative L2 " + ", ".join(f"{k} {v:.2e}" for k, v in errs.items()), flush=True)
    assert errs["sumup"] <= 1e-5 and errs["atten"] <= 1e-5 and errs["atten_lors"] <= 1e-5
    assert errs["fraction"] <= 1e-3 and errs["scale"] <= 1e-6 and errs["corrected"] <= 1e-3
    assert np.all(~((gc[deg] > 0) & (gc[deg] < 1e8)))          # dropped by sino2lm
    # the listmode: the same crystal pairs, except where a bin's count and its
    # subtracted scatter share K f_i agree within 1% (there the 1e-3 fraction
    # bar allows the sign, and so the 0 cut, to differ)
    got_lm = lm.lors.data.cpu().numpy().astype(np.float64)
    pos = crystal_positions(ecat()).astype(np.float64)
    key = lambda r: tuple(np.round(r[:6], 3))
    wk = {key(r) for r in want["lm"]}
    gk = {key(r) for r in got_lm}
    y = want["sino"][want["index"]]
    share = want["frac"] / want["frac"].sum() * y.sum() * SF
    near = np.abs(y - share) <= 1e-2 * share
    i, j = _pairs(want["index"])
    near_keys = {key(np.concatenate([pos[a], pos[b]])) for a, b in zip(i[near], j[near])}
    diff = wk ^ gk
    print(f"listmode: {len(gk)} events (oracle {len(wk)}), {len(diff)} differ, {int(near.sum())} bins within 1% "
          f"of the cut", flush=True)
    assert diff <= near_keys and len(gk) > 0.5 * len(want["index"])
    assert np.array_equal(got_lm[:, 6], lm.data.cpu().numpy())     # value repeated in column 6
    # ScatterCorrect with the preprocessing dict, then through the reference's files
    via_dict = ScatterCorrect(LOW, HIGH, RES, SF)(emission, umap, sc, pre)
    scatter_preprocess(sc, lm_in, emission, umap, outdir=str(tmp_path), out_append="_t")
    names = sorted(p.name for p in tmp_path.iterdir())
    assert names == sorted(f"{k}_t.npy" for k in ("lors", "sinogram", "index", "scatter_position",
                                                   "crystal_position", "sumup_emission", "atten", "atten_lors"))
    assert np.load(tmp_path / "sinogram_t.npy").shape == (sc.nb_crystals * (sc.nb_crystals - 1) // 2, 1)
    via_files = ScatterCorrect(LOW, HIGH, RES, SF)(emission, umap, sc, str(tmp_path), "_t")
    torch.cuda.synchronize()
    for out in (via_dict, via_files):
        assert torch.equal(out.data, lm.data) and torch.equal(out.lors.data, lm.lors.data)


def pts_count(want):
    return int(want["pts"].shape[0])
