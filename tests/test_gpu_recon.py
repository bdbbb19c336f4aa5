"""End-to-end reconstruction from a recon description on the GPU: GATE-style
listmode files snapped to crystals, the crystal-pair sensitivity image, MLEM
(srf_amd.recon.recon_full) equal to the same MlemFull built by hand."""
import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from test_gpu_geometry import small_ecat  # noqa: E402


def gate_events(sc, n, seed):
    """Crystal pairs whose LOR passes within 12 mm of the axis are kept five
    times as often (a hot centre), endpoints moved inside their crystals."""
    from srf_amd.scanner import crystal_positions
    rng = np.random.default_rng(seed)
    pos = crystal_positions(sc).astype(np.float64)
    out = []
    while sum(len(o) for o in out) < n:
        a, b = rng.integers(0, pos.shape[0], 4 * n), rng.integers(0, pos.shape[0], 4 * n)
        p, q = pos[a], pos[b]
        d = q[:, :2] - p[:, :2]
        L = np.hypot(d[:, 0], d[:, 1])
        ok = L > 30.0
        dist = np.abs(p[:, 0] * d[:, 1] - p[:, 1] * d[:, 0]) / np.maximum(L, 1e-9)
        keep = ok & (rng.random(4 * n) < np.where(dist < 12.0, 1.0, 0.2))
        jit = rng.uniform(-0.4, 0.4, (4 * n, 2)) * np.array([sc.blocks.unit_size[2], sc.blocks.unit_size[2]])
        p[:, 2] += jit[:, 0]
        q[:, 2] += jit[:, 1]
        extra = rng.random((4 * n, 1))
        out.append(np.hstack((p, q, extra))[keep])
    return np.vstack(out)[:n].astype(np.float32)


def test_gate_listmode_snapped(tmp_path):
    from srf_amd.recon import listmode_from_gate_out
    from srf_amd.data_transform import snap_lors
    sc = small_ecat()
    ev = [gate_events(sc, 3000, s) for s in (1, 2)]
    for i, e in enumerate(ev):
        np.save(tmp_path / f"coinc_{i}.npy", e)
    lm = listmode_from_gate_out(str(tmp_path / "coinc_?.npy"), sc, 2)
    assert lm.lors.data.shape == (6000, 6) and float(lm.data.sum()) == 6000.0
    ref, outside = snap_lors(sc, np.vstack(ev)[:, :6])
    assert outside == 0
    assert np.array_equal(lm.lors.data, ref.cpu().numpy())


def test_recon_full_end_to_end(tmp_path):
    from srf_amd.data import ImageConfig, Listmode, Lors
    from srf_amd.functions import MlemFull
    from srf_amd.recon import listmode_from_gate_out, recon_full
    sc = small_ecat()
    np.save(tmp_path / "coinc_0.npy", gate_events(sc, 40000, 7))
    lm = listmode_from_gate_out(str(tmp_path / "coinc_?.npy"), sc, 1)
    np.savez(tmp_path / "lm.npz", data=lm.data, lors=lm.lors.data)
    b = sc.blocks
    cfg = {"n_iter": 4,
           "image_config": {"shape": [32, 32, 8], "center": [0.0, 0.0, 0.0], "size": [128.0, 128.0, 32.0]},
           "scanner": {"inner_radius": sc.inner_radius, "outer_radius": sc.outer_radius, "nb_rings": sc.nb_rings,
                       "nb_blocks_per_ring": sc.nb_blocks_per_ring, "gap": sc.gap,
                       "blocks": {"size": list(b.size), "shape": list(b.shape), "interval": [0, 0, 0]},
                       "tof": None},
           "listmode": str(tmp_path / "lm.npz"), "atten_corr": None, "scatter_corr": None, "psf_corr": None,
           "emap": None}
    (tmp_path / "recon.json").write_text(json.dumps(cfg))
    img = recon_full(str(tmp_path / "recon.json"), out_path=str(tmp_path / "img.npy"))
    x = np.asarray(img.data)
    assert x.shape == (32, 32, 8) and np.isfinite(x).all() and (x >= 0).all()
    assert np.array_equal(np.load(tmp_path / "img.npy"), x)
    # the hot centre (LORs within 12 mm of the axis kept 5x as often)
    r = np.hypot(*np.meshgrid((np.arange(32) + 0.5) * 4 - 64, (np.arange(32) + 0.5) * 4 - 64, indexing="ij"))
    inner, ring = x[r < 8].mean(), x[(r > 20) & (r < 40)].mean()
    print(f"centre {inner:.4g} vs ring {ring:.4g}", flush=True)
    assert inner > 2 * ring
    # the same MlemFull built by hand
    ref = MlemFull(4, ImageConfig([32, 32, 8], [0.0, 0.0, 0.0], [128.0, 128.0, 32.0]), sc,
                   Listmode(lm.data, Lors(lm.lors.data)), None, None, None, None)()
    assert np.array_equal(np.asarray(ref.data), x)


def test_doc_v7_description_end_to_end_vs_oracle(orc, tmp_path):
    """The reference's recon description (usage_docs/doc_v7:7-79, fixture
    tests/golden/doc_v7_recon.json) run at its own scale: the mCT scanner, a
    200^3 x 2.05 mm image, 15 MLEM iterations, attenuation correction on a
    400 x 400 x 500 mm u-map, the all-crystal-pair sensitivity image
    (emap null).  Its HDF5 paths point at synthetic .npz / .npy inputs here:
    4e5 coincidences from a uniform water cylinder (r 180 mm, filling the
    axial field of view) with endpoints snapped to their crystals, and the
    cylinder's u-map.  psf_corr is off (it would fit point-source images).
    The image equals the oracle's MLEM (driver 'full': div_no_nan, emap 0 ->
    1e8) on the same listmode, attenuation-corrected values and sensitivity
    image within 1e-4 relative L2.  (Sparse random crystal pairs are not used:
    there the reference's div_no_nan meets estimates that have collapsed to
    ~1e-30 on some lines, and float rounding decides where inf and NaN appear.)"""
    import os
    import torch
    from srf_amd import phantom, scanner as S
    from srf_amd.data_transform import snap_lors
    from srf_amd.recon import json_load, recon_from_dict
    cfg = json_load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "doc_v7_recon.json"))
    sc = S.scanner_mct
    ic, um = cfg["image_config"], cfg["atten_corr"]["u_map"]
    ushape = (80, 80, 100)
    cu = [(np.arange(n) + 0.5) * (sz / n) - sz / 2 for n, sz in zip(ushape, um["size"])]
    X, Y, Z = np.meshgrid(*cu, indexing="ij")
    body = (X ** 2 + Y ** 2 < 180.0 ** 2) & (np.abs(Z) < 110.0)
    mu = np.where(body, 0.0096, 0.0).astype(np.float32)
    act = torch.from_numpy(body.astype(np.float32))
    ev = phantom.sample_events(act, um["center"], um["size"], 400000, sc.inner_radius, 213.2, 11, device="cuda")
    lors_t, outside = snap_lors(sc, ev[:, :6].contiguous())
    # endpoints off the crystal table become NaN (dropped here, as a GATE
    # listmode holds none), and so do same-column pairs' axial lines
    keep = torch.isfinite(lors_t).all(1) & (torch.hypot(lors_t[:, 3] - lors_t[:, 0], lors_t[:, 4] - lors_t[:, 1]) > 1.0)
    lors = lors_t[keep].contiguous().cpu().numpy()
    np.savez(tmp_path / "lm.npz", data=np.ones(len(lors), np.float32), lors=lors)
    np.save(tmp_path / "umap.npy", mu)
    cfg["listmode"] = str(tmp_path / "lm.npz")
    um["data"] = str(tmp_path / "umap.npy")
    cfg["psf_corr"] = None
    print(f"doc_v7: mCT, 200^3, 15 iterations, {len(lors)} events ({outside} endpoints outside)", flush=True)
    m = recon_from_dict(cfg)
    x = np.asarray(m().data)
    emap = m.emap.data
    emap = emap.cpu().numpy() if isinstance(emap, torch.Tensor) else np.asarray(emap)
    assert x.shape == (200, 200, 200) and np.isfinite(x).all()
    vals = (1.0 / orc.attenuation_factor(mu, lors, um["center"], um["size"])).astype(np.float32)
    want = orc.mlem(vals, lors, emap, ic["center"], ic["size"], 15, driver="full")
    sel = emap > 0
    assert np.isfinite(want).all()
    err = float(np.linalg.norm((x - want)[sel]) / np.linalg.norm(want[sel]))
    print(f"doc_v7 description vs oracle MLEM: relative L2 {err:.3g}", flush=True)
    assert err <= 1e-4
