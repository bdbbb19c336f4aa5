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
