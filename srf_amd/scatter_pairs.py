"""Crystal-pair single-scatter correction: srfnef's own ScatterCorrect path.

The reference corrects a listmode in four steps
(srfnef/corrections/scattering/):
  1. scatter_preprocess (scatter.py:57-73): the crystal-pair sinogram of the
     listmode (lm2sino), the bins holding counts, scatter points every
     floor(50 mm / voxel) voxels (preprocess.py:6-18), emission and
     attenuation line integrals from every crystal to every scatter point
     (preprocess.py:21-34) and the attenuation of 1.3 mu along every counted
     crystal pair (preprocess.py:37-43); saved as .npy files;
  2. the per-pair single-scatter sum and the pair's scale (get_scatter_fraction,
     scatter_fraction/scatter_fraction.py:12-136);
  3. (sino - fraction / sum(fraction) * sum(sino) * scatter_fraction) * scale /
     (atten_lors + tiny) on the counted bins (scatter.py:51-53);
  4. the corrected bins back to a listmode of crystal-centre LORs (sino2lm,
     listmode2sinogram.py:88-99).

Here every step runs on the GPU: the sinogram by srf_pair_sinogram, the line
integrals by the projector (srf_project), the per-pair sum by the SSS kernels
(srf_sss_prepare / srf_sss_sinogram: crystals as detectors with their block's
normal angle, the reference's scatter points with mu looked up as the reference
does), the correction by srf_scatter_correct.  The per-pair scale and the pair
decoding are small elementwise steps in torch on the GPU.  The .npy files keep
the reference's names, dtypes and shapes, so either side's preprocessing can
feed the other's correction.

Arithmetic kept from the reference, quirks included:
  * Project * |p2 - p1|^2 for the emission integrals, exp(-Project * L * L /
    unit_size[0]) for attenuation (the reference scaling, exact only for 1 mm
    voxels, quirk Q15), 1.3 mu for the pair attenuation (quirk Q17);
  * scatter points placed from -size/2 and mu(S) = u_map[int((S / size + 0.5) *
    shape)], both assuming a zero-centred image (quirk Q16), the KN factor at
    the quantised angle used as a cosine (Q16);
  * a block's normal from get_block_theta (floor of the angle shifted by half a
    block), the obliquity ((D - S) . n_D) / |D - S| with its sign;
  * the constants of scatter_fraction.py:27-36 (window efficiency without
    scatter, crystal area, 5 / sqrt(1022 pi) / resolution, 4 pi).
Deviation: when every pair's scatter sum is 0 (no attenuating material on the
scatter grid) the reference divides 0 by 0 and drops every event; here nothing
is subtracted (srf_scatter_correct with a zero sum).
"""
import math
import os

import numpy as np
import torch

from . import ops
from . import scanner as S
from ._lib import check, lib, stream_handle
from .data import Listmode, Lors
from .functions import device, image_to_kernel, to_gpu
from .sss import energy_tables, scatter_correct

FILES = ("lors", "sinogram", "index", "scatter_position", "crystal_position", "sumup_emission", "atten",
         "atten_lors")


def scatter_positions(shape, size):
    """(ns, 3) float32: preprocess.py:6-18 -- step = floor(50 / voxel) voxels,
    floor(shape / step) points per axis at step (k + 1/2) voxel - size / 2."""
    vox = np.asarray(size, np.float64) / np.asarray(shape, np.float64)
    step = np.floor(50.0 / vox)
    cnt = np.floor(np.asarray(shape, np.float64) / step).astype(int)
    g = np.meshgrid(*[np.arange(c, dtype=np.float64) for c in cnt], indexing="ij")
    pos = np.stack([step[i] * (g[i] + 0.5) * vox[i] for i in range(3)], -1).reshape(-1, 3)
    return (pos - np.asarray(size, np.float64) / 2).astype(np.float32)


def block_params(scanner):
    """(nb_blocks_per_ring, nb_crystals_per_ring, grid_block, size_block) as
    ScatterCorrect derives them (scatter.py:36-46)."""
    if isinstance(scanner, S.PetEcatScanner):
        b = scanner.blocks
        return (int(scanner.nb_blocks_per_ring), int(scanner.nb_blocks_per_ring * b.shape[1]),
                [int(v) for v in b.shape], [float(v) for v in b.size])
    if isinstance(scanner, S.PetCylindricalScanner):
        ny = scanner.nb_module[0] * scanner.nb_submodule[0] * scanner.nb_crystal[0]
        nz = scanner.nb_module[1] * scanner.nb_submodule[1] * scanner.nb_crystal[1]
        return (int(scanner.nb_rsector), int(scanner.nb_rsector * ny), [1, int(ny), int(nz)],
                [20.0, float(scanner.mv_module[0] * scanner.nb_module[0]),
                 float(scanner.mv_module[1] * scanner.nb_module[1])])
    raise NotImplementedError(type(scanner).__name__)


def block_normals(crystal_pos, nb_crystals_per_ring, nb_blocks_per_ring, grid_block_y):
    """Normal angle (rad) of the block holding each crystal: get_block_theta
    (scatter_fraction.py:122-131), in float64 on the host."""
    p = np.asarray(crystal_pos, np.float64)
    norm = np.sqrt(np.sum(p ** 2, axis=1))
    th = np.arccos(p[:, 0] / norm) / math.pi * 180.0
    th = np.where(p[:, 1] < 0, 360.0 - th, th)
    fixed = np.mod(th + 180.0 / nb_crystals_per_ring * grid_block_y, 360.0)
    return np.floor(fixed / 360.0 * nb_blocks_per_ring) / nb_blocks_per_ring * 2 * math.pi


def pairs_of_bins(bins):
    """Crystal pair (i, j), i > j, of sinogram bins k = i (i - 1) / 2 + j
    (get_all_lors_id's order, listmode2sinogram.py:27-35): int64 tensors."""
    k = bins.to(torch.int64)
    i = torch.floor((1.0 + torch.sqrt(1.0 + 8.0 * k.to(torch.float64))) / 2.0).to(torch.int64)
    i = torch.where(i * (i - 1) // 2 > k, i - 1, i)
    i = torch.where((i + 1) * i // 2 <= k, i + 1, i)
    return i, k - i * (i - 1) // 2


def _image_data(img):
    return to_gpu(img.data if not isinstance(img, torch.Tensor) else img)


def _project_len2(image, lors):
    """Project(image, lors) and |p2 - p1|^2 (float32, summed per axis as
    preprocess.py:29-31 does)."""
    data = _image_data(image)
    p = ops.project(lors, image_to_kernel(data), [int(v) for v in data.shape],
                    [float(v) for v in image.center] + [float(v) for v in image.size], False, 0.0, 0.0)
    d = lors[:, 3:6] - lors[:, 0:3]
    len2 = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]
    return p, len2


def _umap_project(u_map, lors, factor=1.0):
    """UmapProject (attenuation_correct.py:31-40) of factor * u_map:
    exp(-Project * L * L / unit_size[0])."""
    um = u_map if factor == 1.0 else u_map.update(data=_image_data(u_map) * np.float32(factor))
    p, len2 = _project_len2(um, lors)
    L = torch.sqrt(len2)
    unit = float(u_map.size[0]) / float(np.asarray(_image_data(u_map).shape)[0])
    return torch.exp(-(p * L * L / unit))


def _crystal_point_lors(crystal_pos, points):
    """(N * ns, 6) LORs from every crystal (p1) to every scatter point (p2),
    crystal-major (get_lors, preprocess.py:46-54)."""
    n, ns = crystal_pos.shape[0], points.shape[0]
    a = crystal_pos[:, None, :].expand(n, ns, 3)
    b = points[None, :, :].expand(n, ns, 3)
    return torch.cat([a, b], -1).reshape(-1, 6).contiguous()


def scatter_preprocess(scanner, listmode, emission_image, umap, outdir=None, out_append=None):
    """scatter.py:57-73 on the GPU.  Returns a dict of GPU tensors with the
    reference's keys (lors, the (M, 2) pair table, only when `outdir` is given);
    with `outdir` the same arrays are also written as <key><out_append>.npy in
    the reference's dtypes and shapes."""
    from .data_transform import crystal_table, lm2sino
    dev = device()
    ncrys = int(scanner.nb_crystals)
    sino = lm2sino(listmode, scanner)                                    # (M,) float64
    index = torch.nonzero(sino > 0).reshape(-1).to(torch.int32)
    shape = [int(v) for v in np.asarray(_image_data(emission_image).shape)]
    pts = torch.from_numpy(scatter_positions(shape, emission_image.size)).to(dev)
    cpos = crystal_table(scanner, dev)
    lors_cp = _crystal_point_lors(cpos, pts)
    p, len2 = _project_len2(emission_image, lors_cp)
    sumup = (p * len2).contiguous()
    atten = _umap_project(umap, lors_cp).contiguous()
    i, j = pairs_of_bins(index)
    pair_lors = torch.cat([cpos[i], cpos[j]], 1).contiguous()
    atten_lors = _umap_project(umap, pair_lors, 1.3).contiguous()
    pre = {"sinogram": sino, "index": index, "scatter_position": pts, "crystal_position": cpos,
           "sumup_emission": sumup, "atten": atten, "atten_lors": atten_lors}
    if outdir is not None:
        ap = "" if out_append is None else str(out_append)
        os.makedirs(outdir, exist_ok=True)
        from .data_transform import get_all_lors_id
        arrays = {"lors": get_all_lors_id(ncrys), "sinogram": sino.cpu().numpy().reshape(-1, 1),
                  "index": index.cpu().numpy(), "scatter_position": pts.cpu().numpy(),
                  "crystal_position": cpos.cpu().numpy(), "sumup_emission": sumup.cpu().numpy(),
                  "atten": atten.cpu().numpy(), "atten_lors": atten_lors.cpu().numpy().reshape(-1, 1)}
        for k in FILES:
            np.save(os.path.join(outdir, f"{k}{ap}.npy"), arrays[k])
    return pre


def load_preprocessed(outdir, out_append=None):
    """The files scatter_preprocess wrote (this module's or the reference's),
    read without pickles, as GPU tensors (scatter.py:28-35)."""
    ap = "" if out_append is None else str(out_append)
    dev = device()
    pre = {}
    for k in FILES:
        if k == "lors":
            continue     # the pair of a bin is decoded from the bin index
        a = np.load(os.path.join(outdir, f"{k}{ap}.npy"), allow_pickle=False)
        pre[k] = torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    pre["sinogram"] = pre["sinogram"].reshape(-1).to(torch.float64)
    pre["index"] = pre["index"].reshape(-1).to(torch.int32)
    pre["atten_lors"] = pre["atten_lors"].reshape(-1).to(torch.float32)
    for k in ("scatter_position", "crystal_position", "sumup_emission", "atten"):
        pre[k] = pre[k].to(torch.float32).contiguous()
    return pre


def pair_scatter(pre, u_map, scanner, low_energy, high_energy, resolution):
    """get_scatter_fraction (scatter_fraction.py:12-37): (fraction, scale), one
    float32 value per counted bin of pre['index']."""
    dev = pre["crystal_position"].device
    nbr, ncr, grid, size_block = block_params(scanner)
    cpos = pre["crystal_position"]
    pts = pre["scatter_position"]
    ncrys, ns = int(cpos.shape[0]), int(pts.shape[0])
    normals = torch.from_numpy(block_normals(cpos.cpu().numpy(), ncr, nbr, grid[1])).to(dev)
    det = torch.cat([cpos, normals.to(torch.float32)[:, None]], 1).contiguous()
    # mu(S) as kn_formula.py:12-14 reads it (a zero-centred image)
    mu = _image_data(u_map)
    shp = list(mu.shape)
    size = [float(v) for v in u_map.size]
    idx = [torch.trunc((pts[:, k].double() / size[k] + 0.5) * shp[k]).long() for k in range(3)]
    inside = (idx[0] >= 0) & (idx[0] < shp[0]) & (idx[1] >= 0) & (idx[1] < shp[1]) & (idx[2] >= 0) & (
        idx[2] < shp[2])
    if not bool(inside.all()):
        raise ValueError("a scatter point lies outside the u_map (its grid must be zero-centred, quirk Q16)")
    mu_s = mu[idx[0], idx[1], idx[2]]
    points = torch.cat([pts, mu_s[:, None]], 1).contiguous()
    em = pre["sumup_emission"].reshape(ncrys, ns).contiguous()
    with torch.no_grad():
        mu_int = (-torch.log(pre["atten"].reshape(ncrys, ns))).contiguous()
    i, j = pairs_of_bins(pre["index"])
    bins = torch.stack([i, j], 1).to(torch.int32).contiguous()
    tables = torch.from_numpy(energy_tables(low_energy, high_energy, resolution)).to(dev)
    sigma = math.sqrt(511.0) * resolution
    ei = np.arange(int(low_energy), int(high_energy), 5, dtype=np.float64)
    e_ws = float(np.sum(np.exp(-(ei - 511.0) ** 2 / 2.0 / sigma ** 2))) * 5.0 / math.sqrt(1022.0 * math.pi) \
        / resolution
    area = size_block[1] * size_block[2] / grid[1] / grid[2]
    const = e_ws * area ** 2 * 5.0 / math.sqrt(1022.0 * math.pi) / resolution / 4 / math.pi
    n = int(bins.shape[0])
    fraction = torch.zeros(n, dtype=torch.float32, device=dev)
    if n > 0 and ns > 0:
        rows = torch.empty(int(lib().srf_sss_rows_bytes(ncrys, ns)) // 4 + 4, dtype=torch.float32, device=dev)
        sref = stream_handle(dev)
        check(lib().srf_sss_prepare(det.data_ptr(), ncrys, points.data_ptr(), ns, em.data_ptr(), mu_int.data_ptr(),
                                    rows.data_ptr(), rows.numel() * 4, sref), "srf_sss_prepare")
        check(lib().srf_sss_sinogram(rows.data_ptr(), ncrys, points.data_ptr(), ns, bins.data_ptr(), n,
                                     tables.data_ptr(), float(const), fraction.data_ptr(), sref), "srf_sss_sinogram")
    # get_scale (scatter_fraction.py:64-67) with the constants of :34-36
    A, B = cpos[i].double(), cpos[j].double()
    nb = normals[j]
    dab = torch.sqrt(((B - A) ** 2).sum(1))
    area_ab = ((B[:, 0] - A[:, 0]) * torch.cos(nb) + (B[:, 1] - A[:, 1]) * torch.sin(nb)) / dab
    scale = ((dab / area_ab) ** 2 * 4 * math.pi / e_ws ** 2 / area ** 2).to(torch.float32)
    return fraction, scale


def correct_pairs(pre, fraction, scale, scatter_fraction):
    """scatter.py:51-55: the corrected counted bins (float32, by
    srf_scatter_correct: y = sino[index], weight = scale / atten_lors) and the
    listmode of the bins with 0 < value < 1e8 (sino2lm): crystal-centre LORs
    (M, 7), the value repeated in column 6 as the reference's Lors(corrected_data)
    holds it."""
    index = pre["index"].to(torch.int64)
    y = pre["sinogram"][index].to(torch.float32).contiguous()
    weight = (scale / pre["atten_lors"]).contiguous()
    n = int(index.numel())
    bins = torch.arange(n, dtype=torch.int32, device=y.device)
    corr = scatter_correct(y, weight, bins, fraction.contiguous(), scatter_fraction) if n else y
    keep = (corr > 0) & (corr < 1e8)
    i, j = pairs_of_bins(index[keep])
    cpos = pre["crystal_position"]
    vals = corr[keep]
    lors = torch.cat([cpos[i], cpos[j], vals[:, None]], 1).contiguous()
    return corr, Listmode(vals, Lors(lors))


__all__ = ["scatter_positions", "block_params", "block_normals", "pairs_of_bins", "scatter_preprocess",
           "load_preprocessed", "pair_scatter", "correct_pairs"]
