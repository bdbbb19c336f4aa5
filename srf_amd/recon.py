"""End-to-end reconstruction from srfnef's JSON recon description, and its
listmode inputs.

- The JSON schema of usage_docs/doc_v7 (an MlemFull: n_iter, image_config,
  scanner, listmode, atten_corr, scatter_corr, psf_corr, emap), read as
  srfnef's json_load does (io/json_io_mixin.py:48-52, nested from_dict);
  recon_full builds and runs it (app/app_test_full.py:30-45,
  app_test_full_tof.py for TOF).
- Array fields hold a path.  The reference stores full objects as deepdish
  HDF5 (io/local_io_mixin.py:45-61); deepdish and h5py are not available
  here, so paths to .npy / .npz files are read (NumPy, no pickles) and an
  HDF5 path raises with that reason.
- listmode_from_gate_out (adapters/listmode_from_gate_out.py:36-64): GATE
  coincidence arrays (N, >= 7) from nb_sub files, endpoints snapped to their
  crystals on the GPU (srf_amd.data_transform).
"""
import json
import os

import numpy as np

from . import scanner as S
from .corrections import AttenuationCorrect
from .data import Emap, ImageConfig, Listmode, Lors, TofConfig, Umap
from .functions import MlemFull, MlemFullTof


def load_array(path, key=None):
    """An array from a .npy file or an .npz archive (key, or its only / 'data'
    entry).  No pickled objects are loaded."""
    ext = os.path.splitext(str(path))[1].lower()
    if ext in (".hdf5", ".h5"):
        raise NotImplementedError(f"{path}: HDF5 object stores need deepdish / h5py, which are not installed; "
                                  "save the array as .npy or .npz")
    if ext == ".npy":
        return np.load(path, allow_pickle=False)
    if ext == ".npz":
        with np.load(path, allow_pickle=False) as z:
            if key is not None:
                return z[key]
            if "data" in z.files:
                return z["data"]
            if len(z.files) == 1:
                return z[z.files[0]]
            raise ValueError(f"{path}: several arrays ({z.files}); name one")
    raise ValueError(f"{path}: expected .npy or .npz")


def scanner_from_dict(d):
    """PetEcatScanner (fields of geometry/pet_ecat_scanner.py:17-26) or
    PetCylindricalScanner (pet_cylindrical_scanner.py:32-49) from a dict; the
    reference's extra 'center' key is accepted and ignored."""
    if d is None:
        return None
    tof = d.get("tof")
    tof = TofConfig(**tof) if isinstance(tof, dict) else None
    if "nb_blocks_per_ring" in d:
        b = d["blocks"]
        blocks = S.Block(list(b["size"]), list(b["shape"]), list(b.get("interval") or [0, 0, 0]))
        return S.PetEcatScanner(float(d["inner_radius"]), float(d["outer_radius"]), int(d["nb_rings"]),
                                int(d["nb_blocks_per_ring"]), float(d["gap"]), blocks, tof)
    if "nb_rsector" in d:
        keys = ["nb_module", "mv_module", "nb_submodule", "mv_submodule", "nb_crystal", "mv_crystal", "sz_crystal"]
        return S.PetCylindricalScanner(float(d["inner_radius"]), float(d["outer_radius"]), int(d["nb_rsector"]),
                                       *[list(d[k]) for k in keys], tof)
    raise ValueError("unknown scanner description")


def listmode_from_gate_out(path, scanner, nb_sub):
    """GATE listmode: the first 7 columns of files path.replace('?', str(i)),
    i < nb_sub, stacked; endpoints snapped to the crystal centres, values 1,
    the 6 position columns kept (as the reference does)."""
    from .data_transform import snap_lors
    parts = [np.asarray(np.load(path.replace("?", str(i)), allow_pickle=False)[:, :7], np.float32)
             for i in range(int(nb_sub))]
    full = np.ascontiguousarray(np.vstack(parts))
    lors, _ = snap_lors(scanner, full[:, :6])
    return Listmode(np.ones((full.shape[0],), np.float32), Lors(lors.cpu().numpy()))


def listmode_from(spec, scanner=None):
    """A listmode field: a path to an .npz with 'data' and 'lors' arrays, or to
    a (N, >= 6) LOR array (.npy; values 1), or a dict {'data': ..., 'lors':
    ...} of paths / arrays."""
    if isinstance(spec, dict):
        data = spec.get("data")
        lors = spec["lors"]
        lors = load_array(lors) if isinstance(lors, str) else np.asarray(lors)
        if isinstance(lors, dict):
            lors = lors["data"]
        data = load_array(data) if isinstance(data, str) else data
        data = np.ones(lors.shape[0], np.float32) if data is None else np.asarray(data, np.float32)
        return Listmode(data.reshape(-1), Lors(np.ascontiguousarray(lors, np.float32)))
    path = str(spec)
    if path.lower().endswith(".npz"):
        with np.load(path, allow_pickle=False) as z:
            lors = np.ascontiguousarray(z["lors"], np.float32)
            data = z["data"] if "data" in z.files else np.ones(lors.shape[0], np.float32)
        return Listmode(np.asarray(data, np.float32).reshape(-1), Lors(lors))
    lors = np.ascontiguousarray(load_array(path), np.float32)
    return Listmode(np.ones(lors.shape[0], np.float32), Lors(lors))


def _image_field(d, cls):
    if d is None:
        return None
    data = d["data"]
    data = load_array(data) if isinstance(data, str) else np.asarray(data, np.float32)
    return cls(np.asarray(data, np.float32), list(d["center"]), list(d["size"]))


def recon_from_dict(cfg, tof=None):
    """The MlemFull (MlemFullTof with tof, or when the scanner has a TOF
    configuration) a recon description describes."""
    ic = cfg["image_config"]
    image_config = ImageConfig(list(ic["shape"]), list(ic["center"]), list(ic["size"]))
    scanner = scanner_from_dict(cfg["scanner"])
    listmode = listmode_from(cfg["listmode"], scanner)
    atten = cfg.get("atten_corr")
    atten = AttenuationCorrect(_image_field(atten["u_map"], Umap)) if atten else None
    if cfg.get("scatter_corr"):
        raise NotImplementedError("scatter_corr in a recon description: use srf_amd.sss.ScatterCorrect directly")
    psf = cfg.get("psf_corr")
    if psf:
        from .psf import Deconvolute, FittedX, FittedY, FittedZ, PsfCorrect
        dec = psf.get("deconv")
        if not dec:
            raise NotImplementedError("psf_corr needs fitted tables (deconv with fitted_x / fitted_y / fitted_z "
                                      "array paths); fitting point-source images is not on this path")
        fx, fy, fz = (load_array(dec[k]) if isinstance(dec[k], str) else np.asarray(dec[k])
                      for k in ("fitted_x", "fitted_y", "fitted_z"))
        psf = PsfCorrect(psf.get("pnt_img_path"), Deconvolute(int(dec.get("n_iter", 10)), FittedX(fx), FittedY(fy),
                                                              FittedZ(fz)))
    emap = _image_field(cfg.get("emap"), Emap)
    use_tof = tof if tof is not None else (scanner is not None and scanner.tof is not None)
    cls = MlemFullTof if use_tof else MlemFull
    return cls(int(cfg["n_iter"]), image_config, scanner, listmode, atten, None, psf, emap)


def json_load(path):
    with open(path) as f:
        return json.load(f)


def recon_full(cfg, out_path=None, tof=None):
    """Run a recon description (a dict, or a path to its JSON file); the image
    is returned and, with out_path, saved as .npy."""
    if not isinstance(cfg, dict):
        cfg = json_load(cfg)
    image = recon_from_dict(cfg, tof)()
    if out_path:
        np.save(out_path, np.asarray(image.data, np.float32))
    return image


__all__ = ["load_array", "scanner_from_dict", "listmode_from_gate_out", "listmode_from", "recon_from_dict",
           "json_load", "recon_full"]
