"""Listmode <-> crystal level on the GPU: snapping LOR endpoints to crystal
centres, crystal-pair sinograms and back.

Mirrors srfnef's
- functions/data_transform/pet_ecat.py and pet_cylindrical.py
  (Pet*ListmodeTrans, Pet*ListmodeToSinogram, Pet*SinogramToListmode);
- corrections/scattering/listmode2sinogram.py (get_all_lors_id, lm2sino,
  sino2lm), the crystal-pair sinogram of its ScatterCorrect.
The per-event work (index of every endpoint, snapping, binning) and the
per-bin work (selecting the bins to list) run in libsrf_hip.so
(srf_crystal_index, srf_snap_lors, srf_pair_sinogram, srf_sinogram_lors); the
crystal-centre table comes from srf_amd.scanner.
"""
import ctypes

import numpy as np
import torch

from . import scanner as S
from ._lib import SrfCyl, SrfEcat, check, lib, stream_handle
from .data import Listmode, Lors, nef_class
from .functions import device, to_gpu


def _params(scanner):
    """(kind, ecat, cyl) arguments of srf_crystal_index for a scanner."""
    e, c = SrfEcat(), SrfCyl()
    if isinstance(scanner, S.PetEcatScanner):
        b = scanner.blocks
        e.angle_per_block = scanner.angle_per_block
        e.axial_length = scanner.axial_length
        e.gap = scanner.gap
        e.block_size_y, e.block_size_z = b.size[1], b.size[2]
        e.unit_y, e.unit_z = b.unit_size[1], b.unit_size[2]
        e.nb_blocks_per_ring = scanner.nb_blocks_per_ring
        e.shape_y = b.shape[1]
        e.nb_crystals_per_block = scanner.nb_crystals_per_block
        e.nb_crystals_per_ring = scanner.nb_crystals_per_ring
        return 0, e, c
    if isinstance(scanner, S.PetCylindricalScanner):
        c.angle_per_rsector = scanner.angle_per_rsector
        for k in (0, 1):
            c.sz_rsector[k] = scanner.sz_rsector[k]
            c.mv_module[k] = scanner.mv_module[k]
            c.sz_module[k] = scanner.sz_module[k]
            c.mv_submodule[k] = scanner.mv_submodule[k]
            c.sz_submodule[k] = scanner.sz_submodule[k]
            c.mv_crystal[k] = scanner.mv_crystal[k]
            c.nb_crystal[k] = scanner.nb_crystal[k]
            c.nb_submodule[k] = scanner.nb_submodule[k]
            c.nb_module[k] = scanner.nb_module[k]
        c.nb_rsector = scanner.nb_rsector
        c.nb_crystal_per_submodule = scanner.nb_crystal_per_submodule
        c.nb_crystal_per_module = scanner.nb_crystal_per_module
        c.nb_crystal_per_rsector = scanner.nb_crystal_per_rsector
        return 1, e, c
    raise NotImplementedError(type(scanner).__name__)


def _lors_gpu(lors):
    L = to_gpu(lors.data if isinstance(lors, Lors) else lors)
    if L.dim() != 2 or L.shape[1] < 6:
        raise ValueError("LORs must be (N, >=6)")
    return L.contiguous()


def crystal_table(scanner, dev=None):
    """Crystal centres (nb_crystals, 3) float32 on the GPU."""
    return torch.from_numpy(S.crystal_positions(scanner)).to(dev or device())


def crystal_indices(scanner, lors):
    """(N, 2) int64 GPU tensor: the crystal index of each LOR endpoint
    (srf_crystal_index)."""
    L = _lors_gpu(lors)
    n = int(L.shape[0])
    ind = torch.empty((n, 2), dtype=torch.int64, device=L.device)
    kind, e, c = _params(scanner)
    with torch.cuda.device(L.device):
        check(lib().srf_crystal_index(L.data_ptr(), n, int(L.shape[1]), kind, ctypes.byref(e), ctypes.byref(c),
                                      ind.data_ptr(), stream_handle(L.device)), "srf_crystal_index")
    return ind


def snap_lors(scanner, lors, ind=None, table=None):
    """LORs with both endpoints moved to their crystals' centres (extra
    columns kept); returns (lors, count of LORs with an endpoint off the
    crystal table, which become NaN)."""
    L = _lors_gpu(lors)
    n = int(L.shape[0])
    ind = crystal_indices(scanner, L) if ind is None else ind
    table = crystal_table(scanner, L.device) if table is None else table
    out = torch.empty_like(L)
    outside = torch.zeros(1, dtype=torch.int64, device=L.device)
    with torch.cuda.device(L.device):
        check(lib().srf_snap_lors(L.data_ptr(), n, int(L.shape[1]), ind.data_ptr(), table.data_ptr(),
                                  int(table.shape[0]), out.data_ptr(), outside.data_ptr(),
                                  stream_handle(L.device)), "srf_snap_lors")
    return out, int(outside.item())


@nef_class
class PetEcatListmodeTrans:
    """Listmode with LOR endpoints snapped to the ECAT crystal centres
    (functions/data_transform/pet_ecat.py:43-60)."""
    scanner: S.PetEcatScanner

    def __call__(self, listmode: Listmode) -> Listmode:
        out, _ = snap_lors(self.scanner, listmode.lors)
        return listmode.update(lors=Lors(out))


@nef_class
class PetCylindricalListmodeTrans:
    """Listmode with LOR endpoints snapped to the cylindrical scanner's crystal
    centres (functions/data_transform/pet_cylindrical.py:43-60)."""
    scanner: S.PetCylindricalScanner

    def __call__(self, listmode: Listmode) -> Listmode:
        out, _ = snap_lors(self.scanner, listmode.lors)
        return listmode.update(lors=Lors(out))


@nef_class
class PetListmodeTrans:
    """Dispatch on the scanner type (functions/data_transform/__init__.py:40-47).
    A scanner without crystals (an ideal cylinder) has nothing to snap to: the
    listmode is returned unchanged (the reference returns None there)."""
    scanner: object

    def __call__(self, listmode: Listmode) -> Listmode:
        if isinstance(self.scanner, S.PetEcatScanner):
            return PetEcatListmodeTrans(self.scanner)(listmode)
        if isinstance(self.scanner, S.PetCylindricalScanner):
            return PetCylindricalListmodeTrans(self.scanner)(listmode)
        return listmode


@nef_class
class Sinogram:
    """Crystal-pair sinogram with its scanner (data/sinogram.py:14-16)."""
    data: object
    scanner: object


def _coo(scanner, listmode):
    from scipy.sparse import coo_matrix
    ind = crystal_indices(scanner, listmode.lors).cpu().numpy()
    w = np.asarray(listmode.data.cpu() if torch.is_tensor(listmode.data) else listmode.data).reshape(-1)
    n = int(scanner.nb_crystals)
    return coo_matrix((w, (ind[:, 0], ind[:, 1])), shape=(n, n))


def _from_coo(scanner, sino):
    """One LOR per stored nonzero entry, duplicates not summed, as the
    reference's nonzero() / .data pair; it pairs the nonzero coordinates with
    every stored value, which only line up when no stored value is 0 -- the
    values of the kept entries are taken here."""
    m = sino.data.tocoo()
    keep = m.data != 0
    pos = S.crystal_positions(scanner)
    lors = np.hstack((pos[m.row[keep]], pos[m.col[keep]]))
    return Listmode(m.data[keep], Lors(lors))


@nef_class
class PetEcatListmodeToSinogram:
    """(nb_crystals x nb_crystals) sparse sinogram of (first, second) crystal
    (functions/data_transform/pet_ecat.py:20-31); indices on the GPU."""
    scanner: S.PetEcatScanner

    def __call__(self, listmode: Listmode) -> Sinogram:
        return Sinogram(_coo(self.scanner, listmode), self.scanner)


@nef_class
class PetEcatSinogramToListmode:
    """Nonzero entries of a sparse sinogram as a listmode
    (functions/data_transform/pet_ecat.py:9-17)."""
    scanner: S.PetEcatScanner

    def __call__(self, sino: Sinogram) -> Listmode:
        return _from_coo(self.scanner, sino)


@nef_class
class PetCylindricalListmodeToSinogram:
    scanner: S.PetCylindricalScanner

    def __call__(self, listmode: Listmode) -> Sinogram:
        return Sinogram(_coo(self.scanner, listmode), self.scanner)


@nef_class
class PetCylindricalSinogramToListmode:
    scanner: S.PetCylindricalScanner

    def __call__(self, sino: Sinogram) -> Listmode:
        return _from_coo(self.scanner, sino)


def get_all_lors_id(total_num_crystal):
    """(N (N - 1) / 2, 2) int32 crystal pairs (i, j), i > j, in bin order
    (listmode2sinogram.py:27-35)."""
    n = int(total_num_crystal)
    i = np.repeat(np.arange(n, dtype=np.int64), np.arange(n, dtype=np.int64))
    start = np.repeat(np.arange(n, dtype=np.int64) * (np.arange(n, dtype=np.int64) - 1) // 2,
                      np.arange(n, dtype=np.int64))
    j = np.arange(n * (n - 1) // 2, dtype=np.int64) - start
    return np.stack([i, j], 1).astype(np.int32)


def lm2sino(listmode, scanner, return_outside=False):
    """Crystal-pair sinogram (N (N - 1) / 2,) float64 on the GPU of a listmode
    (listmode2sinogram.py:76-85, which returns it as an (M, 1) array)."""
    L = _lors_gpu(listmode.lors)
    n = int(L.shape[0])
    ncrys = int(scanner.nb_crystals)
    ind = crystal_indices(scanner, L)
    w = listmode.data
    w = None if w is None else to_gpu(w).reshape(-1).contiguous()
    sino = torch.zeros(ncrys * (ncrys - 1) // 2, dtype=torch.float64, device=L.device)
    outside = torch.zeros(1, dtype=torch.int64, device=L.device)
    with torch.cuda.device(L.device):
        check(lib().srf_pair_sinogram(ind.data_ptr(), None if w is None else w.data_ptr(), n, ncrys,
                                      sino.data_ptr(), outside.data_ptr(), stream_handle(L.device)),
              "srf_pair_sinogram")
    return (sino, int(outside.item())) if return_outside else sino


def sino2lm(scanner, sinogram, lors=None):
    """(M, 7) float64: for every bin with 0 < value < 1e8, in bin order, the
    LOR of its crystal pair and the value (listmode2sinogram.py:88-99; the
    pair of a bin is decoded directly, so `lors` -- get_all_lors_id -- is not
    needed)."""
    s = to_gpu(sinogram, torch.float64).reshape(-1).contiguous()
    ncrys = int(scanner.nb_crystals)
    if s.numel() != ncrys * (ncrys - 1) // 2:
        raise ValueError("sinogram must have N (N - 1) / 2 bins")
    table = crystal_table(scanner, s.device)
    cnt = ctypes.c_int64()
    with torch.cuda.device(s.device):
        check(lib().srf_sinogram_lors(s.data_ptr(), ncrys, table.data_ptr(), None, None, 0, ctypes.byref(cnt),
                                      stream_handle(s.device)), "srf_sinogram_lors")
        m = int(cnt.value)
        lo = torch.empty((max(m, 1), 6), dtype=torch.float32, device=s.device)
        vals = torch.empty(max(m, 1), dtype=torch.float64, device=s.device)
        check(lib().srf_sinogram_lors(s.data_ptr(), ncrys, table.data_ptr(), lo.data_ptr(), vals.data_ptr(), m,
                                      ctypes.byref(cnt), stream_handle(s.device)), "srf_sinogram_lors")
    out = np.zeros((m, 7), np.float64)
    if m:
        out[:, :6] = lo[:m].cpu().numpy()
        out[:, 6] = vals[:m].cpu().numpy()
    return out


__all__ = ["crystal_table", "crystal_indices", "snap_lors", "PetEcatListmodeTrans", "PetCylindricalListmodeTrans",
           "PetListmodeTrans",
           "Sinogram", "PetEcatListmodeToSinogram", "PetEcatSinogramToListmode", "PetCylindricalListmodeToSinogram",
           "PetCylindricalSinogramToListmode", "get_all_lors_id", "lm2sino", "sino2lm"]
