"""Crystal-level scanner geometry: block (ECAT) and cylindrical (rsector /
module / submodule / crystal) PET systems, crystal index <-> position.

The classes keep srfnef's names and fields:
- Block (srfnef/geometry/block.py:15-23), PetEcatScanner
  (geometry/pet_ecat_scanner.py:17-26) with the properties of
  ops/geometry/pet_ecat_scanner.py:13-47;
- PetCylindricalScanner (geometry/pet_cylindrical_scanner.py:32-49) with the
  properties of ops/geometry/pet_cylindrical_scanner.py:13-74;
- the presets of geometry/existing_geometry.py:15-33.

Index <-> position follows functions/geometry/pet_ecat_scanner.py:15-89 and
functions/geometry/pet_cylindrical_scanner.py:15-160: float64 arithmetic,
positions returned as float32 like the reference's arrays.  These are
per-crystal tables (at most ~1e5 crystals), built once per scanner on the host;
what runs per event or per crystal pair (listmode snapping, pair LOR
generation, the backprojection of all pairs) is in the HIP library.

Reference quirks kept:
- CylindricalCrystalPosToIndex strides the module's z index by nb_module[1]
  (pet_cylindrical_scanner.py:37-38) where CylindricalIndexToCrystalPos uses
  nb_module[0] (:132-136); the two agree only when nb_module[0] ==
  nb_module[1] (as in both presets).
- Block.interval is stored but never used by the geometry (block.py:18-23).
"""
import math

import numpy as np

from .data import TofConfig, nef_class


@nef_class
class Block:
    """Crystal block: size (mm) and crystal grid shape, [x, y, z]."""
    size: list
    shape: list
    interval: list

    def __post_init__(self):
        if self.interval is None:
            object.__setattr__(self, "interval", [0, 0, 0])

    @property
    def unit_size(self):
        return [s1 / s2 for s1, s2 in zip(self.size, self.shape)]


@nef_class
class PetEcatScanner:
    """Ring of nb_blocks_per_ring blocks, nb_rings rings with an axial gap."""
    inner_radius: float
    outer_radius: float
    nb_rings: int
    nb_blocks_per_ring: int
    gap: float
    blocks: Block
    tof: TofConfig

    @property
    def average_radius(self):
        return (self.inner_radius + self.outer_radius) / 2

    @property
    def axial_length(self):
        return self.blocks.size[2] * self.nb_rings + self.gap * (self.nb_rings - 1)

    @property
    def nb_crystals_per_block(self):
        return self.blocks.shape[1] * self.blocks.shape[2]

    @property
    def nb_crystals_per_ring(self):
        return self.nb_crystals_per_block * self.nb_blocks_per_ring

    @property
    def nb_crystals(self):
        return self.nb_crystals_per_ring * self.nb_rings

    @property
    def angle_per_block(self):
        return 2 * math.pi / self.nb_blocks_per_ring

    @property
    def nb_thin_rings(self):
        return self.nb_rings * self.blocks.shape[2]

    @property
    def nb_crystals_per_thin_ring(self):
        return self.blocks.shape[1] * self.nb_blocks_per_ring


@nef_class
class PetCylindricalScanner:
    """One-layer cylindrical system: nb_rsector rotational sectors, each a grid
    of modules, submodules and crystals ([transaxial, z] per level); the
    sectors are numbered anticlockwise from +x."""
    inner_radius: float
    outer_radius: float
    nb_rsector: int
    nb_module: list
    mv_module: list
    nb_submodule: list
    mv_submodule: list
    nb_crystal: list
    mv_crystal: list
    sz_crystal: list
    tof: TofConfig

    @property
    def average_radius(self):
        return (self.inner_radius + self.outer_radius) / 2

    @property
    def angle_per_rsector(self):
        return 2 * math.pi / self.nb_rsector

    @property
    def gap_crystal(self):
        return [mv - sz for mv, sz in zip(self.mv_crystal, self.sz_crystal)]

    @property
    def sz_submodule(self):
        return [self.sz_crystal[k] + self.mv_crystal[k] * (self.nb_crystal[k] - 1) for k in (0, 1)]

    @property
    def gap_submodule(self):
        return [mv - sz for mv, sz in zip(self.mv_submodule, self.sz_submodule)]

    @property
    def sz_module(self):
        return [self.sz_submodule[k] + self.mv_submodule[k] * (self.nb_submodule[k] - 1) for k in (0, 1)]

    @property
    def gap_module(self):
        return [mv - sz for mv, sz in zip(self.mv_module, self.sz_module)]

    @property
    def sz_rsector(self):
        return [self.sz_module[k] + self.mv_module[k] * (self.nb_module[k] - 1) for k in (0, 1)]

    @property
    def axial_length(self):
        return self.sz_rsector[1]

    @property
    def nb_crystal_per_submodule(self):
        return self.nb_crystal[0] * self.nb_crystal[1]

    @property
    def nb_submodule_per_module(self):
        return self.nb_submodule[0] * self.nb_submodule[1]

    @property
    def nb_module_per_rsector(self):
        return self.nb_module[0] * self.nb_module[1]

    @property
    def nb_crystal_per_module(self):
        return self.nb_crystal_per_submodule * self.nb_submodule_per_module

    @property
    def nb_crystal_per_rsector(self):
        return self.nb_crystal_per_module * self.nb_module_per_rsector

    @property
    def nb_crystals(self):
        return self.nb_crystal_per_rsector * self.nb_rsector


def _positions(pos):
    """(N, 3) positions in their own floating type.  Listmode LORs are float32,
    and the reference's arithmetic on them stays float32 where no float64 array
    enters (the block / sector angle from arctan2, the ring from the axial
    coordinate); the in-block corrections are float64."""
    pos = np.asarray(pos)
    if pos.dtype.kind != "f":
        pos = pos.astype(np.float64)
    if pos.ndim == 1:
        pos = pos.reshape((1, pos.size))
    return pos


# ----------------------------------------------------------------- ECAT --

@nef_class
class EcatIndexToCrystalPos:
    """Crystal index (or the (iy, iz, iblock, iring) tuple with verbose) ->
    crystal centre (N, 3) float32 (functions/geometry/pet_ecat_scanner.py:65-89)."""
    scanner: PetEcatScanner

    def __call__(self, ind, verbose=False):
        sc = self.scanner
        b = sc.blocks
        if verbose:
            iy, iz, iblock, iring = [np.asarray(v) for v in ind]
        else:
            ind = np.asarray(ind)
            iy = ind % b.shape[1]
            iz = (ind // b.shape[1]) % b.shape[2]
            iblock = (ind // b.shape[1] // b.shape[2]) % sc.nb_blocks_per_ring
            iring = ind // b.shape[1] // b.shape[2] // sc.nb_blocks_per_ring
        iy = np.atleast_1d(iy)
        pos = np.zeros((iy.size, 3), dtype=np.float32)
        x0 = sc.average_radius
        y0 = (iy + 0.5) * b.unit_size[1] - b.size[1] / 2
        theta = sc.angle_per_block * np.atleast_1d(iblock)
        pos[:, 0] = x0 * np.cos(theta) - y0 * np.sin(theta)
        pos[:, 1] = x0 * np.sin(theta) + y0 * np.cos(theta)
        pos[:, 2] = (np.atleast_1d(iz) + 0.5) * b.unit_size[2] + np.atleast_1d(iring) * (sc.gap + b.size[2]) \
            - sc.axial_length / 2
        return pos


@nef_class
class EcatCrystalPosToIndex:
    """Position (N, 3) -> crystal index, or (iy, iz, iblock, iring) with verbose
    (functions/geometry/pet_ecat_scanner.py:15-62)."""
    scanner: PetEcatScanner

    def __call__(self, pos, verbose=False):
        sc = self.scanner
        b = sc.blocks
        pos = _positions(pos)
        xc, yc, zc = pos[:, 0], pos[:, 1], pos[:, 2]
        iblock = np.round(np.arctan2(yc, xc) / sc.angle_per_block).astype(int) % sc.nb_blocks_per_ring
        iring = np.floor((zc + sc.axial_length / 2) / (sc.gap + b.size[2])).astype(int)
        angle = iblock * sc.angle_per_block
        y_corr = -xc * np.sin(angle) + yc * np.cos(angle) + b.size[1] / 2
        iy = np.round(y_corr / b.unit_size[1] - 0.5).astype(int)
        z_corr = zc + sc.axial_length / 2 - iring * (b.size[2] + sc.gap)
        iz = np.round(z_corr / b.unit_size[2] - 0.5).astype(int)
        if verbose:
            return iy, iz, iblock, iring
        return iy + iz * b.shape[1] + iblock * sc.nb_crystals_per_block + iring * sc.nb_crystals_per_ring


# ----------------------------------------------------------- cylindrical --

@nef_class
class CylindricalIndexToCrystalPos:
    """Crystal index (or the 7-tuple with verbose) -> crystal centre (N, 3)
    float32 (functions/geometry/pet_cylindrical_scanner.py:113-160)."""
    scanner: PetCylindricalScanner

    def __call__(self, ind, verbose=False):
        sc = self.scanner
        if verbose:
            cy, cz, sy, sz, my, mz, rs = [np.atleast_1d(np.asarray(v)) for v in ind]
        else:
            i = np.atleast_1d(np.asarray(ind))
            cy = i % sc.nb_crystal[0]
            i = i // sc.nb_crystal[0]
            cz = i % sc.nb_crystal[1]
            i = i // sc.nb_crystal[1]
            sy = i % sc.nb_submodule[0]
            i = i // sc.nb_submodule[0]
            sz = i % sc.nb_submodule[1]
            i = i // sc.nb_submodule[1]
            my = i % sc.nb_module[0]
            i = i // sc.nb_module[0]
            mz = i % sc.nb_module[1]
            rs = i // sc.nb_module[1]
        y = cy * sc.mv_crystal[0] + 0.5 * sc.sz_crystal[0] - sc.sz_submodule[0] / 2
        z = cz * sc.mv_crystal[1] + 0.5 * sc.sz_crystal[1] - sc.sz_submodule[1] / 2
        y = y + (sy * sc.mv_submodule[0] - sc.sz_module[0] / 2 + sc.sz_submodule[0] / 2)
        z = z + (sz * sc.mv_submodule[1] - sc.sz_module[1] / 2 + sc.sz_submodule[1] / 2)
        y = y + (my * sc.mv_module[0] - sc.sz_rsector[0] / 2 + sc.sz_module[0] / 2)
        z = z + (mz * sc.mv_module[1] - sc.sz_rsector[1] / 2 + sc.sz_module[1] / 2)
        x = sc.average_radius
        angle = sc.angle_per_rsector * rs
        pos = np.zeros((cy.size, 3), np.float32)
        pos[:, 0] = x * np.cos(angle) - y * np.sin(angle)
        pos[:, 1] = x * np.sin(angle) + y * np.cos(angle)
        pos[:, 2] = z
        return pos


@nef_class
class CylindricalCrystalPosToIndex:
    """Position (N, 3) -> crystal index, or the 7-tuple (crystal y, z,
    submodule y, z, module y, z, rsector) with verbose
    (functions/geometry/pet_cylindrical_scanner.py:15-110)."""
    scanner: PetCylindricalScanner

    def __call__(self, pos, verbose=False):
        sc = self.scanner
        pos = _positions(pos)
        rs = np.round(np.arctan2(pos[:, 1], pos[:, 0]) / sc.angle_per_rsector).astype(int) % sc.nb_rsector
        angle = rs * sc.angle_per_rsector
        y = -pos[:, 0] * np.sin(angle) + pos[:, 1] * np.cos(angle)
        z = pos[:, 2].astype(np.float64)
        my = np.floor((y + sc.sz_rsector[0] / 2) / sc.mv_module[0]).astype(int)
        mz = np.floor((z + sc.sz_rsector[1] / 2) / sc.mv_module[1]).astype(int)
        y = y - (my * sc.mv_module[0] - sc.sz_rsector[0] / 2 + sc.sz_module[0] / 2)
        z = z - (mz * sc.mv_module[1] - sc.sz_rsector[1] / 2 + sc.sz_module[1] / 2)
        sy = np.floor((y + sc.sz_module[0] / 2) / sc.mv_submodule[0]).astype(int)
        sz = np.floor((z + sc.sz_module[1] / 2) / sc.mv_submodule[1]).astype(int)
        y = y - (sy * sc.mv_submodule[0] - sc.sz_module[0] / 2 + sc.sz_submodule[0] / 2)
        z = z - (sz * sc.mv_submodule[1] - sc.sz_module[1] / 2 + sc.sz_submodule[1] / 2)
        cy = np.floor((y + sc.sz_submodule[0] / 2) / sc.mv_crystal[0]).astype(int)
        cz = np.floor((z + sc.sz_submodule[1] / 2) / sc.mv_crystal[1]).astype(int)
        if verbose:
            return cy, cz, sy, sz, my, mz, rs
        # module z strided by nb_module[1], as the reference does
        return cy + cz * sc.nb_crystal[0] + (sy + sz * sc.nb_submodule[0]) * sc.nb_crystal_per_submodule \
            + (my + mz * sc.nb_module[1]) * sc.nb_crystal_per_module + rs * sc.nb_crystal_per_rsector


def crystal_positions(scanner):
    """Every crystal centre of a scanner, in index order: (nb_crystals, 3) float32."""
    ind = np.arange(int(scanner.nb_crystals))
    if isinstance(scanner, PetEcatScanner):
        return EcatIndexToCrystalPos(scanner)(ind)
    if isinstance(scanner, PetCylindricalScanner):
        return CylindricalIndexToCrystalPos(scanner)(ind)
    raise NotImplementedError(type(scanner).__name__)


def crystal_pos_to_index(scanner, pos):
    if isinstance(scanner, PetEcatScanner):
        return EcatCrystalPosToIndex(scanner)(pos)
    if isinstance(scanner, PetCylindricalScanner):
        return CylindricalCrystalPosToIndex(scanner)(pos)
    raise NotImplementedError(type(scanner).__name__)


# presets (geometry/existing_geometry.py:15-33)
scanner_1_4m = PetEcatScanner(400.0, 420.0, 26, 48, 3.42, Block([20.0, 51.3, 51.3], [1, 15, 15]))
scanner_mct = PetEcatScanner(424.5, 444.5, 4, 48, 0.0, Block([20.0, 53.3, 53.3], [1, 13, 13]))
scanner_8panel = PetCylindricalScanner(146.7, 166.7, 8, [2, 2], [60, 80], [3, 4], [20, 20], [6, 6], [3.2, 3.2],
                                       [3, 3])
scanner_20panel = PetCylindricalScanner(389.72, 411.9, 20, [2, 2], [60, 80], [3, 4], [20, 20], [6, 6],
                                        [3.2, 3.2], [3, 3])

__all__ = ["Block", "PetEcatScanner", "PetCylindricalScanner", "EcatIndexToCrystalPos", "EcatCrystalPosToIndex",
           "CylindricalIndexToCrystalPos", "CylindricalCrystalPosToIndex", "crystal_positions",
           "crystal_pos_to_index", "scanner_1_4m", "scanner_mct", "scanner_8panel", "scanner_20panel"]
