"""srf_amd -- MI355X-native list-mode PET reconstruction hot path.

A drop-in for SRF-NEF's projector / backprojector / MLEM / OSEM / attenuation
path (srfnef/functions, srfnef/corrections): the same class names and call
signatures, computed by hand-written gfx950 HIP kernels in libsrf_hip.so
through PyTorch custom ops (namespace ``srf``).
"""
from .data import Emap, Image, ImageConfig, Listmode, Lors, TofConfig, Umap, nef_class  # noqa: F401

__version__ = "0.1.0"

_LAZY = {
    "Project", "ProjectTof", "BackProject", "BackProjectTof", "Mlem", "MlemFull", "MlemFullTof", "OsemFull",
    "OsemTof", "EmapGenerator", "CylinderScanner", "AttenuationCorrect", "UmapProject",
}


def __getattr__(name):
    # operators import torch and load libsrf_hip.so on first use
    if name in _LAZY:
        from . import corrections, functions
        for mod in (functions, corrections):
            if hasattr(mod, name):
                return getattr(mod, name)
    raise AttributeError(name)
