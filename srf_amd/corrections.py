"""Attenuation correction with srfnef's names (the scatter estimate lives in
srf_amd.sss).

  AttenuationCorrect(u_map)(listmode) -> Listmode
      srfnef/corrections/attenuation/attenuation_correct.py:16-28
  UmapProject()(u_map, lors) -> Listmode
      srfnef/corrections/attenuation/attenuation_correct.py:31-40

Both compute a = exp(-Project(u) * L^2 / unit_size[0]) on the GPU.  Project
returns dx_^2/L^2 * integral(u), so the exponent is dx_ * integral(u): the
reference's scaling, exact only for 1 mm voxels.  ``physical=True`` divides by
unit_size[0] once more and uses integral(u) itself.
"""
import torch

from .data import Listmode, Umap, nef_class
from .functions import _out, project_data, to_gpu


def attenuation_factors(u_map, lors_data, physical=False):
    """GPU tensor a[e] = exp(-exponent[e])."""
    lors = to_gpu(lors_data)
    p = project_data(u_map, lors[:, :6].contiguous())
    d = lors[:, 0:3] - lors[:, 3:6]
    L = torch.sqrt(d[:, 0] ** 2 + d[:, 1] ** 2 + d[:, 2] ** 2)
    unit = float(u_map.unit_size[0])
    u = p * L * L / unit
    if physical:
        u = u / unit
    return torch.exp(-u)


@nef_class
class AttenuationCorrect:
    u_map: Umap
    physical: bool = False

    def __call__(self, listmode: Listmode) -> Listmode:
        a = attenuation_factors(self.u_map, listmode.lors.data, self.physical)
        if isinstance(listmode.data, torch.Tensor):
            return listmode.update(data=to_gpu(listmode.data) / a)
        return listmode.update(data=(to_gpu(listmode.data) / a).cpu().numpy())


@nef_class
class UmapProject:
    physical: bool = False

    def __call__(self, u_map: Umap, lors) -> Listmode:
        a = attenuation_factors(u_map, lors.data, self.physical)
        return Listmode(_out(a, "tf-eager") if not isinstance(lors.data, torch.Tensor) else a, lors)


__all__ = ["AttenuationCorrect", "UmapProject", "attenuation_factors"]
