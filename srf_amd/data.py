"""Data containers with srfnef's names, fields and behaviour.

``nef_class`` mirrors srfnef/base/base.py:154-177: a frozen, slotted record
whose fields default to None, with ``update(**kw)`` returning a modified copy,
``asdict()`` / ``from_dict()`` and ``keys/values/items``.  ``data`` may be a
numpy array or a torch tensor (host or GPU); the operators in
``srf_amd.functions`` accept either.

Containers (reference file:line):
  Image / Emap / Umap   srfnef/data/image.py:20-37, emap.py:20-37, umap.py:13-14
  Lors                  srfnef/data/lors.py:6-26  (N,6|7) [x1,y1,z1,x2,y2,z2,(tof)] mm
  Listmode              srfnef/data/listmode.py:17-28  per-event values + Lors
  ImageConfig           srfnef/geometry/image_config.py:14-22
  TofConfig             srfnef/geometry/tof_config.py:14-17  (3.66 mm, 1000 mm^2)
"""
import dataclasses

import numpy as np

try:
    import torch
except ImportError:  # pragma: no cover
    torch = None


def nef_class(cls):
    """Frozen record with Optional fields (srfnef/base/base.py:154-177)."""
    ann = cls.__dict__.get("__annotations__", {})
    for name in ann:
        if not hasattr(cls, name):
            setattr(cls, name, None)
    dc = dataclasses.dataclass(frozen=True, eq=False, repr=False)(cls)

    def update(self, **kwargs):
        return dataclasses.replace(self, **kwargs)

    def asdict(self):
        return {f.name: getattr(self, f.name) for f in dataclasses.fields(self)}

    @classmethod
    def from_dict(klass, d):
        names = {f.name for f in dataclasses.fields(klass)}
        return klass(**{k: v for k, v in d.items() if k in names})

    def keys(self):
        return [f.name for f in dataclasses.fields(self)]

    def values(self):
        return [getattr(self, k) for k in self.keys()]

    def items(self):
        return list(zip(self.keys(), self.values()))

    def __repr__(self):
        parts = []
        for k, v in self.items():
            if hasattr(v, "shape"):
                parts.append(f"{k}=<{type(v).__name__} {tuple(v.shape)}>")
            else:
                parts.append(f"{k}={v!r}")
        return f"{type(self).__name__}({', '.join(parts)})"

    dc.update = update
    dc.asdict = asdict
    dc.from_dict = from_dict
    dc.keys = keys
    dc.values = values
    dc.items = items
    dc.__repr__ = __repr__
    return dc


class ArithmeticMixin:
    """+ - * / // % ** on .data (srfnef/ops/common/arithmetic_mixins.py:12-87)."""

    def _binop(self, other, op):
        o = other.data if hasattr(other, "data") else other
        return self.update(data=op(self.data, o))

    def __neg__(self):
        return self.update(data=-self.data)

    def abs(self):
        return self.update(data=abs(self.data))

    def __add__(self, other):
        return self._binop(other, lambda a, b: a + b)

    def __sub__(self, other):
        return self._binop(other, lambda a, b: a - b)

    def __mul__(self, other):
        return self._binop(other, lambda a, b: a * b)

    def __truediv__(self, other):
        return self._binop(other, lambda a, b: a / b)

    def __floordiv__(self, other):
        return self._binop(other, lambda a, b: a // b)

    def __mod__(self, other):
        return self._binop(other, lambda a, b: a % b)

    def __pow__(self, other):
        return self._binop(other, lambda a, b: a ** b)


class ShapeMixin:
    @property
    def shape(self):
        return tuple(self.data.shape)

    def __getitem__(self, item):
        return self.data[item]


class UnitSizeMixin:
    @property
    def unit_size(self):
        # srfnef/ops/common/property_mixins.py:64-67
        return [s1 / s2 for s1, s2 in zip(self.size, self.shape)]


@nef_class
class Image(ShapeMixin, UnitSizeMixin, ArithmeticMixin):
    data: object
    center: list
    size: list

    @property
    def central_slices(self):
        d = _np(self.data)
        return (d[d.shape[0] // 2, :, :], d[:, d.shape[1] // 2, :], d[:, :, d.shape[2] // 2])


@nef_class
class Emap(ShapeMixin, UnitSizeMixin, ArithmeticMixin):
    data: object
    center: list
    size: list


@nef_class
class Umap(Image):
    """Attenuation map (srfnef/data/umap.py:13-14: a subclass of Image)."""


@nef_class
class Lors(ShapeMixin):
    data: object

    @property
    def fst(self):
        return self.data[:, :3]

    @property
    def snd(self):
        return self.data[:, 3:6]

    @property
    def tof_values(self):
        return self.data[:, 6]

    def __len__(self):
        return int(self.data.shape[0])

    def append(self, lors):
        if torch is not None and isinstance(self.data, torch.Tensor):
            return self.update(data=torch.cat((self.data, lors.data), 0))
        return self.update(data=np.vstack((self.data, lors.data)))


@nef_class
class Listmode(ShapeMixin, ArithmeticMixin):
    data: object
    lors: Lors

    def __len__(self):
        return int(self.data.shape[0])

    def append(self, listmode):
        lors = self.lors.append(listmode.lors)
        if torch is not None and isinstance(self.data, torch.Tensor):
            return self.update(data=torch.cat((self.data, listmode.data)), lors=lors)
        return self.update(data=np.hstack((self.data, listmode.data)), lors=lors)


@nef_class
class ImageConfig(UnitSizeMixin):
    shape: list
    center: list
    size: list


@nef_class
class TofConfig:
    tof_bin: float = 3.66
    tof_sigma2: float = 1000.0


def _np(a):
    if torch is not None and isinstance(a, torch.Tensor):
        return a.detach().cpu().numpy()
    return np.asarray(a)


__all__ = ["nef_class", "Image", "Emap", "Umap", "Lors", "Listmode", "ImageConfig", "TofConfig"]
