"""PyTorch-ROCm custom ops over libsrf_hip.so.

Each op is a ``torch.library.custom_op`` in the ``srf`` namespace whose body is
a single C-ABI call on torch's current stream.  Tensors must already live on
the GPU (float32, contiguous); nothing here copies host data or falls back to
the CPU.

Images are in the kernel layout: shape (nz, ny, nx), x fastest -- the
``tf.transpose(image.data)`` the reference hands to its ops
(srfnef/functions/project.py:27).  ``to_kernel_layout`` / ``from_kernel_layout``
convert from / to the Python (nx, ny, nz) layout on the GPU.
"""
import ctypes
from typing import List, Optional, Tuple

import torch

from . import _lib
from ._lib import SrfGrid, SrfPlanOpts, SrfTof, check, lib, stream_handle

Tensor = torch.Tensor


def _grid(shape, geom):
    nx, ny, nz = (int(v) for v in shape)
    cx, cy, cz, sx, sy, sz = (float(v) for v in geom)
    return SrfGrid(nx, ny, nz, cx, cy, cz, sx, sy, sz)


def _tof(use_tof, tof_bin, tof_sigma2):
    if not use_tof:
        return None
    return ctypes.pointer(SrfTof(float(tof_bin), float(tof_sigma2)))


def _req(t: Tensor, name: str, dtype=torch.float32):
    if not t.is_cuda:
        raise ValueError(f"{name} must be a GPU tensor (got {t.device})")
    if t.dtype != dtype:
        raise ValueError(f"{name} must be {dtype} (got {t.dtype})")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")


def _lor_args(lors: Tensor, tof_values: Optional[Tensor], use_tof: bool):
    _req(lors, "lors")
    if lors.dim() != 2 or lors.shape[1] < 6:
        raise ValueError(f"lors must be (N, >=6), got {tuple(lors.shape)}")
    n, stride = lors.shape
    tp, ts = None, 0
    if use_tof:
        if tof_values is not None:
            _req(tof_values, "tof_values")
            if tof_values.numel() != n:
                raise ValueError("tof_values must have one entry per LOR")
            tp, ts = tof_values.data_ptr(), 1
        else:
            if stride < 7:
                raise ValueError("TOF needs a 7th LOR column or explicit tof_values")
            tp, ts = lors.data_ptr() + 6 * 4, stride
    return n, stride, tp, ts


def _check_image(img: Tensor, shape, name="image"):
    _req(img, name)
    nx, ny, nz = (int(v) for v in shape)
    if img.numel() != nx * ny * nz:
        raise ValueError(f"{name} has {img.numel()} voxels, grid says {nx}x{ny}x{nz}")


# ------------------------------------------------------------------ project --

@torch.library.custom_op("srf::project", mutates_args=())
def project(lors: Tensor, image_k: Tensor, shape: List[int], geom: List[float], use_tof: bool,
            tof_bin: float, tof_sigma2: float, tof_values: Optional[Tensor] = None) -> Tensor:
    """Forward projection p[e] (Projection / ProjectionTof)."""
    n, stride, tp, ts = _lor_args(lors, tof_values, use_tof)
    _check_image(image_k, shape)
    out = torch.empty(n, dtype=torch.float32, device=lors.device)
    with torch.cuda.device(lors.device):
        rc = lib().srf_project(lors.data_ptr(), n, stride, tp, ts, image_k.data_ptr(), _grid(shape, geom),
                               _tof(use_tof, tof_bin, tof_sigma2), out.data_ptr(), stream_handle(lors.device))
    check(rc, "srf_project")
    return out


@project.register_fake
def _(lors, image_k, shape, geom, use_tof, tof_bin, tof_sigma2, tof_values=None):
    return lors.new_empty(lors.shape[0])


# -------------------------------------------------------------- backproject --

@torch.library.custom_op("srf::backproject", mutates_args=())
def backproject(lors: Tensor, values: Optional[Tensor], shape: List[int], geom: List[float], use_tof: bool,
                tof_bin: float, tof_sigma2: float, mode: int, tof_values: Optional[Tensor] = None) -> Tensor:
    """Backprojection into a fresh kernel-layout image (Backprojection / BackprojectionTof).
    values None = all ones.  mode: 0 reference weights, 1 exact adjoint of project."""
    n, stride, tp, ts = _lor_args(lors, tof_values, use_tof)
    vp = None
    if values is not None:
        _req(values, "values")
        if values.numel() != n:
            raise ValueError("values must have one entry per LOR")
        vp = values.data_ptr()
    nx, ny, nz = (int(v) for v in shape)
    img = torch.empty((nz, ny, nx), dtype=torch.float32, device=lors.device)
    with torch.cuda.device(lors.device):
        rc = lib().srf_backproject(lors.data_ptr(), n, stride, tp, ts, vp, _grid(shape, geom),
                                   _tof(use_tof, tof_bin, tof_sigma2), int(mode), 0, img.data_ptr(),
                                   stream_handle(lors.device))
    check(rc, "srf_backproject")
    return img


@backproject.register_fake
def _(lors, values, shape, geom, use_tof, tof_bin, tof_sigma2, mode, tof_values=None):
    nx, ny, nz = shape
    return lors.new_empty((nz, ny, nx))


@torch.library.custom_op("srf::backproject_ratio", mutates_args=())
def backproject_ratio(lors: Tensor, y: Tensor, p: Tensor, kind: int, eps: float, shape: List[int],
                      geom: List[float], use_tof: bool, tof_bin: float, tof_sigma2: float, mode: int,
                      tof_values: Optional[Tensor] = None) -> Tensor:
    """Backprojection of ratio(y, p) with the ratio fused into the event load."""
    n, stride, tp, ts = _lor_args(lors, tof_values, use_tof)
    _req(y, "y")
    _req(p, "p")
    if y.numel() != n or p.numel() != n:
        raise ValueError("y and p must have one entry per LOR")
    nx, ny, nz = (int(v) for v in shape)
    img = torch.empty((nz, ny, nx), dtype=torch.float32, device=lors.device)
    with torch.cuda.device(lors.device):
        rc = lib().srf_backproject_ratio(lors.data_ptr(), n, stride, tp, ts, y.data_ptr(), p.data_ptr(), int(kind),
                                         float(eps), _grid(shape, geom), _tof(use_tof, tof_bin, tof_sigma2),
                                         int(mode), 0, img.data_ptr(), stream_handle(lors.device))
    check(rc, "srf_backproject_ratio")
    return img


@backproject_ratio.register_fake
def _(lors, y, p, kind, eps, shape, geom, use_tof, tof_bin, tof_sigma2, mode, tof_values=None):
    nx, ny, nz = shape
    return lors.new_empty((nz, ny, nx))


# --------------------------------------------------------------- elementwise --

@torch.library.custom_op("srf::ratio", mutates_args=())
def ratio(y: Tensor, p: Tensor, kind: int, eps: float) -> Tensor:
    _req(y, "y")
    _req(p, "p")
    if y.numel() != p.numel():
        raise ValueError("y and p differ in size")
    q = torch.empty_like(y)
    with torch.cuda.device(y.device):
        rc = lib().srf_ratio(y.data_ptr(), p.data_ptr(), q.data_ptr(), y.numel(), int(kind), float(eps),
                             stream_handle(y.device))
    check(rc, "srf_ratio")
    return q


@ratio.register_fake
def _(y, p, kind, eps):
    return torch.empty_like(y)


@torch.library.custom_op("srf::em_update", mutates_args=("x",))
def em_update(x: Tensor, bp: Tensor, emap: Tensor, zero_guard: float) -> None:
    """x <- x * bp / guard(emap), in place."""
    for t, nm in ((x, "x"), (bp, "bp"), (emap, "emap")):
        _req(t, nm)
    if not (x.numel() == bp.numel() == emap.numel()):
        raise ValueError("x, bp, emap differ in size")
    with torch.cuda.device(x.device):
        rc = lib().srf_em_update(x.data_ptr(), bp.data_ptr(), emap.data_ptr(), x.numel(), float(zero_guard),
                                 stream_handle(x.device))
    check(rc, "srf_em_update")


@em_update.register_fake
def _(x, bp, emap, zero_guard):
    return None


@torch.library.custom_op("srf::lor_digest", mutates_args=())
def lor_digest(lors: Tensor, shape: List[int], geom: List[float], use_tof: bool, tof_bin: float,
               tof_sigma2: float, mode: int, tof_values: Optional[Tensor] = None) -> Tuple[Tensor, Tensor]:
    """Per-event (hash, count) of the visited (voxel, branch) terms.
    mode 1: projector predicates; mode 0: backprojector predicates."""
    n, stride, tp, ts = _lor_args(lors, tof_values, use_tof)
    h = torch.empty(n, dtype=torch.int64, device=lors.device)
    c = torch.empty(n, dtype=torch.int32, device=lors.device)
    with torch.cuda.device(lors.device):
        rc = lib().srf_lor_digest(lors.data_ptr(), n, stride, tp, ts, _grid(shape, geom),
                                  _tof(use_tof, tof_bin, tof_sigma2), int(mode), h.data_ptr(), c.data_ptr(),
                                  stream_handle(lors.device))
    check(rc, "srf_lor_digest")
    return h, c


@lor_digest.register_fake
def _(lors, shape, geom, use_tof, tof_bin, tof_sigma2, mode, tof_values=None):
    n = lors.shape[0]
    return lors.new_empty(n, dtype=torch.int64), lors.new_empty(n, dtype=torch.int32)


@torch.library.custom_op("srf::to_kernel_layout", mutates_args=())
def to_kernel_layout(img_xyz: Tensor) -> Tensor:
    """(nx, ny, nz) -> (nz, ny, nx) contiguous, on the GPU."""
    _req(img_xyz, "image")
    if img_xyz.dim() != 3:
        raise ValueError("image must be 3-D (nx, ny, nz)")
    nx, ny, nz = img_xyz.shape
    out = torch.empty((nz, ny, nx), dtype=torch.float32, device=img_xyz.device)
    with torch.cuda.device(img_xyz.device):
        rc = lib().srf_image_to_kernel_layout(img_xyz.data_ptr(), out.data_ptr(), nx, ny, nz,
                                              stream_handle(img_xyz.device))
    check(rc, "srf_image_to_kernel_layout")
    return out


@to_kernel_layout.register_fake
def _(img_xyz):
    nx, ny, nz = img_xyz.shape
    return img_xyz.new_empty((nz, ny, nx))


@torch.library.custom_op("srf::from_kernel_layout", mutates_args=())
def from_kernel_layout(img_k: Tensor) -> Tensor:
    """(nz, ny, nx) -> (nx, ny, nz) contiguous, on the GPU."""
    _req(img_k, "image")
    if img_k.dim() != 3:
        raise ValueError("image must be 3-D (nz, ny, nx)")
    nz, ny, nx = img_k.shape
    out = torch.empty((nx, ny, nz), dtype=torch.float32, device=img_k.device)
    with torch.cuda.device(img_k.device):
        rc = lib().srf_image_from_kernel_layout(img_k.data_ptr(), out.data_ptr(), nx, ny, nz,
                                                stream_handle(img_k.device))
    check(rc, "srf_image_from_kernel_layout")
    return out


@from_kernel_layout.register_fake
def _(img_k):
    nz, ny, nx = img_k.shape
    return img_k.new_empty((nx, ny, nz))


# -------------------------------------------------------------------- plans --

def trim_plan_pool(keep_bytes=0, device_=None):
    """Return the plan pool's freed memory above keep_bytes on a device to the
    system (srf_plan_pool_trim)."""
    dev = device_ if device_ is not None else torch.device("cuda", torch.cuda.current_device())
    with torch.cuda.device(dev):
        check(lib().srf_plan_pool_trim(int(keep_bytes)), "srf_plan_pool_trim")


class Plan:
    """A binned event list (srf_plan_*): the LDS-tiled projector/backprojector.

    Built once per fixed event list (an OSEM subset, a sensitivity LOR set) and
    reused by every iteration; results equal the one-shot ops above to float
    round-off.  Device memory is owned by the plan and freed by close()/GC."""

    def __init__(self, lors: Tensor, shape, geom, use_tof: bool, tof_bin: float = 0.0, tof_sigma2: float = 0.0,
                 tof_values: Optional[Tensor] = None, item_records: int = 8192, tof_window: float = 0.0):
        n, stride, tp, ts = _lor_args(lors, tof_values, use_tof)
        self.item_records = int(item_records)
        self.tof_window = float(tof_window)
        self.shape = [int(v) for v in shape]
        self.geom = [float(v) for v in geom]
        self.use_tof, self.tof_bin, self.tof_sigma2 = bool(use_tof), float(tof_bin), float(tof_sigma2)
        self.n = n
        self.device = lors.device
        self._h = ctypes.c_void_p()
        with torch.cuda.device(lors.device):
            opts = SrfPlanOpts(self.item_records, self.tof_window)
            rc = lib().srf_plan_create_ex(lors.data_ptr(), n, stride, tp, ts, _grid(self.shape, self.geom),
                                          _tof(use_tof, tof_bin, tof_sigma2), ctypes.byref(opts),
                                          ctypes.byref(self._h), stream_handle(lors.device))
        check(rc, "srf_plan_create_ex")
        self._lib = lib()

    @property
    def handle(self) -> int:
        if not self._h:
            raise ValueError("plan is closed")
        return int(self._h.value)

    def rebuild(self, lors: Tensor, tof_values: Optional[Tensor] = None):
        """Bin a new event list (same grid, TOF and options) into this plan,
        reusing its device memory (srf_plan_rebuild)."""
        n, stride, tp, ts = _lor_args(lors, tof_values, self.use_tof)
        with torch.cuda.device(lors.device):
            opts = SrfPlanOpts(self.item_records, self.tof_window)
            rc = lib().srf_plan_rebuild(self.handle, lors.data_ptr(), n, stride, tp, ts, _grid(self.shape, self.geom),
                                        _tof(self.use_tof, self.tof_bin, self.tof_sigma2), ctypes.byref(opts),
                                        stream_handle(lors.device))
        self.n = 0
        check(rc, "srf_plan_rebuild")
        self.n = n
        return self

    def info(self):
        ne, ns, ni = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        nb = ctypes.c_size_t()
        check(lib().srf_plan_info(self.handle, ctypes.byref(ne), ctypes.byref(ns), ctypes.byref(ni),
                                  ctypes.byref(nb)), "srf_plan_info")
        return {"events": ne.value, "segments": ns.value, "items": ni.value, "device_bytes": nb.value}

    def project(self, image_k: Tensor) -> Tensor:
        return plan_project(self.handle, image_k, self.n, self.shape)

    def project_ratio(self, image_k: Tensor, y: Tensor, kind: int, eps: float) -> Tensor:
        """q = ratio(y, project(image)) in one pass (the EM forward half)."""
        return plan_project_ratio(self.handle, image_k, y, int(kind), float(eps), self.n, self.shape)

    def backproject(self, values: Optional[Tensor] = None, mode: int = 0) -> Tensor:
        return plan_backproject(self.handle, values, self.n, self.shape, int(mode))

    def backproject_update(self, values: Tensor, x_k: Tensor, emap_k: Tensor, zero_guard: float = 0.0):
        """x_k <- x_k * backproject(values) / emap_k in place (srf_plan_backproject_update)."""
        _req(values, "values")
        if values.numel() != self.n:
            raise ValueError("values must have one entry per event")
        _check_image(x_k, self.shape)
        _check_image(emap_k, self.shape)
        with torch.cuda.device(x_k.device):
            rc = lib().srf_plan_backproject_update(self.handle, values.data_ptr(), x_k.data_ptr(), emap_k.data_ptr(),
                                                   float(zero_guard), stream_handle(x_k.device))
        check(rc, "srf_plan_backproject_update")
        return x_k

    def backproject_ratio(self, y: Tensor, p: Tensor, kind: int, eps: float, mode: int = 0) -> Tensor:
        return plan_backproject_ratio(self.handle, y, p, int(kind), float(eps), self.n, self.shape, int(mode))

    def time_next(self, before, after):
        """Record torch.cuda.Event before / after the tiled kernel of the next
        projection or backprojection of this plan (srf_plan_time_next).  Both
        events are recorded once here first, so that torch has created them."""
        stream = torch.cuda.current_stream(self.device)
        before.record(stream)
        after.record(stream)
        check(lib().srf_plan_time_next(self.handle, before.cuda_event, after.cuda_event), "srf_plan_time_next")

    def digest(self, mode: int):
        h = torch.empty(self.n, dtype=torch.int64, device=self.device)
        c = torch.empty(self.n, dtype=torch.int32, device=self.device)
        with torch.cuda.device(self.device):
            rc = lib().srf_plan_digest(self.handle, int(mode), h.data_ptr(), c.data_ptr(),
                                       stream_handle(self.device))
        check(rc, "srf_plan_digest")
        return h, c

    def close(self):
        if self._h:
            torch.cuda.synchronize(self.device)
            self._lib.srf_plan_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@torch.library.custom_op("srf::plan_project", mutates_args=())
def plan_project(handle: int, image_k: Tensor, n: int, shape: List[int]) -> Tensor:
    """Projection of every event of a plan (srf_plan_project)."""
    _check_image(image_k, shape)
    out = torch.empty(n, dtype=torch.float32, device=image_k.device)
    with torch.cuda.device(image_k.device):
        rc = lib().srf_plan_project(handle, image_k.data_ptr(), out.data_ptr(), stream_handle(image_k.device))
    check(rc, "srf_plan_project")
    return out


@plan_project.register_fake
def _(handle, image_k, n, shape):
    return image_k.new_empty(n)


@torch.library.custom_op("srf::plan_project_ratio", mutates_args=())
def plan_project_ratio(handle: int, image_k: Tensor, y: Tensor, kind: int, eps: float, n: int,
                       shape: List[int]) -> Tensor:
    """q[e] = ratio(y[e], p[e]), p = projection of every event of a plan."""
    _check_image(image_k, shape)
    _req(y, "y")
    if y.numel() != n:
        raise ValueError("y must have one entry per event")
    out = torch.empty(n, dtype=torch.float32, device=image_k.device)
    with torch.cuda.device(image_k.device):
        rc = lib().srf_plan_project_ratio(handle, image_k.data_ptr(), y.data_ptr(), int(kind), float(eps),
                                          out.data_ptr(), stream_handle(image_k.device))
    check(rc, "srf_plan_project_ratio")
    return out


@plan_project_ratio.register_fake
def _(handle, image_k, y, kind, eps, n, shape):
    return image_k.new_empty(n)


@torch.library.custom_op("srf::plan_backproject", mutates_args=())
def plan_backproject(handle: int, values: Optional[Tensor], n: int, shape: List[int], mode: int) -> Tensor:
    nx, ny, nz = (int(v) for v in shape)
    dev = values.device if values is not None else torch.device("cuda", torch.cuda.current_device())
    vp = None
    if values is not None:
        _req(values, "values")
        if values.numel() != n:
            raise ValueError("values must have one entry per event")
        vp = values.data_ptr()
    img = torch.empty((nz, ny, nx), dtype=torch.float32, device=dev)
    with torch.cuda.device(dev):
        rc = lib().srf_plan_backproject(handle, vp, int(mode), 0, img.data_ptr(), stream_handle(dev))
    check(rc, "srf_plan_backproject")
    return img


@plan_backproject.register_fake
def _(handle, values, n, shape, mode):
    nx, ny, nz = shape
    return torch.empty((nz, ny, nx), device="meta")


@torch.library.custom_op("srf::plan_backproject_ratio", mutates_args=())
def plan_backproject_ratio(handle: int, y: Tensor, p: Tensor, kind: int, eps: float, n: int, shape: List[int],
                           mode: int) -> Tensor:
    _req(y, "y")
    _req(p, "p")
    if y.numel() != n or p.numel() != n:
        raise ValueError("y and p must have one entry per event")
    nx, ny, nz = (int(v) for v in shape)
    img = torch.empty((nz, ny, nx), dtype=torch.float32, device=y.device)
    with torch.cuda.device(y.device):
        rc = lib().srf_plan_backproject_ratio(handle, y.data_ptr(), p.data_ptr(), int(kind), float(eps), int(mode), 0,
                                              img.data_ptr(), stream_handle(y.device))
    check(rc, "srf_plan_backproject_ratio")
    return img


@plan_backproject_ratio.register_fake
def _(handle, y, p, kind, eps, n, shape, mode):
    nx, ny, nz = shape
    return y.new_empty((nz, ny, nx))


# gradients of srf::project / srf::plan_project (adjoint-mode backprojector)
from . import autograd as _autograd  # noqa: E402,F401

__all__ = ["Plan", "plan_project", "plan_backproject", "plan_backproject_ratio", "project", "backproject", "backproject_ratio", "ratio", "em_update", "lor_digest",
           "to_kernel_layout", "from_kernel_layout", "_lib"]
