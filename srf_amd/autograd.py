"""Differentiable projector.

The forward is the hand-written projector (srf::project / srf::plan_project);
the backward is the hand-written backprojector in adjoint mode
(SRF_BP_ADJOINT): the projector's own predicates, cuts and 1/L^2 weight,
scattered instead of gathered -- the exact transpose A^T.  The reference's
backprojector (srfnef/cc/siddon/siddon.cu.cc:342-656) is not A^T: its weight
lacks 1/L^2 and its cuts and dominance test differ (siddon.cu.cc:38,46,56-57
vs :372), so it equals A^T only after scaling the event values by 1/L^2 and
only where those predicates agree.

Gradients are registered on the custom ops themselves, so
``ops.project(...)`` and ``Plan.project(...)`` are differentiable with respect
to the image wherever they are used.
"""
import torch

from . import ops


def _project_setup(ctx, inputs, output):
    lors, image_k, shape, geom, use_tof, tof_bin, tof_sigma2, tof_values = inputs
    ctx.save_for_backward(lors, tof_values if tof_values is not None else lors.new_empty(0))
    ctx.meta = (list(shape), list(geom), bool(use_tof), float(tof_bin), float(tof_sigma2), tof_values is not None)


def _project_backward(ctx, grad):
    lors, tv = ctx.saved_tensors
    shape, geom, use_tof, tof_bin, tof_sigma2, has_tv = ctx.meta
    g = ops.backproject(lors, grad.contiguous(), shape, geom, use_tof, tof_bin, tof_sigma2, 1,
                        tv if has_tv else None)
    return None, g, None, None, None, None, None, None


torch.library.register_autograd("srf::project", _project_backward, setup_context=_project_setup)


def _plan_setup(ctx, inputs, output):
    handle, image_k, n, shape = inputs
    ctx.meta = (int(handle), int(n), list(shape))


def _plan_backward(ctx, grad):
    handle, n, shape = ctx.meta
    g = ops.plan_backproject(handle, grad.contiguous(), n, shape, 1)
    return None, g, None, None


torch.library.register_autograd("srf::plan_project", _plan_backward, setup_context=_plan_setup)


class Projector:
    """p = A x for a fixed LOR set, differentiable in x.

    lors: (N, 6|7) float32 GPU tensor; shape (nx, ny, nz); center, size in mm;
    tof: None or (tof_bin, tof_sigma2).  ``tiled`` bins the LOR set once into a
    plan (LDS-tiled kernels); otherwise the one-event-per-lane kernels run.

    __call__(x): x of shape (nx, ny, nz) or (B, nx, ny, nz) in the Python
    layout -> (N,) or (B, N)."""

    def __init__(self, lors, shape, center, size, tof=None, tiled=True):
        self.lors = lors.contiguous()
        self.shape = [int(v) for v in shape]
        self.geom = [float(v) for v in center] + [float(v) for v in size]
        self.use_tof = tof is not None
        self.tof_bin, self.tof_sigma2 = (float(tof[0]), float(tof[1])) if tof is not None else (0.0, 0.0)
        self.plan = ops.Plan(self.lors, self.shape, self.geom, self.use_tof, self.tof_bin,
                             self.tof_sigma2) if tiled else None

    def _one(self, x):
        xk = x.permute(2, 1, 0).contiguous()           # (nz, ny, nx): the kernel layout
        if self.plan is not None:
            return self.plan.project(xk)
        return ops.project(self.lors, xk, self.shape, self.geom, self.use_tof, self.tof_bin, self.tof_sigma2)

    def __call__(self, x):
        if x.dim() == 3:
            return self._one(x)
        return torch.stack([self._one(x[b]) for b in range(x.shape[0])], 0)

    def transpose(self, y):
        """A^T y in the Python layout (nx, ny, nz)."""
        y = y.contiguous()
        if self.plan is not None:
            g = self.plan.backproject(y, 1)
        else:
            g = ops.backproject(self.lors, y, self.shape, self.geom, self.use_tof, self.tof_bin, self.tof_sigma2, 1)
        return g.permute(2, 1, 0)


__all__ = ["Projector"]
