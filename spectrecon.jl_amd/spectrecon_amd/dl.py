"""The projector inside torch autograd, and the regularised EM step of the
reference's deep-learning example (docs/lit/examples/6-dl.jl:150-198).

Tensors live on the GPU in float32 and are handed to the library by device
pointer (no copies).  Their torch shapes are the reverse of the reference's
index order, so that a C-contiguous tensor has the reference's column-major
memory: an image [nx, ny, nz] is a tensor (nz, ny, nx), views [nx, nz, nview]
are a tensor (nview, nz, nx).

As in the reference (docs/lit/examples/6-dl.jl:150-154), the projector's
Jacobians are the operators themselves: the backward of `project` is
`backproject` and vice versa; nothing is differentiated through the system
matrix.

Import torch (or this module) before the library is first used: torch ships
its own HIP runtime, which the library then shares.  The other way round, the
process holds two HIP runtimes and torch cannot open the device; the calls
here raise a RuntimeError that says so.
"""
from __future__ import annotations

import torch

from . import _lib as L
from ._lib import check, lib


def image_shape(plan):
    """Torch shape of an image of the plan: (nz, ny, nx)."""
    nx, ny, nz = plan.imgsize
    return (nz, ny, nx)


def views_shape(plan):
    """Torch shape of all of the plan's views: (nview, nz, nx)."""
    nx, _, nz = plan.imgsize
    return (plan.nview, nz, nx)


def _runtime_check():
    if not torch.cuda.is_available():
        raise RuntimeError("torch sees no GPU: if the library was loaded before torch, the process holds two HIP "
                           "runtimes; import torch (or spectrecon_amd.dl) before the library's first use")


def _dev(t, shape, what):
    _runtime_check()
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float32):
        raise TypeError(f"{what}: a float32 GPU tensor is needed")
    if tuple(t.shape) != tuple(shape):
        raise L.DimensionMismatch(f"{what}: shape {tuple(t.shape)}, expected {tuple(shape)}")
    return t.contiguous()


def _project(x, plan):
    x = _dev(x, image_shape(plan), "image")
    y = torch.empty(views_shape(plan), dtype=torch.float32, device=x.device)
    torch.cuda.synchronize(x.device)  # torch's queued work on x lands first
    check(lib().sr_project(plan.handle, y.data_ptr(), x.data_ptr(), None, plan.nview, L.SR_MEM_DEVICE))
    return y


def _backproject(y, plan):
    y = _dev(y, views_shape(plan), "views")
    x = torch.empty(image_shape(plan), dtype=torch.float32, device=y.device)
    torch.cuda.synchronize(y.device)
    check(lib().sr_backproject(plan.handle, x.data_ptr(), y.data_ptr(), None, plan.nview, L.SR_MEM_DEVICE))
    return x


class _Project(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, plan):
        ctx.plan = plan
        return _project(x, plan)

    @staticmethod
    def backward(ctx, gy):
        return _backproject(gy, ctx.plan), None


class _Backproject(torch.autograd.Function):
    @staticmethod
    def forward(ctx, y, plan):
        ctx.plan = plan
        return _backproject(y, plan)

    @staticmethod
    def backward(ctx, gx):
        return _project(gx, ctx.plan), None


def project(x, plan):
    """A x for an image tensor (nz, ny, nx): all views, (nview, nz, nx);
    differentiable (projectb, docs/lit/examples/6-dl.jl:150-151)."""
    return _Project.apply(x, plan)


def backproject(y, plan):
    """A' y for a views tensor (nview, nz, nx): an image (nz, ny, nx);
    differentiable (backprojectb, docs/lit/examples/6-dl.jl:153-154)."""
    return _Backproject.apply(y, plan)


def bregem(y, r, asum, x, cnn, beta, plan, niter=1):
    """Regularised EM with a learned prior (docs/lit/examples/6-dl.jl:179-198):
    u = cnn(x), then niter times
        x <- max(0, (-(asum - beta u) + sqrt((asum - beta u)^2 + 4 beta x A'(y / (A x + r)))) / (2 beta)),
    differentiable in x, the cnn's parameters and beta.  `cnn` maps an image
    tensor (nz, ny, nx) to one of the same shape (the caller adds the batch
    and channel dimensions its model needs, as the reference's unsqueeze45)."""
    u = cnn(x)
    asumu = asum - beta * u
    asumu2 = asumu * asumu
    for _ in range(niter):
        eterm = backproject(y / (project(x, plan) + r), plan)
        eterm_beta = 4 * beta * (x * eterm)
        x = torch.clamp_min((-asumu + torch.sqrt(asumu2 + eterm_beta)) / (2 * beta), 0.0)
    return x
