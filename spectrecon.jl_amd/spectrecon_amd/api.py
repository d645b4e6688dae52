"""Host-side mirror of SPECTrecon.jl's projector API over the HIP C-ABI.

Names follow the reference (a trailing ``_`` marks the mutating ``!`` forms);
arrays use the reference's index order and are handled as Fortran-ordered
numpy arrays (or DeviceArray for device-resident data):
    image, mumap [nx, ny, nz];  views [nx, nz, nview];  psfs [px, pz, ny, nview].
View indices are 0-based.  Every compute call goes to the HIP library; there
is no CPU path here.
"""
from __future__ import annotations

import ctypes as C
import math
import time
import warnings

import numpy as np

from . import _lib as L
from ._lib import DimensionMismatch, SPECTError, check, lib, ptr
from .device import DeviceArray

__all__ = [
    "SPECTplan", "project", "project_", "backproject", "backproject_",
    "mlem", "mlem_", "osem", "osem_", "Ablock", "LinearMapAA", "spect_operator",
    "PlanRotate", "plan_rotate", "imrotate", "imrotate_", "imrotate_adj", "imrotate_adj_",
    "PlanPSF", "plan_psf", "fft_conv", "fft_conv_", "fft_conv_adj", "fft_conv_adj_", "fft_conv_adj2_",
    "psf_gauss", "phantom", "mumap_phantom", "poisson", "gradient", "project_vjp_mu", "osem_strided",
]

_METHODS = {"one": L.SR_INTERP_ONE, "two": L.SR_INTERP_TWO}
_MODES = {"fast": L.SR_MODE_FAST, "mem": L.SR_MODE_MEM}


def _sym(s):
    return s[1:] if isinstance(s, str) and s.startswith(":") else s


def _mem(*arrays):
    dev = [isinstance(a, DeviceArray) for a in arrays]
    if all(dev):
        return L.SR_MEM_DEVICE
    if any(dev):
        raise TypeError("mix of host and device arrays")
    return L.SR_MEM_HOST


def _f32(a):
    return np.asfortranarray(a, dtype=np.float32)


def _writeback(dst, src):
    """Copy a Fortran float32 result into the caller's output array."""
    if dst is not src:
        dst[...] = src


# ---------------------------------------------------------------------------
# psf_gauss (src/psf-gauss.jl:24-55)
# ---------------------------------------------------------------------------
def psf_gauss(*, ny=128, px=11, pz=None, fwhm_start=1.0, fwhm_end=4.0, fwhm=None, fwhm_x=None, fwhm_z=None,
              T=np.float32):
    """Depth-dependent Gaussian PSFs [px, pz, ny], each summing to 1."""
    pz = px if pz is None else pz
    if px % 2 == 0:
        warnings.warn(f"even px = {px} ?")
    if pz % 2 == 0:
        warnings.warn(f"even pz = {pz} ?")
    fwhm = np.linspace(fwhm_start, fwhm_end, ny) if fwhm is None else np.asarray(fwhm, np.float64)
    fx = np.ascontiguousarray(fwhm if fwhm_x is None else fwhm_x, np.float64)
    fz = np.ascontiguousarray(fx if fwhm_z is None else fwhm_z, np.float64)
    out = np.zeros((px, pz, ny), np.float32, order="F")
    check(lib().sr_psf_gauss(out.ctypes.data, ny, px, pz, fx.ctypes.data, fz.ctypes.data))
    return out.astype(T, order="F") if np.dtype(T) != np.float32 else out


def phantom(nx, ny, nz, seed=0):
    """Seeded sum of 6 random ellipsoids (intensities U(0.5,4)), [nx, ny, nz]."""
    out = np.zeros((nx, ny, nz), np.float32, order="F")
    check(lib().sr_phantom(out.ctypes.data, nx, ny, nz, seed))
    return out


def mumap_phantom(nx, ny, nz):
    """0.15 outer ellipsoid with two 0.05 inner ellipsoids, [nx, ny, nz]."""
    out = np.zeros((nx, ny, nz), np.float32, order="F")
    check(lib().sr_mumap(out.ctypes.data, nx, ny, nz))
    return out


def poisson(mean, seed=0):
    mean = _f32(mean)
    out = np.zeros_like(mean)
    check(lib().sr_poisson(out.ctypes.data, mean.ctypes.data, mean.size, seed))
    return out


# ---------------------------------------------------------------------------
# SPECTplan (src/spectplan.jl:35-157)
# ---------------------------------------------------------------------------
def default_viewangle(nview, T=np.float32):
    """(0:nview-1)/nview * T(2π) as a Float32 StepRangeLen evaluates it."""
    twopi = float(np.float32(2 * math.pi))
    return np.array([np.float32(k * twopi / nview) for k in range(nview)], dtype=np.float32)


class SPECTplan:
    """Device-resident SPECT system model: SPECTplan(mumap, psfs, dy; ...)."""

    def __init__(self, mumap, psfs, dy, *, T=np.float32, viewangle=None, interpmeth="two", nthread=None,
                 mode="fast", conv="auto"):
        if np.dtype(T) != np.float32:
            raise SPECTError(f"only T=Float32 is built on the device (got {np.dtype(T)})")
        # mumap .= T.(mumap); psfs .= T.(psfs)  (src/spectplan.jl:71-72)
        for a in (mumap, psfs):
            if isinstance(a, np.ndarray) and a.dtype != np.float32 and np.issubdtype(a.dtype, np.floating):
                a[...] = a.astype(np.float32)
        self.T = np.float32
        self.dy = float(np.float32(dy))
        mumap = _f32(mumap)
        psfs = _f32(psfs)
        if mumap.ndim != 3:
            raise DimensionMismatch("mumap must be 3D")
        if psfs.ndim != 4:
            raise DimensionMismatch("psfs must be 4D")
        nx, ny, nz = mumap.shape
        px, pz, _, nview = psfs.shape
        self.imgsize = (nx, ny, nz)
        self.px, self.pz, self.nview = px, pz, nview
        self.interpmeth = _sym(interpmeth)
        self.mode = _sym(mode)
        self.nthread = 1 if nthread is None else nthread
        va = default_viewangle(nview) if viewangle is None else np.asarray(viewangle, np.float32)
        if va.shape != (nview,):
            raise DimensionMismatch("viewangle length must equal size(psfs, 4)")
        self.viewangle = np.ascontiguousarray(va)
        self.mumap = mumap
        self.psfs = psfs
        # checks whose messages live on the Python side (the rest are in the C-ABI)
        if nx == ny and nx % 2 == 0 and px % 2 == 1 and pz % 2 == 1:
            if self.interpmeth not in _METHODS:
                raise SPECTError("bad interpmeth")
            if self.mode not in _MODES:
                raise SPECTError("bad mode")
        desc = L.PlanDesc(nx, ny, nz, px, pz, nview, self.dy, _METHODS.get(self.interpmeth, 0),
                          _MODES.get(self.mode, -1),
                          {"auto": L.SR_CONV_AUTO, "direct": L.SR_CONV_DIRECT, "separable": L.SR_CONV_SEPARABLE,
                           "fft": L.SR_CONV_FFT}[conv])
        h = C.c_void_p()
        check(lib().sr_plan_create(C.byref(desc), mumap.ctypes.data, psfs.ctypes.data, self.viewangle.ctypes.data,
                                   C.byref(h)))
        self._h = h.value

    @property
    def handle(self):
        if not self._h:
            raise SPECTError("plan destroyed")
        return self._h

    def info(self):
        nbytes, conv, batch = C.c_size_t(), C.c_int32(), C.c_int32()
        check(lib().sr_plan_info(self.handle, C.byref(nbytes), C.byref(conv), C.byref(batch)))
        layout = C.c_int32()
        check(lib().sr_plan_layout(self.handle, C.byref(layout)))
        return {"device_bytes": nbytes.value, "conv": {1: "direct", 2: "separable", 3: "fft"}[conv.value],
                "max_batch": batch.value, "layout": {0: "x-fastest", 1: "z-fastest"}[layout.value]}

    def set_mumap(self, mumap):
        """A new attenuation map for this plan: every view's attenuation
        factors are rebuilt on the device (geometry, PSFs and angles stay)."""
        if isinstance(mumap, DeviceArray):
            if tuple(mumap.shape) != self.imgsize:
                raise DimensionMismatch("mumap size")
            check(lib().sr_plan_set_mumap(self.handle, mumap.ptr, L.SR_MEM_DEVICE))
            return self
        mu = _f32(mumap)
        if mu.shape != self.imgsize:
            raise DimensionMismatch("mumap size")
        check(lib().sr_plan_set_mumap(self.handle, mu.ctypes.data, L.SR_MEM_HOST))
        self.mumap = mu
        return self

    def set_batch(self, n):
        check(lib().sr_plan_set_batch(self.handle, n))

    def synchronize(self):
        check(lib().sr_plan_synchronize(self.handle))

    def set_timing(self, enable=True, classes=None):
        """Bracket kernel launches of this plan with HIP events: all of them,
        or only those of the named classes (names of KERNEL_CLASSES)."""
        mask = 0
        if enable:
            mask = 1 if classes is None else sum(2 << L.KERNEL_CLASSES.index(c) for c in classes)
        check(lib().sr_plan_set_timing(self.handle, mask))

    def set_graph(self, enable=True):
        """Replay the EM kernel sequence of mlem/osem calls from a HIP graph
        (captured on the second call with the same buffers); results equal
        the eager ones bit for bit."""
        check(lib().sr_plan_set_graph(self.handle, 1 if enable else 0))

    def graph_stats(self):
        """(graphs captured, graph replays) since the plan was made."""
        cap, rep = C.c_int64(), C.c_int64()
        check(lib().sr_plan_graph_stats(self.handle, C.byref(cap), C.byref(rep)))
        return cap.value, rep.value

    CHOICES = ("psf_fwd_family", "psf_fwd_chunk_percent", "psf_adj_family", "psf_adj_chunk_percent",
               "gather_slices", "psf_method", "em_norm")

    def choice(self):
        """The plan's kernel choices: PSF families (forward: 0 built-in
        producer/consumer waves, 1 tile, 2 column-owned, 3 / 4 producer/consumer
        with 8 / 4 producers, 5 producer/consumer with column-pair consumers;
        adjoint: 0 column-owned, 1 tile), chunk counts in
        percent of the built-in ones, gather view slices (0 built-in), PSF
        method (0 built-in, 1 direct taps, 2 plane FFTs), where OS-EM sums its
        subset normalisers (0 built-in, 1 in each subset's first gather, 2 all
        subsets in one launch before the iterations)."""
        c = (C.c_int32 * len(self.CHOICES))()
        check(lib().sr_plan_get_choice(self.handle, c))
        return dict(zip(self.CHOICES, list(c)))

    def set_choice(self, **kw):
        """Set some kernel choices by name (the rest keep their values)."""
        cur = self.choice()
        for k, v in kw.items():
            if k not in cur:
                raise KeyError(k)
            cur[k] = int(v)
        c = (C.c_int32 * len(self.CHOICES))(*[cur[k] for k in self.CHOICES])
        check(lib().sr_plan_set_choice(self.handle, c))

    def autotune(self, nsub):
        """Time the candidates of each kernel choice on batches of nsub views
        and keep the fastest; returns the record of every candidate's time
        (nsub <= 0 restores the built-in choices)."""
        import json
        buf = C.create_string_buffer(8192)
        check(lib().sr_plan_autotune(self.handle, int(nsub), buf, len(buf)))
        return json.loads(buf.value.decode())

    def timing(self):
        """{kernel class: (total device ms, launches)} since set_timing(True)."""
        out = {}
        for k, name in enumerate(L.KERNEL_CLASSES):
            ms, n = C.c_double(), C.c_int64()
            check(lib().sr_plan_get_timing(self.handle, k, C.byref(ms), C.byref(n)))
            out[name] = (ms.value, n.value)
        return out

    def close(self):
        if getattr(self, "_h", None):
            lib().sr_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __repr__(self):
        return (f"SPECTplan{{Float32}} imgsize={self.imgsize} px={self.px} pz={self.pz} nview={self.nview} "
                f"interpmeth=:{self.interpmeth} mode=:{self.mode} dy={self.dy}")


def _index(plan, index, n_out=None):
    if index is None:
        return None, plan.nview
    idx = np.ascontiguousarray(np.asarray(list(index) if not isinstance(index, np.ndarray) else index, np.int32))
    return idx, len(idx)


# ---------------------------------------------------------------------------
# project! / project (src/project.jl)
# ---------------------------------------------------------------------------
def _view_args(args, name):
    """The trailing (viewidx) or (buffer_id, viewidx) of the single-view forms
    (src/project.jl:11-69, :77-135; src/backproject.jl:9-62, :69-124).  The
    reference's buffer_id picks a thread's work buffers; the device plan owns
    its buffers, so it is accepted and not used."""
    if len(args) == 0:
        return None
    if len(args) == 1:
        return int(args[0])
    if len(args) == 2:
        return int(args[1])
    raise TypeError(f"{name}: expected (viewidx) or (buffer_id, viewidx), got {len(args)} arguments")


def project_(views, image, plan, *args, index=None, asynchronous=False):
    """project!(views, image, plan; index) -- views[:,:,ii] = A_{index[ii]} image.
    With a 2-D view: project!(view, image, plan, viewidx) and
    project!(view, image, plan, buffer_id, viewidx) (0-based viewidx)."""
    viewidx = _view_args(args, "project!")
    if viewidx is not None:
        if np.ndim(views) != 2:
            raise DimensionMismatch("single-view form needs a 2D view")
        tmp = np.zeros((views.shape[0], views.shape[1], 1), np.float32, order="F")
        project_(tmp, image, plan, index=[viewidx])
        views[...] = tmp[:, :, 0]
        return views
    idx, n = _index(plan, index)
    nx, ny, nz = plan.imgsize
    if tuple(image.shape) != plan.imgsize:
        raise DimensionMismatch("image size")
    if tuple(views.shape) != (nx, nz, n):
        raise DimensionMismatch("views size")
    mem = _mem(views, image)
    if mem == L.SR_MEM_DEVICE:
        check(lib().sr_project(plan.handle, views.ptr, image.ptr, idx.ctypes.data if idx is not None else None, n,
                               mem | (L.SR_ASYNC if asynchronous else 0)))
        return views
    out = views if (views.dtype == np.float32 and views.flags.f_contiguous) else np.zeros(views.shape, np.float32,
                                                                                          order="F")
    img = _f32(image)
    check(lib().sr_project(plan.handle, out.ctypes.data, img.ctypes.data, idx.ctypes.data if idx is not None else None,
                           n, mem))
    _writeback(views, out)
    return views


def project(image, *args, index=None, interpmeth="two", mode="fast", **kw):
    """project(image, plan; kwargs...) or project(image, mumap, psfs, dy; interpmeth, mode, kwargs...)."""
    if len(args) == 1:
        plan = args[0]
        idx, n = _index(plan, index)
        views = np.zeros((plan.imgsize[0], plan.imgsize[2], n), np.float32, order="F")
        return project_(views, image, plan, index=index)
    mumap, psfs, dy = args
    if tuple(np.shape(mumap)) != tuple(np.shape(image)):
        raise DimensionMismatch("image/mumap size")
    if np.shape(image)[1] != np.shape(psfs)[2]:
        raise DimensionMismatch("image/psfs size")
    plan = SPECTplan(mumap, psfs, dy, interpmeth=interpmeth, mode=mode, **kw)
    try:
        return project(image, plan, index=index)
    finally:
        plan.close()


# ---------------------------------------------------------------------------
# backproject! / backproject (src/backproject.jl)
# ---------------------------------------------------------------------------
def backproject_(image, views, plan, *args, index=None, asynchronous=False):
    """backproject!(image, views, plan; index) -- image = Σ_ii A_{index[ii]}' views[:,:,ii]
    (image is overwritten, as the reference zeroes it).  With a 2-D view:
    backproject!(image, view, plan, viewidx) overwrites image with that view's
    back-projection (src/backproject.jl:9-62), and
    backproject!(image, view, plan, buffer_id, viewidx) adds it to image, as
    the reference's per-thread form accumulates (src/backproject.jl:121)."""
    viewidx = _view_args(args, "backproject!")
    if viewidx is not None:
        if np.ndim(views) != 2:
            raise DimensionMismatch("single-view form needs a 2D view")
        if len(args) == 2:
            one = np.zeros(plan.imgsize, np.float32, order="F")
            backproject_(one, np.asarray(views)[:, :, None], plan, index=[viewidx])
            image[...] = np.asarray(image) + one
            return image
        return backproject_(image, np.asarray(views)[:, :, None], plan, index=[viewidx])
    idx, n = _index(plan, index)
    nx, ny, nz = plan.imgsize
    if tuple(image.shape) != plan.imgsize:
        raise DimensionMismatch("image size")
    if tuple(views.shape) != (nx, nz, n):
        raise DimensionMismatch("views size")
    mem = _mem(views, image)
    if mem == L.SR_MEM_DEVICE:
        check(lib().sr_backproject(plan.handle, image.ptr, views.ptr, idx.ctypes.data if idx is not None else None, n,
                                   mem | (L.SR_ASYNC if asynchronous else 0)))
        return image
    out = image if (image.dtype == np.float32 and image.flags.f_contiguous) else np.zeros(image.shape, np.float32,
                                                                                          order="F")
    v = _f32(views)
    check(lib().sr_backproject(plan.handle, out.ctypes.data, v.ctypes.data,
                               idx.ctypes.data if idx is not None else None, n, mem))
    _writeback(image, out)
    return image


def backproject(views, *args, index=None, interpmeth="two", mode="fast", **kw):
    """backproject(views, plan; kwargs...) or backproject(views, mumap, psfs, dy; interpmeth, mode, kwargs...)."""
    if len(args) == 1:
        plan = args[0]
        image = np.zeros(plan.imgsize, np.float32, order="F")
        return backproject_(image, views, plan, index=index)
    mumap, psfs, dy = args
    if not (np.shape(mumap)[0] == np.shape(views)[0]):
        raise DimensionMismatch("nx")
    if np.shape(mumap)[2] != np.shape(views)[1]:
        raise DimensionMismatch("nz")
    plan = SPECTplan(mumap, psfs, dy, interpmeth=interpmeth, mode=mode, **kw)
    try:
        return backproject(views, plan, index=index)
    finally:
        plan.close()


def osem_strided(x0, y, b, plan, nblocks, niter=16):
    """OS-EM over the strided subsets nb:nblocks:nview for any 1 <= nblocks <=
    nview (subset sizes differ by at most one when nblocks does not divide
    nview; Ablock / osem keep the reference's throw).  Host arrays in and out;
    for device arrays call lib().sr_osem_strided with SR_MEM_DEVICE."""
    out = np.zeros(plan.imgsize, np.float32, order="F")
    _check_plan_shapes(plan, out, x0, y, b)
    check(lib().sr_osem_strided(plan.handle, out.ctypes.data, _f32(x0).ctypes.data, _f32(y).ctypes.data,
                                _f32(b).ctypes.data, int(nblocks), int(niter), L.SR_MEM_HOST))
    return out


# ---------------------------------------------------------------------------
# Gradients of L = ||A(x, mu) - y||^2 (no reference counterpart)
# ---------------------------------------------------------------------------
def project_vjp_mu(image, dy, plan, index=None):
    """d<dy, A(x, mu)>/dmu for view cotangents dy (the indexed views in order,
    [nx,nz,n]): the mu part of the projector's vector-Jacobian product (the
    image part is backproject(dy)).  Needs a z-fastest plan."""
    idx, n = _index(plan, index)
    nx, ny, nz = plan.imgsize
    if tuple(np.shape(image)) != plan.imgsize:
        raise DimensionMismatch("image size")
    if tuple(np.shape(dy)) != (nx, nz, n):
        raise DimensionMismatch("views size")
    gmu = np.zeros(plan.imgsize, np.float32, order="F")
    x, d = _f32(image), _f32(dy)
    check(lib().sr_project_vjp_mu(plan.handle, gmu.ctypes.data, x.ctypes.data, d.ctypes.data,
                                  idx.ctypes.data if idx is not None else None, n, L.SR_MEM_HOST))
    return gmu


def gradient(image, y, plan, index=None):
    """(L, dL/dx, dL/dmu) for L = sum over the indexed views of ||A(x, mu) - y||^2,
    mu the plan's mu-map and y the indexed views in order ([nx,nz,n]).  dL/dx is
    the back-projection of 2 (A x - y); dL/dmu runs through the rotated
    exponential cumulative sum of the attenuation.  Needs a z-fastest plan."""
    idx, n = _index(plan, index)
    nx, ny, nz = plan.imgsize
    if tuple(np.shape(image)) != plan.imgsize:
        raise DimensionMismatch("image size")
    if tuple(np.shape(y)) != (nx, nz, n):
        raise DimensionMismatch("views size")
    gx = np.zeros(plan.imgsize, np.float32, order="F")
    gmu = np.zeros(plan.imgsize, np.float32, order="F")
    loss = C.c_double()
    x, yy = _f32(image), _f32(y)
    check(lib().sr_grad(plan.handle, C.byref(loss), gx.ctypes.data, gmu.ctypes.data, x.ctypes.data, yy.ctypes.data,
                        idx.ctypes.data if idx is not None else None, n, L.SR_MEM_HOST))
    return loss.value, gx, gmu


# ---------------------------------------------------------------------------
# LinearMapAA-style operator, Ablock, ML-EM, OS-EM (src/ml-os-em.jl)
# ---------------------------------------------------------------------------
class LinearMapAA:
    """Minimal LinearMapAA(forw!, back!, (M, N); T, odim, idim): A @ x and A.T @ y
    on arrays of shape idim / odim.  `plan`/`index` tag operators that the
    device-resident mlem/osem can run without host round trips."""

    def __init__(self, forw, back, dims, *, T=np.float32, odim=None, idim=None, plan=None, index=None,
                 _adjoint_of=None):
        self.forw, self.back = forw, back
        self.dims = tuple(dims)
        self.T = T
        self.odim, self.idim = tuple(odim), tuple(idim)
        self.plan, self.index = plan, index
        self._adjoint_of = _adjoint_of

    def mul_(self, y, x):
        self.forw(y, x)
        return y

    def __matmul__(self, x):
        y = np.zeros(self.odim, np.float32, order="F")
        self.forw(y, np.asarray(x))
        return y

    @property
    def T_(self):
        return self.adjoint()

    def adjoint(self):
        if self._adjoint_of is not None:
            return self._adjoint_of
        return LinearMapAA(self.back, self.forw, self.dims[::-1], T=self.T, odim=self.idim, idim=self.odim,
                           _adjoint_of=self)

    @property
    def shape(self):
        return self.dims


def spect_operator(plan, index=None):
    """LinearMapAA(forw!, back!, ...) over the plan, as in the reference's
    examples (docs/lit/examples/4-mlem.jl:78-82)."""
    nx, ny, nz = plan.imgsize
    n = plan.nview if index is None else len(index)
    idim, odim = (nx, ny, nz), (nx, nz, n)

    def forw(y, x):
        project_(y, x, plan, index=index)

    def back(x, y):
        backproject_(x, y, plan, index=index)

    return LinearMapAA(forw, back, (int(np.prod(odim)), int(np.prod(idim))), T=np.float32, odim=odim, idim=idim,
                       plan=plan, index=None if index is None else list(index))


def Ablock(plan, nblocks):
    """Vector of LinearMapAA operators for OS-EM: block nb uses views nb:nblocks:nview."""
    if plan.nview % nblocks != 0:
        raise SPECTError("nview must be divisible by nblocks!")
    return [spect_operator(plan, list(range(nb, plan.nview, nblocks))) for nb in range(nblocks)]


def _check_em(out, x0, ynoisy, background):
    if not np.all(np.asarray(background) > 0):
        raise SPECTError("need background > 0")
    if tuple(np.shape(out)) != tuple(np.shape(x0)):
        raise DimensionMismatch("size out and x0 not match")


def _check_plan_shapes(plan, out, x0, ynoisy, background):
    """Before a device-resident EM call: the library reads V floats of x0 and
    P*nview floats of y and b, and writes V floats of out, so every array must
    have the plan's sizes (a wrong size would over-read or overwrite memory)."""
    nx, ny, nz = plan.imgsize
    for name, a in (("out", out), ("x0", x0)):
        if tuple(np.shape(a)) != plan.imgsize:
            raise DimensionMismatch(f"size {name} {tuple(np.shape(a))} and plan image size {plan.imgsize} not match")
    for name, a in (("ynoisy", ynoisy), ("background", background)):
        if tuple(np.shape(a)) != (nx, nz, plan.nview):
            raise DimensionMismatch(f"size {name} {tuple(np.shape(a))} and plan views {(nx, nz, plan.nview)} not match")


def _em_fast(fn, out, x0, ynoisy, background, args, niter, chat, show):
    """Run a device EM call (sr_mlem / sr_osem) on host arrays.  With chat, one
    iteration per call, printing `show(iter, out)` before each: an EM call
    keeps no state besides the image, so the iterates equal those of one call
    with all iterations."""
    res = _f32(out) if isinstance(out, np.ndarray) else out
    x, y, b = _f32(x0), _f32(ynoisy), _f32(background)
    if not chat:
        check(fn(res.ctypes.data, x.ctypes.data, y.ctypes.data, b.ctypes.data, *args, niter, L.SR_MEM_HOST))
        _writeback(out, res)
        return out
    cur = x.copy(order="F")
    t0 = time.time()
    for it in range(niter):
        show(it + 1, cur, time.time() - t0)
        check(fn(res.ctypes.data, cur.ctypes.data, y.ctypes.data, b.ctypes.data, *args, 1, L.SR_MEM_HOST))
        cur[...] = res
    _writeback(out, cur)
    return out


def mlem_(out, x0, ynoisy, background, A, *, niter=20, chat=False):
    """mlem!(out, x0, ynoisy, background, A; niter): x ← x ⊙ A'(y ⊘ (Ax+b)) ⊘ A'1."""
    _check_em(out, x0, ynoisy, background)
    if tuple(np.shape(ynoisy)) != tuple(np.shape(background)):
        raise DimensionMismatch("size ynoisy and background")
    if A.plan is not None and A.index is None:
        plan = A.plan
        _check_plan_shapes(plan, out, x0, ynoisy, background)
        return _em_fast(lambda *a: lib().sr_mlem(plan.handle, *a), out, x0, ynoisy, background, (), niter, chat,
                        lambda it, x, t: print((it, (float(x.min()), float(x.max())), t)))
    At = A.adjoint()
    asum = At @ np.ones(np.shape(ynoisy), np.float32)
    asum[asum == 0] = np.inf
    x = _f32(x0).copy(order="F")
    t0 = time.time()
    for it in range(niter):
        if chat:
            print((it + 1, (float(x.min()), float(x.max())), time.time() - t0))
        ybar = A @ x
        yratio = np.float32(1) * np.asarray(ynoisy, np.float32) / (ybar + np.asarray(background, np.float32))
        back = At @ yratio
        x *= back / asum
    out[...] = x
    return out


def mlem(x0, *args, **kw):
    return mlem_(np.zeros(np.shape(x0), np.float32, order="F"), x0, *args, **kw)


def osem_(out, x0, ynoisy, background, Ab, *, niter=16, chat=False):
    """osem!(out, x0, ynoisy, background, Ab; niter) with Ab = Ablock(plan, nblocks)."""
    _check_em(out, x0, ynoisy, background)
    nblocks = len(Ab)
    nx, nz, nview = np.shape(ynoisy)
    plan = Ab[0].plan
    standard = plan is not None and all(
        a.plan is plan and a.index == list(range(nb, nview, nblocks)) for nb, a in enumerate(Ab))
    if standard and nview == plan.nview:
        _check_plan_shapes(plan, out, x0, ynoisy, background)
        # (chat: one line per iteration, at its first block)
        return _em_fast(lambda *a: lib().sr_osem(plan.handle, *a), out, x0, ynoisy, background, (nblocks,), niter,
                        chat, lambda it, x, t: print((it, 1, (float(x.min()), float(x.max())), t)))
    y = np.asarray(ynoisy, np.float32)
    b = np.asarray(background, np.float32)
    asum = []
    for nb in range(nblocks):
        a = Ab[nb].adjoint() @ np.ones((nx, nz, nview // nblocks), np.float32)
        a[a == 0] = np.inf
        asum.append(a)
    x = _f32(x0).copy(order="F")
    t0 = time.time()
    for it in range(niter):
        for nb in range(nblocks):
            if chat:
                print((it + 1, nb + 1, (float(x.min()), float(x.max())), time.time() - t0))
            ybar = Ab[nb] @ x
            yratio = y[:, :, nb::nblocks] / (ybar + b[:, :, nb::nblocks])
            back = Ab[nb].adjoint() @ yratio
            x *= back / asum[nb]
    out[...] = x
    return out


def osem(x0, *args, **kw):
    return osem_(np.zeros(np.shape(x0), np.float32, order="F"), x0, *args, **kw)


# ---------------------------------------------------------------------------
# Rotation (src/plan-rotate.jl, src/rotatez.jl)
# ---------------------------------------------------------------------------
class PlanRotate:
    def __init__(self, nx, *, T=np.float32, method="two"):
        method = _sym(method)
        if method not in _METHODS:
            warnings.warn(f"method {method}")
        self.nx = nx
        self.T = T
        self.method = method
        self.padsize = math.ceil(1 + nx * math.sqrt(2) / 2 - nx / 2)

    def __repr__(self):
        return f"PlanRotate{{Float32, RotateMode{{:{self.method}}}}} nx={self.nx} padsize={self.padsize}"


def plan_rotate(nx, *, T=np.float32, method="two", nthread=1):
    return [PlanRotate(nx, T=T, method=method) for _ in range(nthread)]


def _rotate(output, img, theta, plan, adj):
    p = plan[0] if isinstance(plan, (list, tuple)) else plan
    a = np.asarray(img)
    if a.ndim == 2:
        if a.shape != np.shape(output):
            raise SPECTError("size")
        if a.shape[0] != a.shape[1]:
            raise SPECTError("row != col")
        a3 = _f32(a)[:, :, None]
    else:
        if a.shape != np.shape(output):
            raise DimensionMismatch()
        if a.shape[0] != p.nx:
            raise SPECTError("nx")
        a3 = _f32(a)
    a3 = np.asfortranarray(a3)
    res = np.zeros(a3.shape, np.float32, order="F")
    check(lib().sr_imrotate(res.ctypes.data, a3.ctypes.data, a3.shape[0], a3.shape[2], float(theta),
                            _METHODS[p.method], int(adj), L.SR_MEM_HOST))
    output[...] = res[:, :, 0] if a.ndim == 2 else res
    return output


def imrotate_(output, img, theta, plan, *_):
    """imrotate!(output, img, θ, plan): counter-clockwise rotation of a 2-D
    image (or of each [:,:,z] slice of a 3-D stack)."""
    return _rotate(output, img, theta, plan, False)


def imrotate_adj_(output, img, theta, plan, *_):
    return _rotate(output, img, theta, plan, True)


def imrotate(img, theta, *, method="two", T=None):
    out = np.zeros(np.shape(img), np.float32, order="F")
    return imrotate_(out, img, theta, plan_rotate(np.shape(img)[0], method=method))


def imrotate_adj(img, theta, *, method="two", T=None):
    out = np.zeros(np.shape(img), np.float32, order="F")
    return imrotate_adj_(out, img, theta, plan_rotate(np.shape(img)[0], method=method))


# ---------------------------------------------------------------------------
# PSF convolution (src/plan-psf.jl, src/fft_convolve.jl)
# ---------------------------------------------------------------------------
def _power2(x):
    return 1 << math.ceil(math.log2(x))


class PlanPSF:
    def __init__(self, *, nx=128, nz=None, px=1, pz=None, T=np.float32):
        nz = nx if nz is None else nz
        pz = px if pz is None else pz
        self.nx, self.nz, self.px, self.pz, self.T = nx, nz, px, pz, T
        up = math.ceil((_power2(nx + px - 1) - nx) / 2)
        down = math.floor((_power2(nx + px - 1) - nx) / 2)
        left = math.ceil((_power2(nz + pz - 1) - nz) / 2)
        right = math.floor((_power2(nz + pz - 1) - nz) / 2)
        self.padsize = (up, down, left, right)

    def __repr__(self):
        return f"PlanPSF{{Float32}} nx={self.nx} nz={self.nz} px={self.px} pz={self.pz} padsize={self.padsize}"


def plan_psf(*, nx=128, nz=None, px=1, pz=None, nthread=1, T=np.float32):
    return [PlanPSF(nx=nx, nz=nz, px=px, pz=pz, T=T) for _ in range(nthread)]


def _conv2(out, img, ker, adj):
    img = _f32(img)
    ker = _f32(ker)
    nx, nz = img.shape
    px, pz = ker.shape
    res = np.zeros((nx, nz), np.float32, order="F")
    check(lib().sr_fft_conv(res.ctypes.data, img.ctypes.data, nx, nz, ker.ctypes.data, px, pz, int(adj),
                            L.SR_MEM_HOST))
    out[...] = res
    return out


def _conv3(out, img, ker, form):
    """All planes of a 3-D form in one sr_fft_conv3 call (out: [nx, ny, nz])."""
    img = _f32(img)
    ker = _f32(ker)
    nx, ny, nz = out.shape
    px, pz = ker.shape[:2]
    res = out if (out.dtype == np.float32 and out.flags.f_contiguous) else np.zeros((nx, ny, nz), np.float32, "F")
    check(lib().sr_fft_conv3(res.ctypes.data, img.ctypes.data, nx, ny, nz, ker.ctypes.data, px, pz, form,
                             L.SR_MEM_HOST))
    if res is not out:
        out[...] = res
    return out


def _plane_checks(p, nx, nz, ker):
    """The size checks each plane's 2-D call makes (src/fft_convolve.jl:39-42)."""
    if (nx, nz) != (p.nx, p.nz):
        raise SPECTError("size img")
    if tuple(np.shape(ker)[:2]) != (p.px, p.pz):
        raise SPECTError(f"size ker {tuple(np.shape(ker)[:2])} {p.px} {p.pz}")


def _conv(output, img, ker, plan, adj):
    p = plan[0] if isinstance(plan, (list, tuple)) else plan
    if np.ndim(img) == 3:  # 3-D form: plane iy of image3 with ker3[:,:,iy] (src/fft_convolve.jl:160-209)
        if np.shape(output) != np.shape(img):
            raise DimensionMismatch()
        nx, ny, nz = np.shape(img)
        _plane_checks(p, nx, nz, ker)
        if np.ndim(ker) != 3 or np.shape(ker)[2] < ny:
            raise IndexError("ker3 has fewer planes than image3")
        return _conv3(output, img, np.asarray(ker)[:, :, :ny], L.SR_CONV3_ADJ if adj else L.SR_CONV3_FWD)
    if np.shape(output) != np.shape(img):
        raise SPECTError("size output")
    if tuple(np.shape(img)) != (p.nx, p.nz):
        raise SPECTError("size img")
    if tuple(np.shape(ker)) != (p.px, p.pz):
        raise SPECTError(f"size ker {tuple(np.shape(ker))} {p.px} {p.pz}")
    return _conv2(output, img, ker, adj)


def fft_conv_(output, img, ker, plan):
    """fft_conv!(output, img, ker, plan): replicate-padded convolution."""
    return _conv(output, img, ker, plan, False)


def fft_conv_adj_(output, img, ker, plan):
    return _conv(output, img, ker, plan, True)


def fft_conv_adj2_(output, image2, ker3, plans):
    """fft_conv_adj2!(output, image2, ker3, plans): adjoint of each plane's kernel on one 2-D image.

    Like src/fft_convolve.jl:217-239 it writes the planes iy = 1..size(image2, 2)
    of output (all of them when ny == nz, leaving the rest as they are when
    nz < ny, and raising IndexError, Julia's BoundsError, when nz > ny).
    """
    if np.shape(output)[0] != np.shape(image2)[0]:
        raise SPECTError("size 1")
    if np.shape(output)[2] != np.shape(image2)[1]:
        raise SPECTError("size 2")
    p = plans[0] if isinstance(plans, (list, tuple)) else plans
    nx, ny, nz = np.shape(output)
    _plane_checks(p, nx, nz, ker3)
    if nz > ny or np.ndim(ker3) != 3 or np.shape(ker3)[2] < nz:
        raise IndexError("fft_conv_adj2!: plane index beyond output or ker3")
    ker = np.asarray(ker3)[:, :, :nz]
    if nz == ny:
        return _conv3(output, image2, ker, L.SR_CONV3_ADJ2)
    part = np.zeros((nx, nz, nz), np.float32, order="F")
    _conv3(part, image2, ker, L.SR_CONV3_ADJ2)
    output[:, :nz, :] = part
    return output


def _check_sym(ker):
    k = np.asarray(ker, np.float64)
    r = k[::-1, ::-1]
    if not np.linalg.norm(k - r) <= math.sqrt(np.finfo(np.float64).eps) * max(np.linalg.norm(k), np.linalg.norm(r)):
        raise SPECTError("asymmetric kernel")


def fft_conv(img, ker, *, T=None):
    _check_sym(ker)
    nx, nz = np.shape(img)
    px, pz = np.shape(ker)
    out = np.zeros((nx, nz), np.float32, order="F")
    return fft_conv_(out, img, ker, plan_psf(nx=nx, nz=nz, px=px, pz=pz))


def fft_conv_adj(img, ker, *, T=None):
    _check_sym(ker)
    nx, nz = np.shape(img)
    px, pz = np.shape(ker)
    out = np.zeros((nx, nz), np.float32, order="F")
    return fft_conv_adj_(out, img, ker, plan_psf(nx=nx, nz=nz, px=px, pz=pz))
