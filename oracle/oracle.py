"""ctypes wrapper of the CPU oracle (``libspect_oracle.so``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s ``cpu_baseline`` leg, as the checker.  The product library
never imports or calls this module.

Arrays follow the reference's (Julia, column-major) index order: ``image[x,y,z]``,
``views[x,z,view]``, ``psfs[px,pz,y,view]``; they are handled here as numpy
arrays in Fortran order so the memory layout equals the C side's.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_f32p = np.ctypeslib.ndpointer(np.float32, flags="F_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="F_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")


def build() -> str:
    """Compile the oracle with its Makefile; returns the library path."""
    import subprocess

    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return os.path.join(_HERE, "libspect_oracle.so")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libspect_oracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        for suf, fp, ft in (("f32", _f32p, C.c_float), ("f64", _f64p, C.c_double)):
            getattr(L, f"or_imrotate_{suf}").argtypes = [fp, fp, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int]
            getattr(L, f"or_fft_conv_{suf}").argtypes = [fp, fp, C.c_int, C.c_int, fp, C.c_int, C.c_int, C.c_int, C.c_int]
            common = [fp, fp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, ft, _f64p, C.c_int]
            getattr(L, f"or_project_{suf}").argtypes = [fp, fp] + common + [_i32p, C.c_int, C.c_int, C.c_int]
            getattr(L, f"or_backproject_{suf}").argtypes = [fp, fp] + common + [_i32p, C.c_int, C.c_int, C.c_int]
            getattr(L, f"or_mlem_{suf}").argtypes = [fp, fp, fp, fp] + common + [C.c_int, C.c_int, C.c_int]
            getattr(L, f"or_osem_{suf}").argtypes = [fp, fp, fp, fp] + common + [C.c_int, C.c_int, C.c_int, C.c_int]
            getattr(L, f"or_grad_{suf}").argtypes = [fp, fp, fp, fp] + common + [_i32p, C.c_int, C.c_int, C.c_int]
            getattr(L, f"or_grad_{suf}").restype = C.c_double
            getattr(L, f"or_atten_{suf}").argtypes = [fp, fp, C.c_int, C.c_int, C.c_int, ft, C.c_double, C.c_int]
            getattr(L, f"or_psf_gauss_{suf}").argtypes = [fp, C.c_int, C.c_int, C.c_int, _f64p, _f64p]
        L.or_rot_padsize.argtypes = [C.c_int]
        L.or_rot_padsize.restype = C.c_int
        L.or_psf_pads.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.or_linearinterp.argtypes = [C.c_double, C.c_double, C.c_int, np.ctypeslib.ndpointer(np.float32),
                                      np.ctypeslib.ndpointer(np.int64)]
        L.or_rot_f90_f32.argtypes = [_f32p, _f32p, C.c_int, C.c_int, C.c_int]
        L.or_rot_f90_f32.restype = C.c_int
        _LIB = L
    return _LIB


def _suf(dtype):
    return "f64" if np.dtype(dtype) == np.float64 else "f32"


def _F(a, dtype):
    return np.asfortranarray(a, dtype=dtype)


def _method(m):
    return {"one": 1, ":one": 1, 1: 1, "two": 2, ":two": 2, 2: 2}[m]


def padsize(nx: int) -> int:
    return lib().or_rot_padsize(nx)


def psf_pads(n: int, p: int):
    up, down = C.c_int(), C.c_int()
    lib().or_psf_pads(n, p, C.byref(up), C.byref(down))
    return up.value, down.value


def linearinterp(s: float, e: float, length: int):
    Cv = np.zeros(2 * length, np.float32)
    J = np.zeros(2 * length, np.int64)
    lib().or_linearinterp(s, e, length, Cv, J)
    return Cv, J


def rot_f90(img, m: int, adj: bool = False):
    img = _F(img, np.float32)
    out = np.zeros_like(img)
    if lib().or_rot_f90_f32(out, img, img.shape[0], m, int(adj)) != 0:
        raise ValueError("invalid m!")
    return out


def imrotate(img, theta: float, method="two", adj=False, dtype=np.float32):
    """Rotate each [:,:,z] slice of a 2-D or 3-D array (src/rotatez.jl)."""
    img = np.asarray(img)
    squeeze = img.ndim == 2
    a = _F(img[:, :, None] if squeeze else img, dtype)
    out = np.zeros_like(a)
    getattr(lib(), f"or_imrotate_{_suf(dtype)}")(out, a, a.shape[0], a.shape[2], float(theta), _method(method), int(adj))
    return out[:, :, 0] if squeeze else out


def fft_conv(img, ker, adj=False, kind="fft", dtype=np.float32):
    """fft_conv / fft_conv_adj (src/fft_convolve.jl); kind 'fft' is the literal
    FFT restatement, 'direct' the closed-form clamped convolution."""
    img = _F(img, dtype)
    ker = _F(ker, dtype)
    out = np.zeros_like(img)
    nz = img.shape[1] if img.ndim > 1 else 1
    pz = ker.shape[1] if ker.ndim > 1 else 1
    getattr(lib(), f"or_fft_conv_{_suf(dtype)}")(out, img, img.shape[0], nz, ker, ker.shape[0], pz, int(adj),
                                                 0 if kind == "fft" else 1)
    return out


class Problem:
    """Inputs of one SPECT system model (the oracle's analogue of SPECTplan)."""

    def __init__(self, mumap, psfs, dy, angles, method="two", dtype=np.float32, conv="fft", nthread=1):
        self.dtype = np.dtype(dtype)
        self.mumap = _F(mumap, dtype)
        self.psfs = _F(psfs, dtype)
        self.dy = float(dy)
        self.angles = np.asfortranarray(np.asarray(angles, dtype=np.float64))
        self.method = _method(method)
        self.conv = 0 if conv == "fft" else 1
        self.nthread = nthread
        self.nx, self.ny, self.nz = self.mumap.shape
        self.px, self.pz, _, self.nview = self.psfs.shape

    def _common(self):
        ft = C.c_double if self.dtype == np.float64 else C.c_float
        return [self.mumap, self.psfs, self.nx, self.ny, self.nz, self.px, self.pz, self.nview, ft(self.dy),
                self.angles, self.method]

    def project(self, image, index=None):
        index = np.arange(self.nview, dtype=np.int32) if index is None else np.ascontiguousarray(index, np.int32)
        views = np.zeros((self.nx, self.nz, len(index)), self.dtype, order="F")
        getattr(lib(), f"or_project_{_suf(self.dtype)}")(views, _F(image, self.dtype), *self._common(), index,
                                                         len(index), self.nthread, self.conv)
        return views

    def backproject(self, views, index=None):
        index = np.arange(self.nview, dtype=np.int32) if index is None else np.ascontiguousarray(index, np.int32)
        image = np.zeros((self.nx, self.ny, self.nz), self.dtype, order="F")
        getattr(lib(), f"or_backproject_{_suf(self.dtype)}")(image, _F(views, self.dtype), *self._common(), index,
                                                             len(index), self.nthread, self.conv)
        return image

    def mlem(self, x0, y, b, niter):
        out = np.zeros((self.nx, self.ny, self.nz), self.dtype, order="F")
        getattr(lib(), f"or_mlem_{_suf(self.dtype)}")(out, _F(x0, self.dtype), _F(y, self.dtype), _F(b, self.dtype),
                                                      *self._common(), niter, self.nthread, self.conv)
        return out

    def osem(self, x0, y, b, nblocks, niter):
        out = np.zeros((self.nx, self.ny, self.nz), self.dtype, order="F")
        getattr(lib(), f"or_osem_{_suf(self.dtype)}")(out, _F(x0, self.dtype), _F(y, self.dtype), _F(b, self.dtype),
                                                      *self._common(), nblocks, niter, self.nthread, self.conv)
        return out

    def grad(self, image, y, index=None):
        """(L, dL/dx, dL/dmu) for L = sum over the indexed views of ||A x - y||^2,
        y holding the indexed views in order; mu is the plan's mu-map."""
        index = np.arange(self.nview, dtype=np.int32) if index is None else np.ascontiguousarray(index, np.int32)
        gx = np.zeros((self.nx, self.ny, self.nz), self.dtype, order="F")
        gmu = np.zeros((self.nx, self.ny, self.nz), self.dtype, order="F")
        loss = getattr(lib(), f"or_grad_{_suf(self.dtype)}")(gx, gmu, _F(image, self.dtype), _F(y, self.dtype),
                                                             *self._common(), index, len(index), self.nthread,
                                                             self.conv)
        return float(loss), gx, gmu

    def atten(self, view: int):
        """Attenuation factor exp(-dy(Σ_{j≤y} μr_j - μr_y/2)) for one view."""
        out = np.zeros((self.nx, self.ny, self.nz), self.dtype, order="F")
        ft = C.c_double if self.dtype == np.float64 else C.c_float
        getattr(lib(), f"or_atten_{_suf(self.dtype)}")(out, self.mumap, self.nx, self.ny, self.nz, ft(self.dy),
                                                       float(self.angles[view]), self.method)
        return out


def psf_gauss(ny, px, pz=None, fwhm_start=1.0, fwhm_end=4.0, fwhm=None, fwhm_x=None, fwhm_z=None,
              dtype=np.float32):
    pz = px if pz is None else pz
    fwhm = np.linspace(fwhm_start, fwhm_end, ny) if fwhm is None else np.asarray(fwhm, np.float64)
    fx = np.asfortranarray(fwhm if fwhm_x is None else fwhm_x, np.float64)
    fz = np.asfortranarray(fx if fwhm_z is None else fwhm_z, np.float64)
    out = np.zeros((px, pz, ny), dtype, order="F")
    getattr(lib(), f"or_psf_gauss_{_suf(dtype)}")(out, ny, px, pz, fx, fz)
    return out
