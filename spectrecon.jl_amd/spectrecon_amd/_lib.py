"""ctypes binding of ``libspectrecon_amd.so`` (the C-ABI in include/spectrecon_amd.h).

The library is the only compute path: if it is missing this module raises at
import-time use, and there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_PKG), "lib", "libspectrecon_amd.so")

SR_OK, SR_ERR_ARG, SR_ERR_DIM, SR_ERR_HIP, SR_ERR_ALLOC, SR_ERR_UNSUPPORTED = range(6)
SR_MEM_HOST, SR_MEM_DEVICE, SR_ASYNC, SR_NO_CHECK, SR_BATCH_STREAMS = 0, 1, 0x100, 0x200, 0x400
SR_COPY_H2D, SR_COPY_D2H, SR_COPY_D2D = 1, 2, 3
SR_CONV3_FWD, SR_CONV3_ADJ, SR_CONV3_ADJ2 = 0, 1, 2
SR_INTERP_ONE, SR_INTERP_TWO = 1, 2
SR_MODE_FAST, SR_MODE_MEM = 0, 1
SR_CONV_AUTO, SR_CONV_DIRECT, SR_CONV_SEPARABLE, SR_CONV_FFT = 0, 1, 2, 3
KERNEL_CLASSES = ("rot_fwd", "psf_fwd", "view_reduce", "psf_adj", "rot_adj", "img_reduce", "atten")


class SPECTError(Exception):
    """An argument check failed; the message is the string the reference throws."""


class DimensionMismatch(ValueError):
    """The reference's DimensionMismatch."""


class DeviceError(RuntimeError):
    """HIP runtime failure, allocation failure, or an unsupported configuration."""


class PlanDesc(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("px", C.c_int32), ("pz", C.c_int32), ("nview", C.c_int32),
        ("dy", C.c_float), ("interpmeth", C.c_int32), ("mode", C.c_int32), ("conv", C.c_int32),
    ]


# every symbol include/spectrecon_amd.h declares, with (restype, argtypes)
_vp, _fp, _dp, _ip = C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_double), C.POINTER(C.c_int32)
SIGNATURES = {
    "sr_last_error": (C.c_char_p, []),
    "sr_version": (C.c_int, []),
    "sr_device_count": (C.c_int, []),
    "sr_malloc": (C.c_int, [C.POINTER(C.c_void_p), C.c_size_t]),
    "sr_free": (C.c_int, [_vp]),
    "sr_memcpy_h2d": (C.c_int, [_vp, _vp, C.c_size_t]),
    "sr_memcpy_d2h": (C.c_int, [_vp, _vp, C.c_size_t]),
    "sr_memcpy_d2d": (C.c_int, [_vp, _vp, C.c_size_t]),
    "sr_memset": (C.c_int, [_vp, C.c_int, C.c_size_t]),
    "sr_device_synchronize": (C.c_int, []),
    "sr_plan_create": (C.c_int, [C.POINTER(PlanDesc), _vp, _vp, _vp, C.POINTER(C.c_void_p)]),
    "sr_plan_destroy": (C.c_int, [_vp]),
    "sr_plan_set_mumap": (C.c_int, [_vp, _vp, C.c_int32]),
    "sr_plan_info": (C.c_int, [_vp, C.POINTER(C.c_size_t), _ip, _ip]),
    "sr_plan_layout": (C.c_int, [_vp, _ip]),
    "sr_grad": (C.c_int, [_vp, C.POINTER(C.c_double), _vp, _vp, _vp, _vp, _vp, C.c_int32, C.c_int32]),
    "sr_project_vjp_mu": (C.c_int, [_vp, _vp, _vp, _vp, _vp, C.c_int32, C.c_int32]),
    "sr_osem_strided": (C.c_int, [_vp, _vp, _vp, _vp, _vp, C.c_int32, C.c_int32, C.c_int32]),
    "sr_plan_query": (C.c_int, [_vp, _ip]),
    "sr_plan_memcpy_async": (C.c_int, [_vp, _vp, _vp, C.c_size_t, C.c_int32]),
    "sr_host_alloc": (C.c_int, [C.POINTER(C.c_void_p), C.c_size_t]),
    "sr_host_free": (C.c_int, [_vp]),
    "sr_osem_batch": (C.c_int, [C.c_int32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, C.c_int32]),
    "sr_batch_stats": (C.c_int, [C.POINTER(C.c_int64)] * 4),
    "sr_plan_set_graph": (C.c_int, [_vp, C.c_int32]),
    "sr_plan_graph_stats": (C.c_int, [_vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "sr_plan_set_choice": (C.c_int, [_vp, _ip]),
    "sr_plan_get_choice": (C.c_int, [_vp, _ip]),
    "sr_plan_autotune": (C.c_int, [_vp, C.c_int32, C.c_char_p, C.c_size_t]),
    "sr_plan_set_batch": (C.c_int, [_vp, C.c_int32]),
    "sr_plan_synchronize": (C.c_int, [_vp]),
    "sr_plan_set_timing": (C.c_int, [_vp, C.c_int32]),
    "sr_plan_get_timing": (C.c_int, [_vp, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    "sr_project": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int32, C.c_int32]),
    "sr_backproject": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int32, C.c_int32]),
    "sr_mlem": (C.c_int, [_vp, _vp, _vp, _vp, _vp, C.c_int32, C.c_int32]),
    "sr_osem": (C.c_int, [_vp, _vp, _vp, _vp, _vp, C.c_int32, C.c_int32, C.c_int32]),
    "sr_imrotate": (C.c_int, [_vp, _vp, C.c_int32, C.c_int32, C.c_float, C.c_int32, C.c_int32, C.c_int32]),
    "sr_fft_conv": (C.c_int, [_vp, _vp, C.c_int32, C.c_int32, _vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32]),
    "sr_fft_conv3": (C.c_int, [_vp, _vp, C.c_int32, C.c_int32, C.c_int32, _vp, C.c_int32, C.c_int32, C.c_int32,
                               C.c_int32]),
    "sr_fft_conv_noalloc": (C.c_int, [_vp, _vp, C.c_int32, C.c_int32, _vp, C.c_int32, C.c_int32, C.c_int32,
                                      C.c_int32]),
    "sr_imrotate_adj": (C.c_int, [_vp, _vp, C.c_int32, C.c_int32, C.c_float, C.c_int32, C.c_int32]),
    "sr_fft_conv_adj": (C.c_int, [_vp, _vp, C.c_int32, C.c_int32, _vp, C.c_int32, C.c_int32, C.c_int32]),
    "sr_psf_gauss": (C.c_int, [_vp, C.c_int32, C.c_int32, C.c_int32, _vp, _vp]),
    "sr_phantom": (C.c_int, [_vp, C.c_int32, C.c_int32, C.c_int32, C.c_uint64]),
    "sr_mumap": (C.c_int, [_vp, C.c_int32, C.c_int32, C.c_int32]),
    "sr_poisson": (C.c_int, [_vp, _vp, C.c_size_t, C.c_uint64]),
}

_LIB = None


def lib():
    """Load the library (once); raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `make -C {os.path.dirname(_PKG)}` "
                "(or __graft_entry__.build()); there is no CPU fallback")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def last_error() -> str:
    return lib().sr_last_error().decode()


def check(rc: int):
    """Turn a C-ABI status into the reference's exception type."""
    if rc == SR_OK:
        return
    msg = last_error()
    if rc == SR_ERR_ARG:
        raise SPECTError(msg)
    if rc == SR_ERR_DIM:
        raise DimensionMismatch(msg)
    raise DeviceError(f"[{rc}] {msg}")


def ptr(a) -> int:
    """Address of a numpy array's data, or a DeviceArray's device pointer."""
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return int(a.ptr)


def device_count() -> int:
    return lib().sr_device_count()
