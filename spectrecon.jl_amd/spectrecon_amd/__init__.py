"""spectrecon_amd -- MI355X (gfx950) SPECT projector / reconstructor.

A drop-in for the hot path of SPECTrecon.jl: SPECTplan, project!/backproject!
(here project_/backproject_), mlem/osem, with hand-written HIP kernels behind
the C-ABI of include/spectrecon_amd.h.
"""
from ._lib import (DeviceError, DimensionMismatch, SPECTError, LIB_PATH, device_count, lib)  # noqa: F401
from .device import DeviceArray  # noqa: F401
from .api import *  # noqa: F401,F403
from .api import default_viewangle  # noqa: F401

__version__ = "0.1.0"
