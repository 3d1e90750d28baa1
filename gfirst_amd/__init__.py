"""gfirst_amd — MI355X-native fully-3D OSEM hot path of GFIRST.

The compute path is libgfirst.so (hand-written gfx950 HIP kernels behind the
gfr_* C ABI of include/gfirst.h). This package is the host-side mirror of the
reference's operator interface over that ABI; importing it does not touch the
GPU, and every operator raises if the native library is missing.
"""
from ._lib import (REQ_ACTIVE, REQ_DONE, REQ_FAILED, REQ_PENDING, GFR_ECANCELED, GFR_EINVAL, GFR_EHIP, GFR_EIO, GFR_ENOMEM, GFR_ESTATE,
                   GFR_OK, OUT_PROJ_BLANK,
                   OUT_VLOR, OUT_XD, OUT_XT, STORE_BF16, STORE_F16, STORE_F32, GfrError, lib)
from .core import (Batch, Context, DeviceArray, Recon, Server, Vlor, batch_footprint,
                   batch_params, footprint, recon_params)
from .geometry import C0, C1, PRESETS, REF, SMALL, TINY, Geometry, nseg_for

__all__ = [
    "Context", "DeviceArray", "Recon", "Vlor", "Batch", "batch_params", "batch_footprint", "Server", "Geometry", "recon_params", "footprint", "lib",
    "GfrError", "REF", "C0", "C1", "TINY", "SMALL", "PRESETS", "nseg_for", "STORE_F32",
    "STORE_F16", "STORE_BF16", "OUT_XD", "OUT_XT", "OUT_PROJ_BLANK", "OUT_VLOR", "GFR_OK",
    "GFR_EINVAL", "GFR_EHIP", "GFR_ENOMEM", "GFR_EIO", "GFR_ESTATE", "GFR_ECANCELED", "REQ_PENDING", "REQ_ACTIVE",
    "REQ_DONE", "REQ_FAILED",
]
