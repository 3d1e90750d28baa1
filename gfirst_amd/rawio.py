"""Raw-file layouts of the reference (native little-endian, no headers).

  image_volume.raw, img_*.raw   NVOX float32, index z*RES^2 + y*RES + x
  sinogram files                 int32 or float32 (auto-detected), prompts then
                                 optional background half (2*NDATA words max)
  normalization.raw              same reader; only the first NDATA values used
  VLORs.raw                      int32 pairs (i1+1, i2+1) of 1-based bin indices
  proj_ref.raw, SINOG_VLOR.raw   NDATA float32
Readers go through the library (GFIRST_CPU_Kernels.cpp:58-66, 104-142;
GFIRST.cu:229-240 semantics).
"""
import ctypes as C

import numpy as np

from ._lib import GfrError, lib


def write_raw(path, arr, dtype=np.float32):
    """WriteImage (GFIRST_CPU_Kernels.cpp:25-33)."""
    np.ascontiguousarray(arr, dtype).reshape(-1).tofile(path)


def read_image(path, nvox):
    """ReadImage (GFIRST_CPU_Kernels.cpp:58-66): NVOX float32."""
    a = np.fromfile(path, dtype=np.float32, count=nvox)
    if a.size < nvox:
        a = np.concatenate([a, np.zeros(nvox - a.size, np.float32)])
    return a


def read_sinogram(path, ndata):
    """ReadSinogram_Type: returns (2*NDATA float32, found, is_float, truncated)."""
    out = np.zeros(2 * ndata, np.float32)
    isf, tr = C.c_int(), C.c_int()
    rc = lib().gfr_read_sinogram_file(str(path).encode(), ndata, out.ctypes.data, C.byref(isf),
                                      C.byref(tr))
    if rc < 0:
        raise GfrError(rc, "reading %s" % path)
    return out, rc == 0, bool(isf.value), bool(tr.value)


def read_vlor(path):
    """VLOR records as the reference's read loop processes them (N x 2 int32)."""
    n = lib().gfr_read_vlor_file(str(path).encode(), None, 0)
    if n < 0:
        raise GfrError(int(n), "reading %s" % path)
    out = np.zeros(2 * max(n, 1), np.int32)
    if n:
        k = lib().gfr_read_vlor_file(str(path).encode(), out.ctypes.data, n)
        if k != n:
            raise GfrError(int(k), "reading %s" % path)
    return out[:2 * n].reshape(-1, 2)


def output_image(g, x, scale=np.float32(3.70e-4 * 162.0)):
    """Image as written by the reference (GFIRST.cu:532-549): x*scale, zero
    outside the circle of radius RES/2-1."""
    res, h = g.res, g.res // 2
    yy, xx = np.mgrid[0:res, 0:res]
    dist = (xx - h) ** 2 + (yy - h) ** 2 - (h - 1) ** 2
    m = (dist > 0)
    out = (np.asarray(x, np.float32).reshape(g.nzs, res, res) * np.float32(scale))
    out[:, m] = 0.0
    return out.reshape(-1)
