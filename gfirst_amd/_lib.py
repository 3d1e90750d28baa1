"""ctypes loader for libgfirst.so (the C ABI declared in include/gfirst.h).

The product path is the native library: if it has not been built this module
raises, there is no Python or CPU fallback.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgfirst.so")
CLI_PATH = os.path.join(HERE, "bin", "gfirst")

GFR_OK = 0
GFR_EINVAL = -1
GFR_EHIP = -2
GFR_ENOMEM = -3
GFR_EIO = -4
GFR_ESTATE = -5
GFR_ECANCELED = -6

STORE_F32, STORE_F16, STORE_BF16 = 0, 1, 2
OUT_XD, OUT_XT, OUT_PROJ_BLANK, OUT_VLOR = 0, 1, 2, 3


class GfrGeometry(C.Structure):
    _fields_ = [("res", C.c_int), ("nzs", C.c_int), ("nrad", C.c_int), ("nang", C.c_int),
                ("span", C.c_int), ("nseg", C.c_int), ("ntbins", C.c_int),
                ("pitch", C.c_float), ("diam_det", C.c_float), ("fov", C.c_float)]


class GfrGeometryInfo(C.Structure):
    _fields_ = [("nsinogs", C.c_int), ("ndata", C.c_longlong), ("nvox", C.c_longlong),
                ("lambda0", C.c_int), ("lambda1", C.c_int)]


class GfrReconParams(C.Structure):
    _fields_ = [("mode", C.c_int), ("niter", C.c_int), ("nsub", C.c_int),
                ("psf_fw", C.c_float), ("psf_bw", C.c_float), ("jbf_r", C.c_int),
                ("jbf_sigma3d", C.c_float), ("jbf_sigmai", C.c_float), ("jbf_coef", C.c_float),
                ("storage", C.c_int)]


class GfrTuning(C.Structure):
    _fields_ = [("back_zb", C.c_int), ("back_lds", C.c_int), ("back_waves", C.c_int),
                ("zsum_sg", C.c_int), ("rot_z", C.c_int), ("back_bw", C.c_int),
                ("ms_back", C.c_float), ("ms_zsum", C.c_float), ("back_ahead", C.c_int)]


class GfrBatchParams(C.Structure):
    _fields_ = [("storage", C.c_int), ("max_frames", C.c_int), ("min_nsub", C.c_int),
                ("lanes", C.c_int), ("psf_fw", C.c_float), ("psf_bw", C.c_float),
                ("group", C.c_int)]


class GfrRequestInfo(C.Structure):
    _fields_ = [("state", C.c_int), ("error", C.c_int), ("t_submit", C.c_double),
                ("t_admit", C.c_double), ("t_done", C.c_double), ("passes", C.c_longlong),
                ("bytes", C.c_longlong)]


class GfrServerTotals(C.Structure):
    _fields_ = [("submitted", C.c_longlong), ("pending", C.c_longlong), ("active", C.c_longlong),
                ("done", C.c_longlong), ("failed", C.c_longlong), ("passes", C.c_longlong),
                ("now", C.c_double)]


REQ_PENDING, REQ_ACTIVE, REQ_DONE, REQ_FAILED = 0, 1, 2, 3


class GfrError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("gfr error %d: %s" % (code, msg))
        self.code = code


_V = C.c_void_p
_G = C.POINTER(GfrGeometry)
_F = C.c_float
_I = C.c_int
_LL = C.c_longlong
_SZ = C.c_size_t

# name -> (restype, argtypes); every symbol declared in include/gfirst.h.
SIGNATURES = {
    "gfr_geometry_default": (_I, [_G]),
    "gfr_geometry_check": (_I, [_G, C.POINTER(GfrGeometryInfo)]),
    "gfr_seg_tables": (_I, [_G, _V, _V, _V]),
    "gfr_angle_tables": (_I, [_G, _V, _V]),
    "gfr_psf_taps": (_I, [_F, _V]),
    "gfr_gz_taps": (_I, [_V]),
    "gfr_ctx_create": (_I, [_I, _SZ, C.POINTER(_V)]),
    "gfr_ctx_destroy": (_I, [_V]),
    "gfr_last_error": (C.c_char_p, [_V]),
    "gfr_ctx_stream": (_V, [_V]),
    "gfr_sync": (_I, [_V, _V]),
    "gfr_alloc": (_I, [_V, _SZ, C.POINTER(_V)]),
    "gfr_free": (_I, [_V, _V]),
    "gfr_bytes_in_use": (_SZ, [_V]),
    "gfr_memcpy_h2d": (_I, [_V, _V, _V, _SZ, _V]),
    "gfr_memcpy_d2h": (_I, [_V, _V, _V, _SZ, _V]),
    "gfr_zero": (_I, [_V, _V, _SZ, _V]),
    "gfr_tuning_get": (_I, [_V, _G, _I, C.POINTER(GfrTuning)]),
    "gfr_tuning_set": (_I, [_V, _G, _I, C.POINTER(GfrTuning)]),
    "gfr_fill_cylinder": (_I, [_V, _G, _F, _V, _V]),
    "gfr_blur_xy": (_I, [_V, _G, _F, _V, _V, _V]),
    "gfr_forward_project": (_I, [_V, _G, _V, _V, _I, _I, _V]),
    "gfr_calib_mask": (_I, [_V, _G, _V, _V, _V, _V]),
    "gfr_ratio": (_I, [_V, _G, _V, _V, _V, _V, _I, _I, _V]),
    "gfr_forward_ratio": (_I, [_V, _G, _V, _V, _V, _V, _V, _I, _I, _V]),
    "gfr_back_project": (_I, [_V, _G, _V, _V, _V, _V, _I, _I, _V]),
    "gfr_update": (_I, [_V, _G, _F, _V, _V, _V, _V, _V]),
    "gfr_back_project_noalloc": (_I, [_V, _G, _V, _V, _V, _V, _I, _I, _V]),
    "gfr_sensitivity": (_I, [_V, _G, _V, _V, _I, _I, _V]),
    "gfr_osem_run": (_I, [_V, _G, C.POINTER(GfrReconParams), _V, _V, _V, _V, _LL, _V, _V]),
    "gfr_jbf3d": (_I, [_V, _G, _V, _V, _F, _I, _F, _F, _V, _V]),
    "gfr_blend": (_I, [_V, _G, _F, _V, _V, _V]),
    "gfr_image_max": (_I, [_V, _G, _V, C.POINTER(_F), _V]),
    "gfr_vlor_create": (_I, [_V, _G, _V, _LL, _V, C.POINTER(_V)]),
    "gfr_vlor_assign": (_I, [_V, _V, _V, _V, _I, _I, _V]),
    "gfr_vlor_destroy": (_I, [_V]),
    "gfr_read_sinogram_file": (_I, [C.c_char_p, _LL, _V, C.POINTER(_I), C.POINTER(_I)]),
    "gfr_read_vlor_file": (_LL, [C.c_char_p, _V, _LL]),
    "gfr_lor_bin": (_LL, [_G, _V, _V]),
    "gfr_vlor_bin_hits": (_LL, [_G, _V, _LL, _F, _V]),
    "gfr_recon_create": (_I, [_V, _G, C.POINTER(GfrReconParams), _V, _V, _V, _V, _LL,
                              C.POINTER(_V)]),
    "gfr_recon_footprint": (_LL, [_G, C.POINTER(GfrReconParams)]),
    "gfr_recon_step": (_I, [_V]),
    "gfr_recon_run": (_I, [_V]),
    "gfr_recon_cursor": (_I, [_V, C.POINTER(_I), C.POINTER(_I), C.POINTER(_I),
                              C.POINTER(_LL)]),
    "gfr_recon_get": (_I, [_V, _I, _V]),
    "gfr_recon_loglik": (_I, [_V, C.POINTER(C.c_double)]),
    "gfr_recon_sync": (_I, [_V]),
    "gfr_recon_stream": (_V, [_V]),
    "gfr_recon_destroy": (_I, [_V]),
    "gfr_recon_profile": (_I, [_V, _I]),
    "gfr_recon_phase_times": (_I, [_V, _V, _V]),
    "gfr_batch_footprint": (_LL, [_G, C.POINTER(GfrBatchParams)]),
    "gfr_batch_frame_footprint": (_LL, [_G, C.POINTER(GfrBatchParams), _I]),
    "gfr_batch_create": (_I, [_V, _G, C.POINTER(GfrBatchParams), _V, C.POINTER(_V)]),
    "gfr_batch_add": (_I, [_V, _I, _I, _V, _V, C.POINTER(_I)]),
    "gfr_batch_step": (_I, [_V]),
    "gfr_batch_run": (_I, [_V]),
    "gfr_batch_cursor": (_I, [_V, _I, C.POINTER(_I), C.POINTER(_I), C.POINTER(_LL)]),
    "gfr_batch_get": (_I, [_V, _I, _V]),
    "gfr_batch_loglik": (_I, [_V, _I, C.POINTER(C.c_double)]),
    "gfr_batch_remove": (_I, [_V, _I]),
    "gfr_batch_use_graph": (_I, [_V, _I]),
    "gfr_batch_stats": (_I, [_V, C.POINTER(_LL), C.POINTER(_LL)]),
    "gfr_batch_stream": (_V, [_V]),
    "gfr_batch_destroy": (_I, [_V]),
    "gfr_server_create": (_I, [_V, _I, _I, _I, C.POINTER(_V)]),
    "gfr_server_submit": (_I, [_V, _G, C.POINTER(GfrReconParams), _V, _V, _V, _V, _LL,
                               C.POINTER(_LL)]),
    "gfr_server_info": (_I, [_V, _LL, C.POINTER(GfrRequestInfo)]),
    "gfr_server_wait": (_I, [_V, _LL, C.c_double]),
    "gfr_server_result": (_I, [_V, _LL, _V, _V]),
    "gfr_server_release": (_I, [_V, _LL]),
    "gfr_server_cancel": (_I, [_V, _LL]),
    "gfr_server_stats": (_I, [_V, C.POINTER(GfrServerTotals)]),
    "gfr_server_destroy": (_I, [_V]),
    "gfr_poisson": (_I, [_V, _V, C.c_double, _V, C.c_double, C.c_ulonglong, _V, _LL, _V]),
}

PHASES = ("blur", "forward", "vlor", "back", "update", "jbf")

_LIB = None


def lib():
    """The loaded libgfirst.so; raises if it is missing (no fallback)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libgfirst.so is not built (%s); run __graft_entry__.build()"
                              % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB
