"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker. The product (gfirst_amd/) never
imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class OrcGeom(C.Structure):
    _fields_ = [("res", C.c_int), ("nzs", C.c_int), ("nrad", C.c_int), ("nang", C.c_int),
                ("span", C.c_int), ("nseg", C.c_int), ("ntbins", C.c_int),
                ("pitch", C.c_float), ("diam_det", C.c_float), ("fov", C.c_float)]


class OrcParams(C.Structure):
    _fields_ = [("mode", C.c_int), ("niter", C.c_int), ("nsub", C.c_int),
                ("psf_fw", C.c_double), ("psf_bw", C.c_double), ("jbf_r", C.c_int),
                ("jbf_sigma3d", C.c_double), ("jbf_sigmai", C.c_double), ("jbf_coef", C.c_double)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        _LIB = C.CDLL(path)
        _declare(_LIB)
    return _LIB


_f32 = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64 = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32 = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u32 = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_G = C.POINTER(OrcGeom)
_P = C.POINTER(OrcParams)


def _opt(t):
    """ndpointer that also accepts None."""
    class _O(t):
        @classmethod
        def from_param(cls, obj):
            if obj is None:
                return None
            return t.from_param(obj)
    return _O


def _declare(L):
    sig = {
        "orc_nsinogs": (C.c_int, [_G]),
        "orc_ndata": (C.c_longlong, [_G]),
        "orc_nvox": (C.c_longlong, [_G]),
        "orc_seg_tables": (C.c_int, [_G, _i32, _i32, _i32]),
        "orc_set_threads": (None, [C.c_int]),
        "orc_get_threads": (C.c_int, []),
        "orc_angle_tables": (None, [_G, _f32, _f32]),
        "orc_psf_taps": (None, [C.c_float, _f32]),
        "orc_gz_taps": (None, [_f32]),
        "orc_cylinder": (None, [_G, C.c_float, _f32]),
        "orc_tex3": (None, [_f32, C.c_int, C.c_int, C.c_int, _f32, _f32, C.c_longlong]),
        "orc_blur_xy": (None, [_G, C.c_float, _f32, _f32]),
        "orc_forward": (None, [_G, _f32, _f32, C.c_int, C.c_int]),
        "orc_calib": (None, [_G, _f32, _f32, _f32]),
        "orc_ratio": (None, [_G, _f32, _f32, _opt(_f32), _f32, C.c_int, C.c_int]),
        "orc_back": (None, [_G, _f32, _f32, _f32, _f32, C.c_int, C.c_int]),
        "orc_update": (None, [_G, C.c_float, _f32, _f32, _f32]),
        "orc_jbf": (None, [_G, _f32, _f32, C.c_float, C.c_int, C.c_float, C.c_float, _f32]),
        "orc_blend": (None, [_G, C.c_float, _f32, _f32]),
        "orc_max": (C.c_float, [_f32, C.c_longlong]),
        "orc_vlor_precount": (None, [_G, _i32, C.c_longlong, _f32]),
        "orc_vlor_blankt": (None, [_G, _f32, _f32, _f32]),
        "orc_vlor_assign": (None, [_G, _i32, C.c_longlong, _f32, _f32, _f32, C.c_int, C.c_int]),
        "orc_output": (None, [_G, C.c_float, _f32, _f32]),
        "orc_loglik": (C.c_double, [_G, _f32, _f32, _opt(_f32), _f32, C.c_int, C.c_int,
                                    C.POINTER(C.c_double)]),
        "orc_osem": (C.c_int, [_G, _P, _f32, _opt(_f32), _f32, _opt(_i32), C.c_longlong,
                               _f32, _opt(_f32), _opt(_f32), _opt(_f32)]),
        "orc64_angle_tables": (None, [_G, _f64, _f64]),
        "orc64_psf_taps": (None, [C.c_double, _f64]),
        "orc64_gz_taps": (None, [_f64]),
        "orc64_cylinder": (None, [_G, C.c_double, _f64]),
        "orc64_blur_xy": (None, [_G, C.c_double, _f64, _f64]),
        "orc64_forward": (None, [_G, _f64, _f64, C.c_int, C.c_int]),
        "orc64_calib": (None, [_G, _f64, _f64, _f64]),
        "orc64_ratio": (None, [_G, _f64, _f64, _opt(_f64), _f64, C.c_int, C.c_int]),
        "orc64_back": (None, [_G, _f64, _f64, _f64, _f64, C.c_int, C.c_int]),
        "orc64_update": (None, [_G, C.c_double, _f64, _f64, _f64]),
        "orc64_jbf": (None, [_G, _f64, _f64, C.c_double, C.c_int, C.c_double, C.c_double, _f64]),
        "orc64_osem": (C.c_int, [_G, _P, _f32, _opt(_f32), _f32, _opt(_i32), C.c_longlong,
                                 _f64, _opt(_f64), _opt(_f64), _opt(_f64)]),
        "orc_decode_sinogram": (C.c_int, [_u32, C.c_longlong, C.c_longlong, _f32]),
        "orc_vlor_records": (C.c_longlong, [_i32, C.c_longlong, _i32]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args


def geom(g):
    """OrcGeom from any object with the geometry attributes."""
    return OrcGeom(g.res, g.nzs, g.nrad, g.nang, g.span, g.nseg, g.ntbins,
                   g.pitch, g.diam_det, g.fov)


def params(mode=1, niter=1, nsub=1, psf_fw=1.5, psf_bw=1.5, jbf_r=4, jbf_sigma3d=0.75,
           jbf_sigmai=0.025, jbf_coef=0.1):
    return OrcParams(mode, niter, nsub, psf_fw, psf_bw, jbf_r, jbf_sigma3d, jbf_sigmai, jbf_coef)


def _dt(fp64):
    return np.float64 if fp64 else np.float32


def _p(fp64):
    return "orc64_" if fp64 else "orc_"


def set_threads(n):
    lib().orc_set_threads(int(n))


def get_threads():
    return lib().orc_get_threads()


def nsinogs(g):
    return lib().orc_nsinogs(C.byref(geom(g)))


def ndata(g):
    return lib().orc_ndata(C.byref(geom(g)))


def nvox(g):
    return lib().orc_nvox(C.byref(geom(g)))


def seg_tables(g):
    zmin = np.zeros(g.nseg, np.int32)
    sgn = np.zeros(g.nseg, np.int32)
    nza = np.zeros(g.nseg + 1, np.int32)
    n = lib().orc_seg_tables(C.byref(geom(g)), zmin, sgn, nza)
    return n, zmin, sgn, nza


def angle_tables(g, fp64=False):
    s = np.zeros(g.nang, _dt(fp64))
    c = np.zeros(g.nang, _dt(fp64))
    getattr(lib(), _p(fp64) + "angle_tables")(C.byref(geom(g)), s, c)
    return s, c


def psf_taps(psf, fp64=False):
    out = np.zeros(9, _dt(fp64))
    getattr(lib(), _p(fp64) + "psf_taps")(psf, out)
    return out


def gz_taps(fp64=False):
    out = np.zeros(5, _dt(fp64))
    getattr(lib(), _p(fp64) + "gz_taps")(out)
    return out


def cylinder(g, val=1.0, fp64=False):
    out = np.zeros(nvox(g), _dt(fp64))
    getattr(lib(), _p(fp64) + "cylinder")(C.byref(geom(g)), val, out)
    return out


def tex3(vol, W, H, D, xyz):
    xyz = np.ascontiguousarray(xyz, np.float32)
    out = np.zeros(len(xyz), np.float32)
    lib().orc_tex3(np.ascontiguousarray(vol, np.float32), W, H, D, xyz, out, len(xyz))
    return out


def blur_xy(g, img, psf, fp64=False):
    img = np.ascontiguousarray(img, _dt(fp64))
    out = np.zeros_like(img)
    getattr(lib(), _p(fp64) + "blur_xy")(C.byref(geom(g)), psf, img, out)
    return out


def forward(g, img, subset=0, nsub=1, out=None, fp64=False):
    img = np.ascontiguousarray(img, _dt(fp64))
    if out is None:
        out = np.zeros(ndata(g), _dt(fp64))
    getattr(lib(), _p(fp64) + "forward")(C.byref(geom(g)), img, out, subset, nsub)
    return out


def calib(g, proj_blank, blank, fp64=False):
    dt = _dt(fp64)
    out = np.zeros(ndata(g), dt)
    getattr(lib(), _p(fp64) + "calib")(C.byref(geom(g)), np.ascontiguousarray(proj_blank, dt),
                                       np.ascontiguousarray(blank, dt), out)
    return out


def ratio(g, proj, y, bg, subset=0, nsub=1, out=None, fp64=False):
    dt = _dt(fp64)
    if out is None:
        out = np.zeros(ndata(g), dt)
    bgp = None if bg is None else np.ascontiguousarray(bg, dt)
    getattr(lib(), _p(fp64) + "ratio")(C.byref(geom(g)), np.ascontiguousarray(proj, dt),
                                       np.ascontiguousarray(y, dt), bgp, out, subset, nsub)
    return out


def back(g, corr, cal, subset=0, nsub=1, fp64=False):
    dt = _dt(fp64)
    num = np.zeros(nvox(g), dt)
    den = np.zeros(nvox(g), dt)
    getattr(lib(), _p(fp64) + "back")(C.byref(geom(g)), np.ascontiguousarray(corr, dt),
                                      np.ascontiguousarray(cal, dt), num, den, subset, nsub)
    return num, den


def update(g, x, num, den, psf_bw, fp64=False):
    dt = _dt(fp64)
    x = np.array(x, dt, copy=True)
    getattr(lib(), _p(fp64) + "update")(C.byref(geom(g)), psf_bw, np.ascontiguousarray(num, dt),
                                        np.ascontiguousarray(den, dt), x)
    return x


def jbf(g, inp, guide, max_ref, r=4, sigma3d=0.75, sigmai=0.025, fp64=False):
    dt = _dt(fp64)
    out = np.zeros(nvox(g), dt)
    getattr(lib(), _p(fp64) + "jbf")(C.byref(geom(g)), np.ascontiguousarray(inp, dt),
                                     np.ascontiguousarray(guide, dt), max_ref, r, sigma3d,
                                     sigmai, out)
    return out


def blend(g, x, filt, coef=0.1):
    x = np.array(x, np.float32, copy=True)
    lib().orc_blend(C.byref(geom(g)), coef, np.ascontiguousarray(filt, np.float32), x)
    return x


def vmax(v):
    return lib().orc_max(np.ascontiguousarray(v, np.float32), len(v))


def vlor_precount(g, rec):
    rec = np.ascontiguousarray(rec, np.int32).reshape(-1)
    out = np.zeros(ndata(g), np.float32)
    lib().orc_vlor_precount(C.byref(geom(g)), rec, len(rec) // 2, out)
    return out


def vlor_blankt(g, yT, v0):
    out = np.zeros(ndata(g), np.float32)
    lib().orc_vlor_blankt(C.byref(geom(g)), np.ascontiguousarray(yT, np.float32),
                          np.ascontiguousarray(v0, np.float32), out)
    return out


def vlor_assign(g, rec, blankt, proj, subset=0, nsub=1, dv=None):
    rec = np.ascontiguousarray(rec, np.int32).reshape(-1)
    if dv is None:
        dv = np.zeros(ndata(g), np.float32)
    lib().orc_vlor_assign(C.byref(geom(g)), rec, len(rec) // 2,
                          np.ascontiguousarray(blankt, np.float32),
                          np.ascontiguousarray(proj, np.float32), dv, subset, nsub)
    return dv


def output(g, x, scale):
    out = np.zeros(nvox(g), np.float32)
    lib().orc_output(C.byref(geom(g)), scale, np.ascontiguousarray(x, np.float32), out)
    return out


def loglik(g, proj, y, bg, cal, subset=0, nsub=1):
    """FP64 Poisson log-likelihood of a pass (y, bg RES-scaled; cal the 1/-1000
    calibration array). Returns (ll, sum of |terms|)."""
    a = C.c_double(0.0)
    f = lambda v: np.ascontiguousarray(v, np.float32)
    ll = lib().orc_loglik(C.byref(geom(g)), f(proj), f(y), None if bg is None else f(bg),
                          f(cal), subset, nsub, C.byref(a))
    return float(ll), float(a.value)


def osem(g, prm, d2, t2, norm, rec=None, fp64=False, want_vlor=False):
    """Whole reconstruction. d2/t2: 2*NDATA (prompts, background), unscaled."""
    dt = _dt(fp64)
    nv = nvox(g)
    nd = ndata(g)
    xD = np.zeros(nv, dt)
    xT = np.zeros(nv, dt) if prm.mode >= 2 else None
    pb = np.zeros(nd, dt)
    vo = np.zeros(nd, dt) if want_vlor else None
    recp = None if rec is None else np.ascontiguousarray(rec, np.int32).reshape(-1)
    nrec = 0 if rec is None else len(recp) // 2
    rc = getattr(lib(), _p(fp64) + "osem")(
        C.byref(geom(g)), C.byref(prm), np.ascontiguousarray(d2, np.float32),
        None if t2 is None else np.ascontiguousarray(t2, np.float32),
        np.ascontiguousarray(norm, np.float32), recp, nrec, xD, xT, pb, vo)
    if rc != 0:
        raise RuntimeError("oracle osem failed: %d" % rc)
    return xD, xT, pb, vo


def decode_sinogram(raw_bytes, nd):
    n = len(raw_bytes) // 4
    words = np.frombuffer(raw_bytes[:4 * n], dtype=np.uint32).copy()
    out = np.zeros(2 * nd, np.float32)
    kind = lib().orc_decode_sinogram(words, n, nd, out)
    return out, kind


def vlor_records(raw_bytes):
    n = len(raw_bytes) // 4
    words = np.frombuffer(raw_bytes[:4 * n], dtype=np.int32).copy()
    out = np.zeros(2 * (n // 2 + 1), np.int32)
    k = lib().orc_vlor_records(words, n, out)
    return out[:2 * k].reshape(-1, 2)
