"""Python mirror of the reference's operator interface over the gfr_* C ABI.

Each method of Context corresponds to one host wrapper of the reference (the
C function it calls cites the reference file:line). Arguments are DeviceArray
(device-resident) or numpy arrays (uploaded/downloaded around the call).
Errors raise GfrError with the library's message; nothing falls back to the
CPU.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import GfrBatchParams, GfrError, GfrReconParams, lib


def _check(ctx, rc):
    if rc < 0:
        msg = lib().gfr_last_error(ctx).decode() if ctx else ""
        raise GfrError(rc, msg)
    return rc


class DeviceArray:
    """A device allocation owned by a Context (gfr_alloc/gfr_free)."""

    def __init__(self, ctx, n, dtype=np.float32):
        self.ctx = ctx
        self.dtype = np.dtype(dtype)
        self.n = int(n)
        self.nbytes = self.n * self.dtype.itemsize
        p = C.c_void_p()
        _check(ctx.h, lib().gfr_alloc(ctx.h, self.nbytes, C.byref(p)))
        self.ptr = p.value

    def upload(self, arr):
        a = np.ascontiguousarray(arr, self.dtype).reshape(-1)
        if a.size != self.n:
            raise ValueError("size mismatch: %d vs %d" % (a.size, self.n))
        _check(self.ctx.h, lib().gfr_memcpy_h2d(self.ctx.h, self.ptr, a.ctypes.data, self.nbytes,
                                               None))
        return self

    def download(self):
        out = np.empty(self.n, self.dtype)
        _check(self.ctx.h, lib().gfr_memcpy_d2h(self.ctx.h, out.ctypes.data, self.ptr,
                                               self.nbytes, None))
        return out

    def zero(self):
        _check(self.ctx.h, lib().gfr_zero(self.ctx.h, self.ptr, self.nbytes, None))
        return self

    def free(self):
        if self.ptr:
            lib().gfr_free(self.ctx.h, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            if self.ptr and self.ctx.h:
                self.free()
        except Exception:
            pass


class Context:
    """One device context (replaces the reference's implicit CUDA state)."""

    def __init__(self, device=0, hbm_budget=0):
        h = C.c_void_p()
        rc = lib().gfr_ctx_create(device, hbm_budget, C.byref(h))
        if rc != 0:
            raise GfrError(rc, "gfr_ctx_create(device=%d) failed" % device)
        self.h = h.value

    def close(self):
        if self.h:
            lib().gfr_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- memory ---------------------------------------------------------------
    def array(self, n, dtype=np.float32):
        return DeviceArray(self, n, dtype)

    def to_device(self, arr, dtype=np.float32):
        a = np.ascontiguousarray(arr, dtype).reshape(-1)
        return DeviceArray(self, a.size, dtype).upload(a)

    def sync(self, stream=None):
        _check(self.h, lib().gfr_sync(self.h, stream))

    def bytes_in_use(self):
        return lib().gfr_bytes_in_use(self.h)

    def _dev(self, x, n=None, dtype=np.float32):
        if isinstance(x, DeviceArray):
            return x, False
        if x is None:
            return None, False
        return self.to_device(x, dtype), True

    # -- tile shapes of the projectors -----------------------------------------
    def tuning(self, g, nsub):
        """Launch shapes for (geometry, nsub) as a dict (+ "measured")."""
        t = _lib.GfrTuning()
        rc = _check(self.h, lib().gfr_tuning_get(self.h, C.byref(g.c()), nsub, C.byref(t)))
        d = {k: getattr(t, k) for k, _ in t._fields_}
        d["measured"] = bool(rc)
        return d

    def set_tuning(self, g, nsub, back_zb=8, zsum_sg=8, back_lds=False, back_waves=0, rot_z=16,
                   back_bw=4, back_ahead=1):
        """Fix the launch shapes for (geometry, nsub) instead of autotuning them:
        back_zb slices per backprojector thread (1 or 2 with cached reads; 4, 5,
        6, 8, 16), back_lds the
        per-block LDS row window, back_waves 6 (8 slices, cached reads) or 0,
        zsum_sg sinograms per z-sum wave (4, 6, 8), rot_z planes per rotation
        block (4, 16, 32), back_bw waves per backprojector block (4, or 1 with
        cached reads and back_waves 0), back_ahead views whose rows are loaded
        before the first is summed (2 or 4 with one slice and cached reads).
        Applies to sessions created afterwards; no shape changes any result."""
        t = _lib.GfrTuning(back_zb, int(back_lds), back_waves, zsum_sg, rot_z, back_bw, 0.0, 0.0,
                           back_ahead)
        _check(self.h, lib().gfr_tuning_set(self.h, C.byref(g.c()), nsub, C.byref(t)))

    # -- operators (one per reference wrapper) ---------------------------------
    def fill_cylinder(self, g, val, out=None):
        """CreateInitialCylinder (GFIRST_CPU_Kernels.cpp:35-56)."""
        o = out if out is not None else self.array(g.nvox)
        _check(self.h, lib().gfr_fill_cylinder(self.h, C.byref(g.c()), val, o.ptr, None))
        return o if out is not None else o.download()

    def blur_xy(self, g, img, psf, out=None):
        """convolutionXY + calcGaussianCoefficients (GFIRST_GPU_Kernels.cu:322-364)."""
        d, tmp = self._dev(img)
        o = out if out is not None else self.array(g.nvox)
        _check(self.h, lib().gfr_blur_xy(self.h, C.byref(g.c()), psf, d.ptr, o.ptr, None))
        return o if out is not None else o.download()

    def forward_project(self, g, img, subset=0, nsub=1, out=None):
        """FORWPROJ over the subset's views + scatter (GFIRST_GPU_Kernels.cu:40-107,
        GFIRST.cu:346-363). Unwritten bins are zero when out is None."""
        d, _ = self._dev(img)
        o = out if out is not None else self.array(g.ndata).zero()
        _check(self.h, lib().gfr_forward_project(self.h, C.byref(g.c()), d.ptr, o.ptr, subset,
                                                 nsub, None))
        return o if out is not None else o.download()

    def calib_mask(self, g, proj_blank, norm, out=None):
        """Mask of GFIRST.cu:453-458 (1 valid, 0 gap)."""
        pb, _ = self._dev(proj_blank)
        nm, _ = self._dev(norm)
        o = out if out is not None else self.array(g.ndata, np.uint8)
        _check(self.h, lib().gfr_calib_mask(self.h, C.byref(g.c()), pb.ptr, nm.ptr, o.ptr, None))
        return o if out is not None else o.download()

    def ratio(self, g, proj, y, bg, subset=0, nsub=1, out=None):
        """Ratio of GFIRST.cu:413-430 (y, bg already x RES)."""
        p, _ = self._dev(proj)
        yy, _ = self._dev(y)
        bb, _ = self._dev(bg)
        o = out if out is not None else self.array(g.ndata).zero()
        _check(self.h, lib().gfr_ratio(self.h, C.byref(g.c()), p.ptr, yy.ptr,
                                       bb.ptr if bb else None, o.ptr, subset, nsub, None))
        return o if out is not None else o.download()

    def forward_ratio(self, g, img, y, bg, subset=0, nsub=1, out=None, proj_out=None):
        """Fused forward projection + ratio for the subset's bins."""
        d, _ = self._dev(img)
        yy, _ = self._dev(y)
        bb, _ = self._dev(bg)
        o = out if out is not None else self.array(g.ndata).zero()
        _check(self.h, lib().gfr_forward_ratio(self.h, C.byref(g.c()), d.ptr, yy.ptr,
                                               bb.ptr if bb else None, o.ptr,
                                               proj_out.ptr if proj_out is not None else None,
                                               subset, nsub, None))
        return o if out is not None else o.download()

    def back_project(self, g, corr, mask, subset=0, nsub=1, numer=None, denom=None):
        """Per-view BACKWPROJ loop of GFIRST.cu:436-469 (GFIRST_GPU_Kernels.cu:109-171)."""
        c, _ = self._dev(corr)
        m, _ = self._dev(mask, dtype=np.uint8)
        n = numer if numer is not None else self.array(g.nvox)
        d = denom if denom is not None else self.array(g.nvox)
        _check(self.h, lib().gfr_back_project(self.h, C.byref(g.c()), c.ptr, m.ptr, n.ptr, d.ptr,
                                              subset, nsub, None))
        if numer is not None:
            return n, d
        return n.download(), d.download()

    def back_project_noalloc(self, g, corr, mask, subset, nsub, numer, denom):
        """back_project into caller-given DeviceArrays through the library's
        fixed scratch array (gfr_back_project_noalloc): no device allocation."""
        _check(self.h, lib().gfr_back_project_noalloc(self.h, C.byref(g.c()), corr.ptr, mask.ptr,
                                                      numer.ptr, denom.ptr, subset, nsub, None))

    def update(self, g, x, numer, denom, psf_bw, work=None):
        """Image update of GFIRST.cu:482-502 (in place on a DeviceArray x)."""
        xd, copied = self._dev(x)
        n, _ = self._dev(numer)
        d, _ = self._dev(denom)
        w = work if work is not None else self.array(g.nvox)
        _check(self.h, lib().gfr_update(self.h, C.byref(g.c()), psf_bw, n.ptr, d.ptr, xd.ptr,
                                        w.ptr, None))
        return xd.download() if copied else xd

    def sensitivity(self, g, mask, subset=0, nsub=1, out=None):
        """Calibration half of the backprojection (FP64 accumulation), NVOX doubles."""
        m, _ = self._dev(mask, dtype=np.uint8)
        o = out if out is not None else self.array(g.nvox, np.float64)
        _check(self.h, lib().gfr_sensitivity(self.h, C.byref(g.c()), m.ptr, o.ptr, subset, nsub,
                                             None))
        if out is not None:
            return o
        self.sync()
        return o.download()

    def osem_run(self, g, params, d2, t2, norm, records=None):
        """Whole reconstruction in one call; returns (xD, xT or None)."""
        d2 = np.ascontiguousarray(d2, np.float32).reshape(-1)
        norm = np.ascontiguousarray(norm, np.float32).reshape(-1)
        t2p = None
        if t2 is not None:
            t2 = np.ascontiguousarray(t2, np.float32).reshape(-1)
            t2p = t2.ctypes.data
        rec = None if records is None else np.ascontiguousarray(records, np.int32).reshape(-1)
        xD = np.empty(g.nvox, np.float32)
        xT = np.empty(g.nvox, np.float32) if params.mode >= 2 else None
        _check(self.h, lib().gfr_osem_run(self.h, C.byref(g.c()), C.byref(params), d2.ctypes.data,
                                          t2p, norm.ctypes.data,
                                          rec.ctypes.data if rec is not None else None,
                                          0 if rec is None else rec.size // 2, xD.ctypes.data,
                                          xT.ctypes.data if xT is not None else None))
        return xD, xT

    def jbf3d(self, g, inp, guide, max_ref, r=4, sigma3d=0.75, sigmai=0.025, out=None):
        """jointbilateralfilter3D (GFIRST_GPU_Kernels.cu:508-544)."""
        i, _ = self._dev(inp)
        gd, _ = self._dev(guide)
        o = out if out is not None else self.array(g.nvox)
        _check(self.h, lib().gfr_jbf3d(self.h, C.byref(g.c()), i.ptr, gd.ptr, max_ref, r,
                                       sigma3d, sigmai, o.ptr, None))
        return o if out is not None else o.download()

    def blend(self, g, x, filt, coef=0.1):
        """x = (1-coef) x + coef filt (GFIRST.cu:524), in place on a DeviceArray."""
        xd, copied = self._dev(x)
        f, _ = self._dev(filt)
        _check(self.h, lib().gfr_blend(self.h, C.byref(g.c()), coef, f.ptr, xd.ptr, None))
        return xd.download() if copied else xd

    def image_max(self, g, x):
        xd, _ = self._dev(x)
        m = C.c_float()
        _check(self.h, lib().gfr_image_max(self.h, C.byref(g.c()), xd.ptr, C.byref(m), None))
        return m.value

    def vlor(self, g, records, yT_scaled):
        return Vlor(self, g, records, yT_scaled)

    def recon(self, g, params, d2, t2, norm, records=None):
        return Recon(self, g, params, d2, t2, norm, records)

    def batch(self, g, params, norm):
        return Batch(self, g, params, norm)

    def server(self, max_active=32, inflight=2, admitters=4):
        return Server(self, max_active, inflight, admitters)

    def poisson(self, mean, scale, seed, out=None, mean2=None, scale2=0.0):
        """Seeded Poisson(scale * mean + scale2 * mean2) draws on the device
        (float counts)."""
        m, _ = self._dev(mean)
        m2, _ = self._dev(mean2)
        if m2 is not None and m2.n != m.n:
            raise ValueError("mean2 must match mean")
        o = out if out is not None else self.array(m.n)
        _check(self.h, lib().gfr_poisson(self.h, m.ptr, float(scale),
                                         m2.ptr if m2 is not None else None, float(scale2),
                                         int(seed), o.ptr, m.n, None))
        if out is not None:
            return o
        self.sync()
        return o.download()


class Vlor:
    """VLOR soft-assignment state (GFIRST.cu:229-247, 373-388)."""

    def __init__(self, ctx, g, records, yT_scaled):
        self.ctx, self.g = ctx, g
        rec = np.ascontiguousarray(records, np.int32).reshape(-1)
        y, _ = ctx._dev(yT_scaled)
        h = C.c_void_p()
        _check(ctx.h, lib().gfr_vlor_create(ctx.h, C.byref(g.c()), rec.ctypes.data,
                                            len(rec) // 2, y.ptr, C.byref(h)))
        self.h = h.value

    def assign(self, proj, subset=0, nsub=1, out=None):
        p, _ = self.ctx._dev(proj)
        o = out if out is not None else self.ctx.array(self.g.ndata).zero()
        _check(self.ctx.h, lib().gfr_vlor_assign(self.ctx.h, self.h, p.ptr, o.ptr, subset, nsub,
                                                 None))
        return o if out is not None else o.download()

    def close(self):
        if self.h:
            lib().gfr_vlor_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def recon_params(mode=1, niter=1, nsub=1, psf_fw=1.5, psf_bw=1.5, jbf_r=4, jbf_sigma3d=0.75,
                 jbf_sigmai=0.025, jbf_coef=0.1, storage=_lib.STORE_F32):
    return GfrReconParams(mode, niter, nsub, psf_fw, psf_bw, jbf_r, jbf_sigma3d, jbf_sigmai,
                          jbf_coef, storage)


def footprint(g, params):
    return lib().gfr_recon_footprint(C.byref(g.c()), C.byref(params))


class Recon:
    """Device-resident reconstruction session (osem3d, GFIRST.cu:44-575)."""

    def __init__(self, ctx, g, params, d2, t2, norm, records=None):
        self.ctx, self.g, self.params = ctx, g, params
        nd = g.ndata

        def ptr(a, n, what):
            """Host (numpy) or device (DeviceArray) input of at least n floats."""
            if isinstance(a, DeviceArray):
                if a.n < n:
                    raise ValueError("%s must hold %d floats" % (what, n))
                return a, a.ptr
            a = np.ascontiguousarray(a, np.float32).reshape(-1)
            if a.size < n:
                raise ValueError("%s must hold %d floats" % (what, n))
            return a, a.ctypes.data

        d2, d2p = ptr(d2, 2 * nd, "d2")
        norm, normp = ptr(norm, nd, "norm")
        t2p = None
        if t2 is not None:
            t2, t2p = ptr(t2, 2 * nd, "t2")
        rec = None
        nrec = 0
        if records is not None:
            rec = np.ascontiguousarray(records, np.int32).reshape(-1)
            nrec = len(rec) // 2
        h = C.c_void_p()
        _check(ctx.h, lib().gfr_recon_create(ctx.h, C.byref(g.c()), C.byref(params),
                                             d2p, t2p, normp,
                                             rec.ctypes.data if rec is not None else None, nrec,
                                             C.byref(h)))
        self.h = h.value

    def step(self):
        return _check(self.ctx.h, lib().gfr_recon_step(self.h))

    def run(self):
        _check(self.ctx.h, lib().gfr_recon_run(self.h))

    def cursor(self):
        it, ty, s, n = C.c_int(), C.c_int(), C.c_int(), C.c_longlong()
        done = lib().gfr_recon_cursor(self.h, C.byref(it), C.byref(ty), C.byref(s), C.byref(n))
        return dict(iter=it.value, type=ty.value, subset=s.value, passes=n.value, done=bool(done))

    def get(self, what):
        n = self.g.nvox if what in (_lib.OUT_XD, _lib.OUT_XT) else self.g.ndata
        out = np.empty(n, np.float32)
        _check(self.ctx.h, lib().gfr_recon_get(self.h, what, out.ctypes.data))
        return out

    def loglik(self):
        v = C.c_double()
        _check(self.ctx.h, lib().gfr_recon_loglik(self.h, C.byref(v)))
        return v.value

    def profile(self, enable=True):
        _check(self.ctx.h, lib().gfr_recon_profile(self.h, 1 if enable else 0))

    def phase_times(self):
        """{phase: (total ms, passes)} since the last call (HIP events)."""
        ms = np.zeros(len(_lib.PHASES), np.float64)
        n = np.zeros(len(_lib.PHASES), np.int64)
        _check(self.ctx.h, lib().gfr_recon_phase_times(self.h, ms.ctypes.data, n.ctypes.data))
        return {k: (float(ms[i]), int(n[i])) for i, k in enumerate(_lib.PHASES)}

    @property
    def stream(self):
        return lib().gfr_recon_stream(self.h)

    def sync(self):
        """Wait for every pass enqueued so far (both chains of a joint session)."""
        _check(self.ctx.h, lib().gfr_recon_sync(self.h))

    def close(self):
        if self.h:
            lib().gfr_recon_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_params(storage=_lib.STORE_F32, max_frames=64, min_nsub=8, lanes=2, psf_fw=1.5,
                 psf_bw=1.5, group=0):
    """group: frames at the same subset of the same subset count that share one
    backprojection launch (1 none, 2 or 3; 0 = 3); results do not depend on it."""
    return GfrBatchParams(storage, max_frames, min_nsub, lanes, psf_fw, psf_bw, group)


def batch_footprint(g, params, frames=0, has_bg=False):
    """Device bytes of a batch holding `frames` frames."""
    base = lib().gfr_batch_footprint(C.byref(g.c()), C.byref(params))
    per = lib().gfr_batch_frame_footprint(C.byref(g.c()), C.byref(params), 1 if has_bg else 0)
    return base + frames * per


class Batch:
    """Frames of one geometry advanced one subset per step, each step a
    replay of one captured hipGraph (doubles OSEM, GFIRST.cu:44-575 mode 1)."""

    def __init__(self, ctx, g, params, norm):
        self.ctx, self.g, self.params = ctx, g, params
        norm = np.ascontiguousarray(norm, np.float32).reshape(-1)
        if norm.size < g.ndata:
            raise ValueError("norm must hold NDATA floats")
        h = C.c_void_p()
        _check(ctx.h, lib().gfr_batch_create(ctx.h, C.byref(g.c()), C.byref(params),
                                             norm.ctypes.data, C.byref(h)))
        self.h = h.value

    def add(self, niter, nsub, prompts, bg=None):
        """Adds a frame from NDATA prompts (numpy or DeviceArray) and an
        optional background; returns its id."""
        keep = []

        def ptr(a):
            if a is None:
                return None
            if isinstance(a, DeviceArray):
                return a.ptr
            a = np.ascontiguousarray(a, np.float32).reshape(-1)
            if a.size != self.g.ndata:
                raise ValueError("frame sinograms must hold NDATA floats")
            keep.append(a)
            return a.ctypes.data

        f = C.c_int()
        _check(self.ctx.h, lib().gfr_batch_add(self.h, niter, nsub, ptr(prompts), ptr(bg),
                                               C.byref(f)))
        return f.value

    def step(self):
        return _check(self.ctx.h, lib().gfr_batch_step(self.h))

    def run(self):
        _check(self.ctx.h, lib().gfr_batch_run(self.h))

    def cursor(self, frame):
        it, s, n = C.c_int(), C.c_int(), C.c_longlong()
        done = _check(self.ctx.h, lib().gfr_batch_cursor(self.h, frame, C.byref(it), C.byref(s),
                                                         C.byref(n)))
        return dict(iter=it.value, subset=s.value, passes=n.value, done=bool(done))

    def get(self, frame):
        out = np.empty(self.g.nvox, np.float32)
        _check(self.ctx.h, lib().gfr_batch_get(self.h, frame, out.ctypes.data))
        return out

    def loglik(self, frame):
        v = C.c_double()
        _check(self.ctx.h, lib().gfr_batch_loglik(self.h, frame, C.byref(v)))
        return v.value

    def remove(self, frame):
        _check(self.ctx.h, lib().gfr_batch_remove(self.h, frame))

    def use_graph(self, enable=True):
        _check(self.ctx.h, lib().gfr_batch_use_graph(self.h, 1 if enable else 0))

    def stats(self):
        a, b = C.c_longlong(), C.c_longlong()
        _check(self.ctx.h, lib().gfr_batch_stats(self.h, C.byref(a), C.byref(b)))
        return dict(launches=a.value, builds=b.value)

    @property
    def stream(self):
        return lib().gfr_batch_stream(self.h)

    def sync(self):
        self.ctx.sync(self.stream)

    def close(self):
        if self.h:
            lib().gfr_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Serving loop over this context (gfr_server_*): submit requests at any
    time; each is admitted as a device-resident session when a slot and HBM
    budget allow and advances pass by pass beside the others. Inputs may be
    numpy arrays or DeviceArrays; they are kept referenced here until the
    request has been admitted."""

    def __init__(self, ctx, max_active=32, inflight=2, admitters=4):
        self.ctx = ctx
        h = C.c_void_p()
        _check(ctx.h, lib().gfr_server_create(ctx.h, max_active, inflight, admitters,
                                              C.byref(h)))
        self.h = h.value
        self._inputs = {}
        self._meta = {}

    def _ptr(self, a, n, keep, dtype=np.float32):
        if a is None:
            return None
        if isinstance(a, DeviceArray):
            if a.n < n:
                raise ValueError("device input too small")
            keep.append(a)
            return a.ptr
        a = np.ascontiguousarray(a, dtype).reshape(-1)
        if a.size < n:
            raise ValueError("input too small")
        keep.append(a)
        return a.ctypes.data

    def submit(self, g, params, d2, t2, norm, records=None):
        keep = []
        nd = g.ndata
        p_d2 = self._ptr(d2, 2 * nd, keep)
        p_t2 = self._ptr(t2, 2 * nd, keep)
        p_n = self._ptr(norm, nd, keep)
        nrec = 0
        p_r = None
        if records is not None:
            rec = np.ascontiguousarray(records, np.int32).reshape(-1)
            keep.append(rec)
            p_r = rec.ctypes.data
            nrec = rec.size // 2
        rid = C.c_longlong()
        _check(self.ctx.h, lib().gfr_server_submit(self.h, C.byref(g.c()), C.byref(params), p_d2,
                                                   p_t2, p_n, p_r, nrec, C.byref(rid)))
        self._inputs[rid.value] = keep
        self._meta[rid.value] = (g, params.mode)
        self._drop_admitted()
        return rid.value

    def _drop_admitted(self):
        for rid in list(self._inputs):
            if self.info(rid)["state"] != _lib.REQ_PENDING:
                del self._inputs[rid]

    def info(self, rid):
        i = _lib.GfrRequestInfo()
        _check(self.ctx.h, lib().gfr_server_info(self.h, rid, C.byref(i)))
        return dict(state=i.state, error=i.error, t_submit=i.t_submit, t_admit=i.t_admit,
                    t_done=i.t_done, passes=i.passes, bytes=i.bytes)

    def wait(self, rid, timeout=-1.0):
        st = _check(self.ctx.h, lib().gfr_server_wait(self.h, rid, float(timeout)))
        self._drop_admitted()
        return st

    def result(self, rid):
        """(xD, xT or None) of a done request; raises GfrError for a failed one."""
        g, mode = self._meta[rid]
        xD = np.empty(g.nvox, np.float32)
        xT = np.empty(g.nvox, np.float32) if mode >= 2 else None
        _check(self.ctx.h, lib().gfr_server_result(self.h, rid, xD.ctypes.data,
                                                   xT.ctypes.data if xT is not None else None))
        return xD, xT

    def release(self, rid):
        _check(self.ctx.h, lib().gfr_server_release(self.h, rid))
        self._meta.pop(rid, None)
        self._inputs.pop(rid, None)

    def cancel(self, rid):
        """Cancel a pending or running request (it ends failed, GFR_ECANCELED)."""
        _check(self.ctx.h, lib().gfr_server_cancel(self.h, rid))

    def stats(self):
        t = _lib.GfrServerTotals()
        _check(self.ctx.h, lib().gfr_server_stats(self.h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in t._fields_}

    def close(self):
        if self.h:
            lib().gfr_server_destroy(self.h)
            self.h = None
            self._inputs.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
