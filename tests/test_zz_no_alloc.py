"""gfr_back_project_noalloc (the backprojection for callers that cannot
allocate device memory) through the C ABI: its numerator and denominator
equal gfr_back_project's bit for bit on two calls in a row, the second on a
smaller geometry than the first (the fixed scratch array keeps the first
call's larger contents), and it refuses a subset whose packed rows exceed the
64 MiB array and a geometry the context holds no tables for."""
import ctypes as C

import numpy as np
import pytest

import gfirst_amd as ga
from gfirst_amd import _lib
from gfirst_amd.geometry import C0, C1, TINY

pytestmark = pytest.mark.gpu


def inputs(g, seed):
    rng = np.random.default_rng(seed)
    nd = g.ndata
    corr = rng.uniform(-0.2, 3.0, nd).astype(np.float32)
    corr[rng.uniform(size=nd) < 0.05] = -1000.0
    valid = (rng.uniform(size=nd) < 0.95).astype(np.uint8)
    return corr, valid


def noalloc(ctx, g, dc, dm, n, d, s, S):
    return _lib.lib().gfr_back_project_noalloc(ctx.h, C.byref(g.c()), dc.ptr, dm.ptr, n.ptr, d.ptr,
                                               s, S, None)


def test_noalloc_backprojection_equals_allocating_entry_point(gctx):
    ctx = gctx
    dev = {}
    for g, s, S in ((C1, 3, 16), (C0, 1, 4)):  # the second call is the smaller one
        print("noalloc backprojection: geometry %dx%dx%d, subset %d of %d" % (
            g.res, g.res, g.nzs, s, S), flush=True)
        corr, valid = inputs(g, 7 + S)
        n_ref, d_ref = ctx.back_project(g, corr, valid, s, S)  # creates the geometry's tables
        dc, dm = ctx.to_device(corr), ctx.to_device(valid, np.uint8)
        n, d = ctx.array(g.nvox), ctx.array(g.nvox)
        n.upload(np.full(g.nvox, 7.0, np.float32))
        d.upload(np.full(g.nvox, -3.0, np.float32))
        rc = noalloc(ctx, g, dc, dm, n, d, s, S)
        assert rc == 0, _lib.lib().gfr_last_error(ctx.h).decode()
        ctx.sync()
        assert np.array_equal(n.download(), n_ref)
        assert np.array_equal(d.download(), d_ref)
        dev[g] = (dc, dm, n, d)
    # packed rows of a 4-subset split at the Inveon size: 184 MB > 64 MiB
    dc, dm, n, d = dev[C1]
    assert noalloc(ctx, C1, dc, dm, n, d, 0, 4) == ga.GFR_EINVAL
    assert b"64 MiB" in _lib.lib().gfr_last_error(ctx.h)
    # a context that has not seen the geometry yet
    fresh = ga.Context(0)
    g = TINY.with_(nang=6)
    dc2, dm2 = fresh.array(g.ndata), fresh.array(g.ndata, np.uint8)
    n2, d2 = fresh.array(g.nvox), fresh.array(g.nvox)
    assert noalloc(fresh, g, dc2, dm2, n2, d2, 0, 2) == ga.GFR_ESTATE
    fresh.close()
