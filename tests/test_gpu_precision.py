"""Reduced-precision sinogram storage, pinned two ways.

Prompts and background are stored in FP16 or BF16 and widened to FP32 on
load; every sum along a LOR or over views stays FP32, the log-likelihood
FP64. (1) Exactness: a reconstruction from FP16/BF16 storage equals, bit for
bit, the FP32 oracle run on the inputs rounded to that format (round to
nearest even), and measured integer counts up to 2048 (FP16) / 256 (BF16)
are stored exactly, so there reduced storage changes nothing. (2) Error
against the FP64 oracle, reported per format: the FP32 algorithm itself sits
about 1e-4 relative L2 per iteration from FP64 on the doubles image (the
reference's -1000 sentinel in the ratio interpolation and its 0.01/0.001
thresholds amplify rounding near the mask edges); rounding a non-integer
background to FP16 adds almost nothing, to BF16 a few percent of that.
Bound on the doubles image: 1e-3.
"""
import numpy as np
import pytest

import gfirst_amd as ga
from gfirst_amd.geometry import C0, SMALL

pytestmark = pytest.mark.gpu

FORMATS = {"f32": ga.STORE_F32, "f16": ga.STORE_F16, "bf16": ga.STORE_BF16}


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def study(g, seed, counts, bg_level, mode=1):
    rng = np.random.default_rng(seed)
    nd = g.ndata
    lam = rng.uniform(0.2, 3.0, nd) * (counts / nd)
    d2 = np.zeros(2 * nd, np.float32)
    d2[:nd] = rng.poisson(lam)
    if bg_level > 0:
        d2[nd:] = rng.uniform(0, bg_level, nd)
    t2 = None
    if mode >= 2:
        t2 = np.zeros(2 * nd, np.float32)
        t2[:nd] = rng.poisson(lam * 0.1)
        if bg_level > 0:
            t2[nd:] = rng.uniform(0, 0.1 * bg_level, nd)
    norm = rng.uniform(0.8, 1.2, nd).astype(np.float32)
    return d2, t2, norm


def run(gctx, g, d2, t2, norm, mode, S, niter, storage):
    r = gctx.recon(g, ga.recon_params(mode=mode, niter=niter, nsub=S, storage=storage), d2, t2,
                   norm)
    r.run()
    xD = r.get(ga.OUT_XD)
    xT = r.get(ga.OUT_XT) if mode >= 2 else None
    r.close()
    return xD, xT


def bf16_round(a):
    b = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    r = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


ROUND = {"f32": lambda a: a, "f16": lambda a: a.astype(np.float16).astype(np.float32),
         "bf16": bf16_round}


@pytest.mark.parametrize("g,mode,S,niter,counts", [(C0, 1, 4, 2, 1e6), (SMALL, 2, 3, 2, 2e5)],
                         ids=["c0-m1", "small-m2"])
def test_reduced_storage_equals_oracle_on_rounded_inputs(gctx, orc, g, mode, S, niter, counts):
    d2, t2, norm = study(g, 37, counts, bg_level=0.7, mode=mode)
    for name in ("f16", "bf16"):
        rd2 = ROUND[name](d2)
        rt2 = ROUND[name](t2) if t2 is not None else None
        assert not np.array_equal(rd2, d2)  # the background really is rounded
        exD, exT, _, _ = orc.osem(g, orc.params(mode=mode, niter=niter, nsub=S), rd2, rt2, norm)
        xD, xT = run(gctx, g, d2, t2, norm, mode, S, niter, FORMATS[name])
        print("%s: xD %d/%d bit-exact vs oracle on rounded inputs" % (
            name, int(np.sum(xD == exD)), xD.size))
        assert np.array_equal(xD, exD)
        if mode >= 2:
            assert np.array_equal(xT, exT)


@pytest.mark.parametrize("g,mode,S,niter,counts", [(C0, 1, 4, 2, 1e6), (SMALL, 2, 3, 2, 2e5)],
                         ids=["c0-m1", "small-m2"])
def test_storage_error_vs_fp64_oracle(gctx, orc, g, mode, S, niter, counts):
    d2, t2, norm = study(g, 31, counts, bg_level=0.7, mode=mode)
    x64, t64, _, _ = orc.osem(g, orc.params(mode=mode, niter=niter, nsub=S), d2, t2, norm, None,
                              fp64=True)
    err, errT = {}, {}
    for name, st in FORMATS.items():
        xD, xT = run(gctx, g, d2, t2, norm, mode, S, niter, st)
        err[name] = rel_l2(xD, x64)
        msg = "%s storage, %d iterations: xD rel L2 vs FP64 oracle %.3g" % (name, niter,
                                                                             err[name])
        if xT is not None:
            errT[name] = rel_l2(xT, t64)
            msg += ", xT %.3g" % errT[name]
        print(msg)
        assert err[name] <= 1e-3
    assert err["f32"] <= 1.5e-4 * niter
    # rounding the background adds at most a little on top of FP32 arithmetic
    assert err["f16"] <= err["f32"] + 1e-4 and err["bf16"] <= err["f32"] + 3e-4
    if errT:
        # the triples image of the joint mode (sparse counts, joint bilateral
        # filter) sits further from FP64 in FP32 itself; storage adds nothing
        assert errT["f32"] <= 5e-3
        assert max(errT.values()) <= errT["f32"] + 3e-4


@pytest.mark.parametrize("name,maxcount", [("f16", 2048), ("bf16", 256)])
def test_integer_counts_in_exact_range_are_lossless(gctx, name, maxcount):
    g = SMALL
    rng = np.random.default_rng(41)
    nd = g.ndata
    d2 = np.zeros(2 * nd, np.float32)
    d2[:nd] = rng.integers(0, maxcount + 1, nd)
    d2[rng.integers(0, nd, 8)] = maxcount  # the largest exact count
    norm = rng.uniform(0.8, 1.2, nd).astype(np.float32)
    x32, _ = run(gctx, g, d2, None, norm, 1, 3, 2, ga.STORE_F32)
    xr, _ = run(gctx, g, d2, None, norm, 1, 3, 2, FORMATS[name])
    assert np.array_equal(x32, xr)
    # one past the exact range is rounded (2049 -> 2048 in FP16, 257 -> 256 in BF16)
    d2[:nd] = maxcount + 1
    x32, _ = run(gctx, g, d2, None, norm, 1, 3, 1, ga.STORE_F32)
    xr, _ = run(gctx, g, d2, None, norm, 1, 3, 1, FORMATS[name])
    assert not np.array_equal(x32, xr)
