"""GPU parity: every HIP kernel and the whole reconstruction against the CPU
oracle on the same seeded inputs, through the C ABI.

Tolerances: the kernels and the oracle share one rounding model (explicit
fmaf, IEEE division, host-tabled sin/cos), so operators and whole
reconstructions are compared bit for bit (rtol = 0); the relative L2 error
against the oracle (specified bound 1e-4 per iteration in FP32) is printed
beside. The only tolerances are where the arithmetic differs by design: the
FP64 sensitivity against the FP32 oracle, and the FP64 log-likelihood
(device log and a tree-ordered sum against a sequential host sum).
"""
import numpy as np
import pytest

import gfirst_amd as ga
from gfirst_amd.geometry import C0, C1, REF, SMALL, TINY, Geometry

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    d = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (d if d > 0 else 1.0)


def close(got, exp, rtol=0.0, atol=0.0, what=""):
    got = np.asarray(got)
    exp = np.asarray(exp)
    assert got.shape == exp.shape
    bad = ~np.isclose(got, exp, rtol=rtol, atol=atol)
    nexact = int(np.sum(got == exp))
    print("%s: %d/%d bit-exact, max|d|=%.3g" % (what, nexact, got.size,
                                                float(np.max(np.abs(got.astype(np.float64) - exp)))
                                                if got.size else 0.0))
    assert not bad.any(), "%s: %d mismatches, first at %d: %r vs %r" % (
        what, bad.sum(), np.argmax(bad), got.ravel()[np.argmax(bad)], exp.ravel()[np.argmax(bad)])


def rand_img(g, seed, lo=0.0, hi=2.0):
    return np.random.default_rng(seed).uniform(lo, hi, g.nvox).astype(np.float32)


@pytest.mark.parametrize("g", [TINY, C0], ids=["tiny", "c0"])
def test_blur_xy(gctx, orc, g):
    x = rand_img(g, 1)
    for psf in (1.5, 0.7):
        close(gctx.blur_xy(g, x, psf), orc.blur_xy(g, x, psf), what="blur")


@pytest.mark.parametrize("g", [TINY, SMALL, C0], ids=["tiny", "small", "c0"])
def test_forward_project(gctx, orc, g):
    x = rand_img(g, 2)
    close(gctx.forward_project(g, x), orc.forward(g, x), what="fwd S=1")
    close(gctx.forward_project(g, x, 1, 4), orc.forward(g, x, 1, 4), what="fwd s=1/4")


def test_forward_constant_image_c1(gctx):
    """Size-independent property at Inveon scale: RES+1 on every bin."""
    g = C1
    x = gctx.to_device(np.ones(g.nvox, np.float32))
    out = gctx.array(g.ndata).zero()
    gctx.forward_project(g, x, 3, 16, out=out)
    s = out.download().reshape(-1, g.nang, g.nrad)
    assert np.all(s[:, 3::16] == np.float32(g.res + 1))
    mask = np.ones(g.nang, bool)
    mask[3::16] = False
    assert np.all(s[:, mask] == 0)


def test_forward_c1_single_angle_vs_oracle(gctx, orc):
    g = C1
    x = rand_img(g, 3)
    close(gctx.forward_project(g, x, 37, 160), orc.forward(g, x, 37, 160), what="fwd c1 1 angle")


def test_forward_wide_bins_vs_oracle(gctx, orc):
    """Bins 4 voxels wide on a 96-voxel image: a rotation tile's texel box
    exceeds the LDS tile at most angles (gathers from global memory), fits at
    a few; every rotation shape gives the oracle's bits."""
    g = SMALL.with_(res=96, nzs=8, nseg=3, fov=24.0)
    x = rand_img(g, 12)
    exp = orc.forward(g, x, 1, 3)
    for rz in (4, 16, 32):
        gctx.set_tuning(g, 3, rot_z=rz)
        close(gctx.forward_project(g, x, 1, 3), exp, rtol=0, what="fwd wide bins rz=%d" % rz)


# image side not a multiple of the 16-column tiles, odd plane count, bins not
# a multiple of the 32-bin rotation tiles: partial tiles in every kernel
ODD = Geometry(40, 13, 30, 18, 3, 5, 2.0, 53.0, 40.0)


def test_odd_geometry_operators(gctx, orc):
    g = ODD
    x = rand_img(g, 13)
    close(gctx.forward_project(g, x), orc.forward(g, x), rtol=0, what="fwd odd S=1")
    close(gctx.forward_project(g, x, 2, 5), orc.forward(g, x, 2, 5), rtol=0, what="fwd odd 2/5")
    rng = np.random.default_rng(14)
    nd = g.ndata
    corr = rng.uniform(0, 3, nd).astype(np.float32)
    corr[rng.uniform(size=nd) < 0.05] = -1000.0
    valid = rng.uniform(size=nd) < 0.95
    for s, S in ((0, 1), (4, 5)):
        ne, de = orc.back(g, corr, np.where(valid, 1, -1000).astype(np.float32), s, S)
        for zb in (1, 2, 4, 5, 6, 8, 16):
            for kind, blk, ahead in ((0, 4, 1), (0, 1, 1), (0, 1, 4), (0, 4, 2)):
                if ahead > 1 and zb != 1:
                    continue
                set_back_shape(gctx, g, S, (zb, kind, 0, blk, ahead))
                n, d = gctx.back_project(g, corr, valid.astype(np.uint8), s, S)
                what = "odd zb=%d kind=%d blk=%d ahead=%d S=%d" % (zb, kind, blk, ahead, S)
                close(n, ne, rtol=0, what="numer " + what)
                close(d, de, rtol=0, what="denom " + what)
    close(gctx.blur_xy(g, x, 1.5), orc.blur_xy(g, x, 1.5), rtol=0, what="blur odd")


# every backprojector candidate of the autotuner: (slices, kind: 0 cached
# reads, 1 LDS window; register budget; waves per block)
BACK_SHAPES = [(4, 0, 0, 4), (5, 0, 0, 4), (6, 0, 0, 4), (8, 0, 0, 4), (8, 0, 6, 4),
               (16, 0, 0, 4), (4, 0, 0, 1), (6, 0, 0, 1), (8, 0, 0, 1), (4, 1, 0, 4),
               (8, 1, 0, 4), (16, 1, 0, 4), (1, 0, 0, 1), (2, 0, 0, 1), (2, 0, 0, 4),
               (1, 0, 0, 1, 2), (1, 0, 0, 1, 4), (1, 0, 0, 4, 4)]


def set_back_shape(gctx, g, S, shape):
    """shape: (slices, kind, register budget, waves per block[, views loaded ahead])"""
    zb, kind, waves, blk = shape[:4]
    ahead = shape[4] if len(shape) > 4 else 1
    gctx.set_tuning(g, S, zb, 8, back_lds=kind == 1, back_waves=waves, back_bw=blk,
                    back_ahead=ahead)


def shape_name(shape):
    return "zb=%d kind=%d waves=%d blk=%d" % shape[:4] + (
        " ahead=%d" % shape[4] if len(shape) > 4 else "")


def test_reference_geometry_operators(gctx, orc):
    """The reference's compiled geometry (GFIRST.h:4-16), the drop-in CLI's
    default: NRAD 147 (odd, a partial second 128-bin z-sum block), NANG 168,
    span 11 with 15 segments (zmin up to 72, the steepest tilts), 1293
    sinograms. Forward projection of one angle and of a 2-angle subset with
    every z-sum and rotation shape, and backprojection of a 2-angle subset
    with every backprojector shape, against the oracle bit for bit."""
    g = REF
    assert g.nsinogs == 1293
    x = rand_img(g, 70)
    close(gctx.forward_project(g, x, 101, 168), orc.forward(g, x, 101, 168),
          what="fwd ref 1 angle")
    s, S = 5, 84
    exp = orc.forward(g, x, s, S)
    for sg in (4, 6, 8):
        for rz in (4, 16, 32):
            gctx.set_tuning(g, S, 8, sg, rot_z=rz)
            close(gctx.forward_project(g, x, s, S), exp, what="fwd ref sg=%d rz=%d" % (sg, rz))
    rng = np.random.default_rng(71)
    nd = g.ndata
    corr = rng.uniform(-0.2, 3, nd).astype(np.float32)
    corr[rng.uniform(size=nd) < 0.05] = -1000.0
    r = np.arange(g.nrad)
    inside = np.abs(r - (g.nrad - 1) / 2) < 0.42 * g.nrad
    valid = np.broadcast_to(inside, (g.nsinogs, g.nang, g.nrad)).copy().reshape(-1)
    valid[rng.integers(0, nd, size=nd // 50)] = False
    ne, de = orc.back(g, corr, np.where(valid, 1, -1000).astype(np.float32), s, S)
    for shape in BACK_SHAPES:
        set_back_shape(gctx, g, S, shape)
        n, d = gctx.back_project(g, corr, valid.astype(np.uint8), s, S)
        what = "ref " + shape_name(shape)
        close(n, ne, what="numer " + what)
        close(d, de, what="denom " + what)
    gctx.set_tuning(g, S, 8, 8)


def test_calib_mask_and_ratio(gctx, orc):
    g = C0
    rng = np.random.default_rng(4)
    nd = g.ndata
    pb = rng.uniform(0, 0.03, nd).astype(np.float32)
    nm = rng.uniform(0, 0.03, nd).astype(np.float32)
    m = gctx.calib_mask(g, pb, nm)
    cal = orc.calib(g, pb, nm)
    assert np.array_equal(m.astype(bool), cal > 0)
    proj = rng.uniform(-0.02, 5, nd).astype(np.float32)
    y = (rng.poisson(2.0, nd) * g.res).astype(np.float32)
    bg = rng.uniform(-1, 1, nd).astype(np.float32)
    for s, S in ((0, 1), (2, 4)):
        got = gctx.ratio(g, proj, y, bg, s, S)
        exp = orc.ratio(g, proj, y, bg, s, S)
        close(got, exp, rtol=0, what="ratio")


def test_forward_ratio_fused(gctx, orc):
    g = C0
    x = rand_img(g, 5)
    rng = np.random.default_rng(6)
    nd = g.ndata
    y = (rng.poisson(3.0, nd) * g.res).astype(np.float32)
    bg = rng.uniform(0, 2, nd).astype(np.float32)
    got = gctx.forward_ratio(g, x, y, bg, 1, 4)
    exp = orc.ratio(g, orc.forward(g, x, 1, 4), y, bg, 1, 4)
    close(got, exp, what="forward+ratio")


@pytest.mark.parametrize("g", [SMALL, C0], ids=["small", "c0"])
def test_back_project(gctx, orc, g):
    rng = np.random.default_rng(7)
    nd = g.ndata
    corr = rng.uniform(0, 3, nd).astype(np.float32)
    corr[rng.uniform(size=nd) < 0.05] = -1000.0
    valid = rng.uniform(size=nd) < 0.93
    cal = np.where(valid, 1, -1000).astype(np.float32)
    for s, S in ((0, 1), (3, 4)):
        n, d = gctx.back_project(g, corr, valid.astype(np.uint8), s, S)
        ne, de = orc.back(g, corr, cal, s, S)
        close(n, ne, what="numer")
        close(d, de, what="denom")


@pytest.mark.parametrize("g,s,S", [(C0, 1, 4), (C1, 3, 80)], ids=["c0", "c1"])
def test_back_project_structured_mask(gctx, orc, g, s, S):
    """Masks as a scanner makes them (valid inside the FOV, whole radial runs
    masked at the edges) plus scattered dead bins: the oracle's bits. Ratios
    include the -1000 gap value and negative counts."""
    rng = np.random.default_rng(17)
    nd = g.ndata
    corr = rng.uniform(-0.2, 3, nd).astype(np.float32)
    corr[rng.uniform(size=nd) < 0.05] = -1000.0
    r = np.arange(g.nrad)
    inside = np.abs(r - (g.nrad - 1) / 2) < 0.42 * g.nrad
    valid = np.broadcast_to(inside, (g.nsinogs, g.nang, g.nrad)).copy()
    dead = rng.integers(0, nd, size=max(8, nd // 20000))
    valid.reshape(-1)[dead] = False
    valid = valid.reshape(-1)
    ne, de = orc.back(g, corr, np.where(valid, 1, -1000).astype(np.float32), s, S)
    for shape in ((6, 0, 0, 4), (8, 1, 0, 4), (8, 0, 0, 1), (4, 0, 0, 4), (1, 0, 0, 1, 4)):
        set_back_shape(gctx, g, S, shape)
        n, d = gctx.back_project(g, corr, valid.astype(np.uint8), s, S)
        close(n, ne, rtol=0, what="numer " + shape_name(shape))
        close(d, de, rtol=0, what="denom " + shape_name(shape))
    gctx.set_tuning(g, S, 8, 8)


def test_back_project_c1_subset_vs_oracle(gctx, orc):
    g = C1
    rng = np.random.default_rng(8)
    nd = g.ndata
    corr = rng.uniform(0, 3, nd).astype(np.float32)
    valid = rng.uniform(size=nd) < 0.97
    cal = np.where(valid, 1, -1000).astype(np.float32)
    n, d = gctx.back_project(g, corr, valid.astype(np.uint8), 5, 80)
    ne, de = orc.back(g, corr, cal, 5, 80)
    close(n, ne, what="numer c1")
    close(d, de, what="denom c1")
    # the LDS-window backprojector (windows at full tilt) and the one-wave
    # blocks of the cached-read kernel on the same views
    for shape in ((8, 1, 0, 4), (6, 0, 0, 1), (1, 0, 0, 4, 4)):
        set_back_shape(gctx, g, 80, shape)
        n2, d2 = gctx.back_project(g, corr, valid.astype(np.uint8), 5, 80)
        close(n2, ne, rtol=0, what="numer c1 " + shape_name(shape))
        close(d2, de, rtol=0, what="denom c1 " + shape_name(shape))
    gctx.set_tuning(g, 80, 8, 8)


# back kind: 0 cached reads, 1 per-block LDS row window; bw: register budget
# (waves per SIMD); blk: waves per direct backprojector block
@pytest.mark.parametrize("zb,sg,lds,bw,rz,blk", [
    (4, 4, 0, 0, 32, 4), (5, 8, 0, 0, 16, 4), (6, 4, 0, 0, 4, 4), (8, 6, 0, 0, 4, 4),
    (16, 8, 0, 0, 16, 4), (8, 8, 0, 6, 4, 4), (4, 8, 1, 0, 16, 4), (6, 8, 1, 0, 32, 4),
    (8, 4, 1, 0, 32, 4), (16, 6, 1, 0, 4, 4), (4, 4, 0, 0, 16, 1), (6, 8, 0, 0, 4, 1),
    (8, 6, 0, 0, 32, 1), (5, 8, 0, 0, 16, 1), (16, 4, 0, 0, 16, 1), (1, 8, 0, 0, 16, 1),
    (2, 4, 0, 0, 4, 1), (2, 6, 0, 0, 32, 4)])
def test_tile_shapes_do_not_change_results(gctx, orc, zb, sg, lds, bw, rz, blk):
    """Every autotuner candidate gives the oracle's bits (forward and back)."""
    g = C0
    gctx.set_tuning(g, 4, zb, sg, lds, bw, rz, back_bw=blk)
    t = gctx.tuning(g, 4)
    assert (t["back_zb"], t["zsum_sg"], t["back_lds"], t["back_waves"], t["rot_z"],
            t["back_bw"], t["measured"]) == (zb, sg, lds, bw, rz, blk, True)
    x = rand_img(g, 30)
    close(gctx.forward_project(g, x, 2, 4), orc.forward(g, x, 2, 4), what="fwd sg=%d" % sg)
    rng = np.random.default_rng(31)
    nd = g.ndata
    corr = rng.uniform(0, 3, nd).astype(np.float32)
    corr[rng.uniform(size=nd) < 0.05] = -1000.0
    valid = rng.uniform(size=nd) < 0.93
    n, d = gctx.back_project(g, corr, valid.astype(np.uint8), 2, 4)
    ne, de = orc.back(g, corr, np.where(valid, 1, -1000).astype(np.float32), 2, 4)
    close(n, ne, rtol=0, what="numer zb=%d" % zb)
    close(d, de, rtol=0, what="denom zb=%d" % zb)
    gctx.set_tuning(g, 4, 8, 8)


@pytest.mark.parametrize("g", [SMALL, C0], ids=["small", "c0"])
def test_sensitivity_fp64(gctx, orc, g):
    """Calibration half of the backprojection in FP64: the oracle's DENOM for
    a ratio of 1 up to FP32-vs-FP64 accumulation (most voxels count whole
    views exactly; segment-edge rows add fractions)."""
    nd = g.ndata
    ones = np.ones(nd, np.float32)
    for s, S in ((0, 1), (2, 4)):
        sens = gctx.sensitivity(g, np.ones(nd, np.uint8), s, S)
        _, de = orc.back(g, ones, ones, s, S)
        assert sens.dtype == np.float64
        close(sens, de, rtol=1e-6, atol=1e-6, what="all-valid sensitivity S=%d" % S)
        whole = (de == np.round(de)) & (sens == np.round(sens))
        assert whole.mean() > 0.5 and np.array_equal(sens[whole], de[whole].astype(np.float64))
    rng = np.random.default_rng(33)
    valid = rng.uniform(size=nd) < 0.9
    sens = gctx.sensitivity(g, valid.astype(np.uint8), 1, 4)
    _, de = orc.back(g, ones, np.where(valid, 1, -1000).astype(np.float32), 1, 4)
    _, de64 = orc.back(g, ones.astype(np.float64), np.where(valid, 1.0, -1000.0), 1, 4, fp64=True)
    close(sens, de, rtol=1e-5, atol=1e-6, what="sensitivity vs FP32 oracle")
    # the FP64 oracle also interpolates in FP64: where a -1000 gap bin has a
    # small weight, the x1001 amplification of weight rounding shows (up to
    # ~1e-4 relative with 10% of the bins invalid)
    close(sens, de64, rtol=2e-3, atol=1e-6, what="sensitivity vs FP64 oracle")


def test_osem_run_one_call(gctx, orc):
    g = SMALL
    d2, t2, norm, _ = make_study(g, 41, 2)
    p = ga.recon_params(mode=2, niter=1, nsub=3)
    xD, xT = gctx.osem_run(g, p, d2, t2, norm)
    exD, exT, _, _ = orc.osem(g, orc.params(mode=2, niter=1, nsub=3), d2, t2, norm)
    assert np.array_equal(xD, exD) and np.array_equal(xT, exT)


@pytest.mark.parametrize("g", [SMALL, C0], ids=["small", "c0"])
def test_update(gctx, orc, g):
    rng = np.random.default_rng(9)
    num = rng.uniform(0, 50, g.nvox).astype(np.float32)
    den = rng.uniform(0, 50, g.nvox).astype(np.float32)
    den[::13] = 0.0
    x = rng.uniform(0.1, 2, g.nvox).astype(np.float32)
    got = gctx.update(g, x, num, den, 1.5)
    exp = orc.update(g, x, num, den, 1.5)
    close(got, exp, what="update")


def test_jbf_blend_max(gctx, orc):
    g = SMALL
    rng = np.random.default_rng(10)
    inp = rng.uniform(0, 2, g.nvox).astype(np.float32)
    guide = rng.uniform(0, 1, g.nvox).astype(np.float32)
    mx = gctx.image_max(g, guide)
    assert mx == orc.vmax(guide)
    got = gctx.jbf3d(g, inp, guide, mx)
    exp = orc.jbf(g, inp, guide, mx)
    close(got, exp, what="jbf")
    b = gctx.blend(g, inp, exp, 0.1)
    close(b, orc.blend(g, inp, exp, 0.1), rtol=0, what="blend")


def near_midpoint_t(n, kI, seed=0):
    """Guide differences t whose intensity weight exp(-t*t/kI) lies within
    2^-42 (and beyond 2^-50) relative of a midpoint between two floats:
    the weights whose float rounding a less accurate exp could flip."""
    rng = np.random.default_rng(seed)
    t = rng.uniform(0.001, 0.45, n).astype(np.float32)
    E = np.exp(-(t * t).astype(np.float64) / kI)
    H = E.astype(np.float32)
    Hd = H.astype(np.float64)
    up = np.nextafter(H, np.float32(np.inf)).astype(np.float64)
    dn = np.nextafter(H, np.float32(0)).astype(np.float64)
    d = np.minimum(np.abs(E - (Hd + up) / 2), np.abs(E - (Hd + dn) / 2)) / E
    return t[(d > 2.0 ** -50) & (d < 2.0 ** -42) & (H > 0)]


@pytest.mark.parametrize("scale", [1.0, 2.0 ** 24], ids=["mref1", "mref2^24"])
def test_jbf_weights_near_rounding_midpoints(gctx, orc, scale):
    """Radius-4 JBF on guides whose weights sit next to float rounding
    midpoints (and a few denormal guide values), against the oracle bit for
    bit. With max_ref = 2^24 every weight takes the reference expression."""
    g = SMALL
    t = near_midpoint_t(8_000_000, (2.0 * 0.025) ** 2)
    assert len(t) >= 30
    guide = np.zeros((g.nzs, g.res, g.res), np.float32)
    slots = [(z, y, x) for z in (5, 10, 15) for y in range(5, 28, 5) for x in range(5, 28, 5)]
    for (z, y, x), v in zip(slots, t):
        guide[z, y, x] = v
    guide[12, 6, 7] = np.float32(1e-42)   # denormal
    guide[8, 22, 21] = np.float32(3e-39)
    guide = (guide * np.float32(scale)).reshape(-1)
    rng = np.random.default_rng(12)
    inp = rng.uniform(0.5, 2, g.nvox).astype(np.float32)
    mref = float(np.float32(scale))
    got = gctx.jbf3d(g, inp, guide, mref)
    exp = orc.jbf(g, inp, guide, mref)
    close(got, exp, rtol=0, atol=0, what="jbf near midpoints")
    assert np.array_equal(got, exp)


def test_vlor_assign(gctx, orc):
    g = SMALL
    nd = g.ndata
    rng = np.random.default_rng(11)
    rec = rng.integers(-3, nd + 3, size=(4000, 2)).astype(np.int32)
    rec[:50, 1] = rec[:50, 0]  # same bin twice
    yT = (rng.poisson(1.0, nd) * g.res).astype(np.float32)
    proj = rng.uniform(0, 2, nd).astype(np.float32)
    proj[::5] = 0.0
    v = gctx.vlor(g, rec, yT)
    bt = orc.vlor_blankt(g, yT, orc.vlor_precount(g, rec))
    for s, S in ((0, 1), (1, 3)):
        got = v.assign(proj, s, S)
        exp = orc.vlor_assign(g, rec, bt, proj, s, S)
        close(got, exp, what="vlor")
    v.close()


def make_study(g, seed, mode, counts=2e5):
    from gfirst_amd import synth
    rng = np.random.default_rng(seed)
    nd = g.ndata
    lam = rng.uniform(0.2, 3.0, nd) * (counts / nd)
    d2 = synth.as_file_pair(rng.poisson(lam).astype(np.float32),
                            rng.uniform(0, 0.3, nd).astype(np.float32))
    t2 = None
    if mode >= 2:
        t2 = synth.as_file_pair(rng.poisson(lam * 0.1).astype(np.float32))
    norm = rng.uniform(0.8, 1.2, nd).astype(np.float32)
    norm[rng.uniform(size=nd) < 0.02] = 0.0
    rec = None
    if mode == 3:
        rec = rng.integers(1, nd, size=(3000, 2)).astype(np.int32)
        rec = np.concatenate([rec, rec[-1:]])
    return d2, t2, norm, rec


# c1-m1-S16: one whole iteration (16 subset passes) at the Inveon size, about
# 15 s for the oracle on 16 host threads
# The joint cases with 2-3 iterations cross iteration boundaries, where the
# triples passes of one iteration run beside the next iteration's doubles.
@pytest.mark.parametrize("g,mode,S,niter,counts", [(C0, 1, 4, 2, 2e5), (SMALL, 2, 3, 2, 2e5),
                                                    (SMALL, 3, 1, 2, 2e5), (SMALL, 3, 4, 1, 2e5),
                                                    (ODD, 2, 5, 2, 1e5), (C1, 1, 16, 1, 5e7),
                                                    (C0, 3, 4, 3, 3e5), (C0, 2, 8, 2, 3e5)],
                         ids=["c0-m1-S4", "small-m2-S3", "small-m3-S1", "small-m3-S4",
                              "odd-m2-S5", "c1-m1-S16", "c0-m3-S4-3it", "c0-m2-S8-2it"])
def test_recon_vs_oracle(gctx, orc, g, mode, S, niter, counts):
    d2, t2, norm, rec = make_study(g, 12 + mode, mode, counts)
    p = ga.recon_params(mode=mode, niter=niter, nsub=S)
    r = gctx.recon(g, p, d2, t2, norm, rec)
    pb = r.get(ga.OUT_PROJ_BLANK)
    r.run()
    xD = r.get(ga.OUT_XD)
    op = orc.params(mode=mode, niter=niter, nsub=S)
    exD, exT, epb, _ = orc.osem(g, op, d2, t2, norm, rec)
    close(pb, epb, what="proj_blank")
    e = rel_l2(xD, exD)
    print("xD rel L2 vs oracle FP32: %.3g (bit-exact %d/%d)" % (e, np.sum(xD == exD), xD.size))
    assert e <= 1e-4
    close(xD, exD, what="xD")
    if mode >= 2:
        xT = r.get(ga.OUT_XT)
        eT = rel_l2(xT, exT)
        print("xT rel L2 vs oracle FP32: %.3g" % eT)
        assert eT <= 1e-4
        close(xT, exT, what="xT")
    r.close()


def test_recon_loglik_and_cursor(gctx, orc):
    g = TINY
    d2, _, norm, _ = make_study(g, 20, 1, counts=5e4)
    r = gctx.recon(g, ga.recon_params(mode=1, niter=2, nsub=2), d2, None, norm)
    assert r.cursor()["passes"] == 0
    assert r.step() == 0
    ll1 = r.loglik()
    r.step()
    r.step()
    ll3 = r.loglik()
    c = r.cursor()
    assert c["passes"] == 3 and c["iter"] == 2 and c["subset"] == 1
    assert np.isfinite(ll1) and np.isfinite(ll3)
    assert r.step() == 1 and r.cursor()["done"]
    r.close()


def oracle_pass_loglik(orc, g, x, d2, cal, s, S):
    """The oracle's FP64 log-likelihood of pass (s of S) from estimate x."""
    nd = g.ndata
    res = np.float32(g.res)
    proj = orc.forward(g, orc.blur_xy(g, x, 1.5), s, S)
    return proj, orc.loglik(g, proj, res * d2[:nd], res * d2[nd:], cal, s, S)


@pytest.mark.parametrize("g,S,counts,npass", [(C0, 4, 2e5, 3), (C1, 16, 5e7, 1)],
                         ids=["c0", "c1"])
def test_recon_loglik_vs_oracle(gctx, orc, g, S, counts, npass):
    """The session's FP64 log-likelihood of each pass against the oracle's
    (orc_loglik: sum over the subset's bins with valid calibration and
    den >= 0.01 of (y/RES) ln(den/RES) - den/RES), the estimate following the
    oracle pass by pass. Relative to the sum of |terms|: 1e-12 (device log
    and a tree-ordered sum against a sequential host sum)."""
    d2, _, norm, _ = make_study(g, 50, 1, counts)
    r = gctx.recon(g, ga.recon_params(mode=1, niter=1, nsub=S), d2, None, norm)
    pb = r.get(ga.OUT_PROJ_BLANK)
    cal = orc.calib(g, pb, norm)
    nd = g.ndata
    res = np.float32(g.res)
    x = orc.cylinder(g, 1.0)
    for s in range(npass):
        r.step()
        got = r.loglik()
        proj, (exp, la) = oracle_pass_loglik(orc, g, x, d2, cal, s, S)
        print("pass %d: loglik %.17g oracle %.17g (|terms| %.3g)" % (s, got, exp, la))
        assert abs(got - exp) <= 1e-12 * la
        if s + 1 < npass:
            num, den = orc.back(g, orc.ratio(g, proj, res * d2[:nd], res * d2[nd:], s, S),
                                cal, s, S)
            x = orc.update(g, x, num, den, 1.5)
    r.close()


def test_errors_are_status_codes(gctx):
    g = TINY
    x = gctx.array(g.nvox).zero()
    with pytest.raises(ga.GfrError) as e:
        gctx.forward_project(g, x, 3, 2)
    assert e.value.code == ga.GFR_EINVAL
    with pytest.raises((ga.GfrError, ValueError)):
        gctx.forward_project(g.with_(ntbins=3), x)
    small = ga.Context(0, hbm_budget=1 << 20)
    d2, _, norm, _ = make_study(C0, 1, 1)
    with pytest.raises(ga.GfrError) as e:
        small.recon(C0, ga.recon_params(), d2, None, norm)
    assert e.value.code == ga.GFR_ENOMEM
    small.close()


@pytest.mark.parametrize("g,S,events,span", [(C1, 16, 2000000, False),
                                              (REF, 12, 1000000, True)],
                         ids=["inveon", "reference-geometry-2-iterations"])
def test_joint_first_triples_passes_vs_oracle(gctx, orc, g, S, events, span):
    """The joint reconstruction (mode 3) with VLOR records binned from triple
    events, pass by pass against the oracle following the reference's loop
    (GFIRST.cu:324-527), bit for bit. At the Inveon size (860k records from
    2e6 events): the first S + 2 passes (S doubles subsets, then triples
    subsets 0 and 1 with VLOR assignment, projection cache, bilateral filter
    and blend). At the reference's compiled geometry (span 11, NRAD 147, 15
    segments; GFIRST.h:4-16): the first 2S + 2 passes of two iterations, so
    iteration 1's triples passes run on the session's second stream beside
    iteration 2's first doubles passes, each still seeing its sequential
    inputs. About 30-40 s of oracle work each on 16 host threads."""
    from gfirst_amd import listmode, synth
    nd = g.ndata

    def expected(img):
        b = gctx.blur_xy(g, img, 1.5)
        return gctx.forward_project(g, b).astype(np.float64)

    chA = synth.phantom_mouse(g, seed=3, channel=0)
    chB = synth.phantom_mouse(g, seed=3, channel=1)
    yDe, yTe = expected(chA), expected(chB)
    rng = np.random.default_rng(31)
    d2 = synth.as_file_pair(rng.poisson(yDe * (1e7 / yDe.sum())).astype(np.float32))
    t2 = synth.as_file_pair(rng.poisson(yTe * (2e5 / yTe.sum())).astype(np.float32))
    norm = synth.norm_uniform(nd, seed=5)
    rec = listmode.bin_hits(g, listmode.triple_events(g, chB, events, seed=8), 80.5)
    rec = rec[(rec[:, 0] > 0) & (rec[:, 1] > 0)]
    assert len(rec) > events // 10
    nT = S if span else 2               # triples passes compared
    nD = S + 2 if span else S           # doubles passes compared
    r = gctx.recon(g, ga.recon_params(mode=3, niter=2 if span else 1, nsub=S), d2, t2, norm,
                   rec)
    for _ in range(nD + nT):
        r.step()
    c = r.cursor()
    assert (c["type"], c["subset"], c["passes"]) == ((0, 2) if span else (1, 2)) + (nD + nT,)
    xD, xT = r.get(ga.OUT_XD), r.get(ga.OUT_XT)
    r.close()

    res = np.float32(g.res)
    cal = orc.calib(g, orc.forward(g, orc.cylinder(g, 1.0), 0, 1), norm)
    yD, bD = res * d2[:nd], res * d2[nd:]
    yT, bT = res * t2[:nd], res * t2[nd:]
    blankt = orc.vlor_blankt(g, yT, orc.vlor_precount(g, rec))
    eD, eT = orc.cylinder(g, 1.0), orc.cylinder(g, 1.0)
    ptc = orc.forward(g, orc.blur_xy(g, eT, 1.5), 0, 1)
    dv = np.zeros(nd, np.float32)
    def doubles(eD, s):
        proj = orc.forward(g, orc.blur_xy(g, eD, 1.5), s, S)
        num, den = orc.back(g, orc.ratio(g, proj, yD, bD, s, S), cal, s, S)
        return orc.update(g, eD, num, den, 1.5)

    for s in range(S):
        eD = doubles(eD, s)
    guide = eD  # iteration 1's doubles image guides iteration 1's triples passes
    for s in range(nT):
        proj = orc.forward(g, orc.blur_xy(g, eT, 1.5), s, S)
        cur = ptc.reshape(g.nsinogs, g.nang, g.nrad)
        cur[:, s::S, :] = proj.reshape(g.nsinogs, g.nang, g.nrad)[:, s::S, :]
        dv = orc.vlor_assign(g, rec, blankt, ptc, s, S, dv)
        num, den = orc.back(g, orc.ratio(g, proj, dv * res, bT, s, S), cal, s, S)
        eT = orc.update(g, eT, num, den, 1.5)
        eT = orc.blend(g, eT, orc.jbf(g, eT, guide, orc.vmax(guide)))
    for s in range(nD - S):  # iteration 2's first doubles passes
        eD = doubles(eD, s)
    close(xD, eD, rtol=0, what="joint xD after %d doubles passes" % nD)
    close(xT, eT, rtol=0, what="joint xT after %d triples passes" % nT)
    assert np.array_equal(xD, eD) and np.array_equal(xT, eT)


def test_joint_sessions_share_geometry_sinograms(gctx, orc):
    """Sessions of one geometry share PROJ_BLANK and, per forward PSF, the
    initial projection cache of the joint mode: sessions with two PSFs, the
    first PSF again after the second, each against the oracle."""
    g = ODD
    d2, t2, norm, rec = make_study(g, 41, 3)
    for psf in (1.5, 1.1, 1.5):
        p = ga.recon_params(mode=3, niter=1, nsub=3, psf_fw=psf)
        r = gctx.recon(g, p, d2, t2, norm, rec)
        pb = r.get(ga.OUT_PROJ_BLANK)
        r.run()
        xD, xT = r.get(ga.OUT_XD), r.get(ga.OUT_XT)
        r.close()
        eD, eT, epb, _ = orc.osem(g, orc.params(mode=3, niter=1, nsub=3, psf_fw=psf), d2, t2,
                                  norm, rec)
        assert np.array_equal(pb, epb)
        close(xD, eD, rtol=0, what="xD psf %.1f" % psf)
        close(xT, eT, rtol=0, what="xT psf %.1f" % psf)


def test_tuning_rejects_invalid_shapes(gctx):
    """gfr_tuning_set validates every field and leaves the shapes in force
    when it refuses."""
    g = SMALL
    gctx.set_tuning(g, 3, 6, 8, back_bw=1)
    before = gctx.tuning(g, 3)
    bad = [dict(back_zb=7), dict(zsum_sg=5), dict(rot_z=8), dict(back_lds=2),
           dict(back_waves=6, back_zb=6), dict(back_waves=6, back_lds=True),
           dict(back_bw=2), dict(back_bw=1, back_lds=True), dict(back_bw=1, back_waves=6),
           dict(back_zb=3), dict(back_zb=2, back_lds=True), dict(back_zb=1, back_waves=6),
           dict(back_zb=2, back_ahead=2), dict(back_zb=1, back_ahead=3),
           dict(back_zb=1, back_ahead=4, back_lds=True)]
    for kw in bad:
        with pytest.raises(ga.GfrError):
            gctx.set_tuning(g, 3, **kw)
        assert gctx.tuning(g, 3) == before
    gctx.set_tuning(g, 3, 8, 8)
    assert gctx.tuning(g, 3)["back_bw"] == 4
