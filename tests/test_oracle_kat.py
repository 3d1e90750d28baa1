"""Known-answer tests that pin the CPU oracle (no GPU).

Each expectation comes from mathematics the reference fixes (closed forms,
invariants, brute force on tiny inputs), not from the oracle's own formulas,
so a dropped term, wrong sign, wrong index or transposed operand fails here.
"""
import math

import numpy as np
import pytest

from gfirst_amd.geometry import C0, C1, REF, SMALL, TINY, Geometry

PI_F = float(np.float32(3.1415926))


def seg_list(g):
    """(zmin, sgn) per segment in the reference's order 0,-1,+1,... from first
    principles: segment k has ring-difference index idz = +-ceil(k/2)."""
    out = []
    for k in range(g.nseg):
        idz = ((k + 1) // 2) * (1 if k % 2 == 0 else -1)
        zm = 0 if idz == 0 else (abs(idz) - 1) * g.span + g.span // 2 + 1
        out.append((zm, int(np.sign(idz))))
    return out


# ---------------------------------------------------------------- geometry
@pytest.mark.parametrize("g,ns", [(REF, 1293), (C1, 4319), (C0, 192)])
def test_nsinogs_matches_reference_counts(orc, g, ns):
    # REF: NSINOGS hand-set in GFIRST.h:9; C1: the Inveon span-3 count (4319)
    assert orc.nsinogs(g) == ns
    assert orc.ndata(g) == ns * g.nang * g.nrad


def test_seg_tables_first_principles(orc):
    n, zmin, sgn, nza = orc.seg_tables(C1)
    exp = seg_list(C1)
    assert [int(z) for z in zmin] == [e[0] for e in exp]
    assert [int(s) for s in sgn] == [e[1] for e in exp]
    assert zmin[1] == 2 and zmin[2] == 2 and zmin[3] == 5
    assert list(np.diff(nza)) == [C1.nzs - 2 * z for z in zmin]
    # sinogram index is a bijection onto [0, NDATA)
    seen = np.zeros(n, np.int32)
    for k in range(C1.nseg):
        for p in range(zmin[k], C1.nzs - zmin[k]):
            seen[nza[k] + p - zmin[k]] += 1
    assert np.all(seen == 1)


def test_angle_tables(orc):
    s, c = orc.angle_tables(C0)
    for it in (0, 1, 17, 32, 63):
        ang = np.float32(np.float32(-it) / np.float32(C0.nang) * np.float32(PI_F) + np.float32(PI_F))
        assert s[it] == np.float32(math.sin(float(ang)))
        assert c[it] == np.float32(math.cos(float(ang)))
    # theta(0) = PI: sin ~ PI - PI_f, cos = -1
    assert abs(s[0] - (math.pi - PI_F)) < 1e-9 and c[0] == -1.0


def test_psf_taps_closed_form(orc):
    for psf in (1.0, 1.5, 3.0):
        g = orc.psf_taps(psf)
        # the argument is formed in float (1.662f, PSF+0.1f), the rest exactly
        a = (np.arange(-4, 5, dtype=np.float32) * np.float32(1.662) /
             (np.float32(psf) + np.float32(0.1))).astype(np.float64)
        w = np.exp(-(a * a))
        np.testing.assert_allclose(g, w / w.sum(), rtol=2e-6)
        assert np.all(g == g[::-1])
    g64 = orc.psf_taps(1.5, fp64=True)
    np.testing.assert_allclose(g64.sum(), 1.0, rtol=1e-14)


def test_gz_taps_closed_form(orc):
    g = orc.gz_taps()
    w = np.exp(-(np.arange(-2, 3) ** 2.0))
    np.testing.assert_allclose(g, w / w.sum(), rtol=2e-7)


def test_cylinder(orc):
    img = orc.cylinder(TINY, 2.5).reshape(TINY.nzs, TINY.res, TINY.res)
    h = TINY.res // 2
    yy, xx = np.mgrid[0:TINY.res, 0:TINY.res]
    inside = (xx - h) ** 2 + (yy - h) ** 2 <= h * h
    assert np.all(img[3][inside] == 2.5) and np.all(img[3][~inside] == 0)
    assert img[0, 0, h] == 2.5 and img[0, 0, 0] == 0  # (x=h, y=0) is on the circle


# ---------------------------------------------------------------- sampler
def tex_textbook(v, c):
    """CUDA texture linear filter, clamp addressing, from the programming
    guide's weighted-sum formula, in float64."""
    D, H, W = v.shape
    idx, wts = [], []
    for coord, n in zip(c, (W, H, D)):
        t = coord - 0.5
        i = math.floor(t)
        a = t - i
        idx.append((min(max(i, 0), n - 1), min(max(i + 1, 0), n - 1)))
        wts.append((1 - a, a))
    out = 0.0
    for kz in (0, 1):
        for ky in (0, 1):
            for kx in (0, 1):
                out += (wts[0][kx] * wts[1][ky] * wts[2][kz] *
                        v[idx[2][kz], idx[1][ky], idx[0][kx]])
    return out


def test_tex3_brute_force(orc):
    rng = np.random.default_rng(1)
    v = rng.uniform(-1, 2, size=(3, 4, 5)).astype(np.float32)
    pts = rng.uniform(-1.5, 6.5, size=(400, 3)).astype(np.float32)
    pts[:20] = np.round(pts[:20] * 2) / 2  # texel centres and edges
    got = orc.tex3(v, 5, 4, 3, pts)
    exp = np.array([tex_textbook(v.astype(np.float64), p.astype(np.float64)) for p in pts])
    np.testing.assert_allclose(got, exp, rtol=0, atol=2e-6)


# ---------------------------------------------------------------- forward
def sino_view(g, sino, zmin, nza, k, it):
    """(plane, r) image of one view from the sinogram (NaN outside planes)."""
    s = sino.reshape(-1, g.nang, g.nrad)
    out = np.full((g.nzs, g.nrad), np.nan)
    for p in range(zmin[k], g.nzs - zmin[k]):
        out[p] = s[nza[k] + p - zmin[k], it]
    return out


def test_forward_constant_image(orc):
    for g in (TINY, C0):
        img = np.ones(g.nvox, np.float32)
        sino = orc.forward(g, img)
        # every LOR has RES+1 samples of a clamped constant (GFIRST_GPU_Kernels.cu:90-98)
        assert np.all(sino == np.float32(g.res + 1))


def test_forward_subset_writes_only_its_angles(orc):
    g = TINY
    img = np.ones(g.nvox, np.float32)
    sino = orc.forward(g, img, subset=1, nsub=3).reshape(-1, g.nang, g.nrad)
    for it in range(g.nang):
        assert np.all(sino[:, it] == (g.res + 1 if it % 3 == 1 else 0))


def test_forward_theta0_closed_form(orc):
    """theta = PI: LORs run along y through column RES-1-r (fx = 1 and NRAD =
    RES at C0); the half-texel offset gives sum_y v + (v_0 + v_last)/2."""
    g = C0
    rng = np.random.default_rng(2)
    v = rng.uniform(0, 1, size=(g.nzs, g.res, g.res)).astype(np.float32)
    sino = orc.forward(g, v.reshape(-1)).reshape(-1, g.nang, g.nrad)
    direct = sino[:g.nzs, 0, :]  # segment 0 occupies the first NZS sinograms
    cols = v[:, :, ::-1]  # column RES-1-r
    exp = cols.sum(axis=1) + 0.5 * (cols[:, 0, :] + cols[:, -1, :])
    np.testing.assert_allclose(direct, exp, rtol=1e-4)


def test_forward_theta_half_closed_form(orc):
    """theta = PI/2: LORs run along x through row r."""
    g = C0
    rng = np.random.default_rng(3)
    v = rng.uniform(0, 1, size=(g.nzs, g.res, g.res)).astype(np.float32)
    sino = orc.forward(g, v.reshape(-1)).reshape(-1, g.nang, g.nrad)
    direct = sino[:g.nzs, g.nang // 2, :]
    exp = v.sum(axis=2) + 0.5 * (v[:, :, 0] + v[:, :, -1])
    np.testing.assert_allclose(direct, exp, rtol=1e-4)


def test_forward_tilt_z_linear_image(orc):
    """Image linear in z: each sample is a + b*clamp(p + l*dz, 0, NZS-1) with
    dz = (FOV/RES)*sgn*zmin/(DIAM/2) (GFIRST_GPU_Kernels.cu:83, 88, 96)."""
    g = SMALL
    a, b = 1.0, 0.25
    v = np.broadcast_to((a + b * np.arange(g.nzs))[:, None, None], (g.nzs, g.res, g.res))
    sino = orc.forward(g, np.ascontiguousarray(v, np.float32).reshape(-1))
    n, zmin, sgn, nza = orc.seg_tables(g)
    lam = np.arange(-(g.res // 2), g.res // 2 + 1)
    for k in range(g.nseg):
        dz = (g.fov / g.res) * sgn[k] * zmin[k] / (g.diam_det / 2)
        for it in (0, 5, 13):
            view = sino_view(g, sino, zmin, nza, k, it)
            for p in range(zmin[k], g.nzs - zmin[k]):
                exp = np.sum(a + b * np.clip(p + lam * dz, 0, g.nzs - 1))
                np.testing.assert_allclose(view[p], exp, rtol=2e-6)


def test_forward_fp64_matches_fp32(orc):
    g = SMALL
    rng = np.random.default_rng(4)
    v = rng.uniform(0, 1, size=g.nvox)
    s32 = orc.forward(g, v.astype(np.float32))
    s64 = orc.forward(g, v, fp64=True)
    np.testing.assert_allclose(s32, s64, rtol=2e-5, atol=1e-4)


# ---------------------------------------------------------------- ratio/mask
def test_ratio_closed_form(orc):
    g = TINY
    rng = np.random.default_rng(5)
    nd = orc.ndata(g)
    proj = rng.uniform(-0.02, 3, nd).astype(np.float32)
    y = rng.integers(0, 5, nd).astype(np.float32) * g.res
    bg = rng.uniform(-1, 1, nd).astype(np.float32)
    corr = orc.ratio(g, proj, y, bg)
    b = np.maximum(bg, 0)
    den = proj + b
    exp = np.where(den >= np.float32(0.01), y / np.where(den == 0, 1, den), -1000).astype(np.float32)
    np.testing.assert_array_equal(corr, exp)
    corr0 = orc.ratio(g, proj, y, None)
    exp0 = np.where(proj >= np.float32(0.01), y / np.where(proj == 0, 1, proj), -1000)
    np.testing.assert_array_equal(corr0, exp0.astype(np.float32))


def test_calib_mask(orc):
    g = TINY
    nd = orc.ndata(g)
    pb = np.linspace(0, 0.02, nd).astype(np.float32)
    nm = np.linspace(0.02, 0, nd).astype(np.float32)
    cal = orc.calib(g, pb, nm)
    exp = np.where((pb.astype(np.float64) > 0.01) & (nm.astype(np.float64) > 0.01), 1, -1000)
    np.testing.assert_array_equal(cal, exp.astype(np.float32))


# ---------------------------------------------------------------- back
def test_back_corr_equals_calib_gives_numer_equals_denom(orc):
    g = SMALL
    nd = orc.ndata(g)
    rng = np.random.default_rng(6)
    cal = np.where(rng.uniform(size=nd) < 0.9, 1, -1000).astype(np.float32)
    num, den = orc.back(g, cal.copy(), cal)
    np.testing.assert_array_equal(num, den)
    assert den.max() <= g.nseg * g.nang + 1e-3


def test_back_linear_in_plane_closed_form(orc):
    """corr(p, r) = a + b p, all bins valid: an interior voxel receives
    a + b (iz + yr fz) per view (bilinear interpolation is exact on linear
    data), with yr = -xv sin + yv cos and fz = (FOV/RES) sgn zmin/(DIAM/2)."""
    g = SMALL
    a, b = 1.0, 0.25
    n, zmin, sgn, nza = orc.seg_tables(g)
    nd = orc.ndata(g)
    corr = np.zeros(nd, np.float32).reshape(-1, g.nang, g.nrad)
    for k in range(g.nseg):
        for p in range(zmin[k], g.nzs - zmin[k]):
            corr[nza[k] + p - zmin[k]] = a + b * p
    cal = np.ones(nd, np.float32)
    num, den = orc.back(g, corr.reshape(-1), cal)
    num = num.reshape(g.nzs, g.res, g.res)
    den = den.reshape(g.nzs, g.res, g.res)
    s, c = orc.angle_tables(g)
    s = s.astype(np.float64)
    c = c.astype(np.float64)
    resm = (g.res - 1) / 2
    iz, iy, ix = np.meshgrid(np.arange(g.nzs), np.arange(g.res), np.arange(g.res), indexing="ij")
    xv, yv = ix - resm, iy - resm
    exp = np.zeros_like(num, dtype=np.float64)
    ok = np.ones(num.shape, bool)
    for k in range(g.nseg):
        fz = (g.fov / g.res) * sgn[k] * zmin[k] / (g.diam_det / 2)
        for it in range(g.nang):
            yr = -xv * s[it] + yv * c[it]
            t = iz + yr * fz  # sample position in plane units (w - 0.5)
            ok &= (t >= zmin[k] + 1) & (t <= g.nzs - zmin[k] - 2)
            exp += a + b * t
    assert ok.sum() > 100
    np.testing.assert_allclose(num[ok], exp[ok], rtol=1e-5)
    assert np.all(den[ok] == g.nseg * g.nang)


def test_back_gap_rule(orc):
    """A view whose planes are all invalid contributes nothing (cal < 0)."""
    g = TINY
    nd = orc.ndata(g)
    corr = np.full(nd, 2.0, np.float32)
    cal = np.full(nd, -1000.0, np.float32)
    num, den = orc.back(g, corr, cal)
    assert np.all(num == 0) and np.all(den == 0)


# ---------------------------------------------------------------- blur/update
def test_blur_impulse_closed_form(orc):
    g = SMALL
    gk = orc.psf_taps(1.5).astype(np.float64)
    v = np.zeros((g.nzs, g.res, g.res), np.float32)
    imp = [(3, 10, 12, 1.0), (3, 20, 1, 2.0), (5, 0, 16, 1.5)]
    for z, y, x, a in imp:
        v[z, y, x] = a
    out = orc.blur_xy(g, v.reshape(-1), 1.5).reshape(v.shape)
    exp = np.zeros_like(out, dtype=np.float64)
    for z in range(g.nzs):
        for y in range(g.res):
            for x in range(g.res):
                if 4 <= x <= g.res - 5:
                    acc = 0.0
                    for (zi, yi, xi, a) in imp:
                        if zi == z and 0 <= xi - x + 4 <= 8 and 0 <= yi - y + 4 <= 8:
                            acc += a * gk[xi - x + 4] * gk[yi - y + 4]
                    exp[z, y, x] = acc
                else:
                    exp[z, y, x] = v[z, y, x]
    np.testing.assert_allclose(out, exp, rtol=1e-6, atol=1e-8)


def test_update_closed_form(orc):
    """numer = 1 + A delta, denom = 1, x = 1: x' = Gz(Gxy(R)) with separable
    taps; slices 0 and NZS-1 copy slices 1 and NZS-2 (GFIRST.cu:491-496)."""
    g = SMALL
    gk = orc.psf_taps(1.5).astype(np.float64)
    gz = orc.gz_taps().astype(np.float64)
    A = 3.0
    z0, y0, x0 = 7, 15, 14
    num = np.ones((g.nzs, g.res, g.res), np.float32)
    num[z0, y0, x0] += A
    den = np.ones(g.nvox, np.float32)
    den[5] = 0.0005  # DENOM < 0.001 -> ratio 1
    num_flat = num.reshape(-1).copy()
    num_flat[5] = 77.0
    x = np.ones(g.nvox, np.float32)
    out = orc.update(g, x, num_flat, den, 1.5).reshape(num.shape)
    sg = gk.sum() ** 2
    for z in (4, 5, 6, 7, 8, 9, 10):
        for (y, xx) in ((15, 14), (16, 13), (13, 16), (8, 8)):
            dz = z - z0
            f = sg
            if abs(dz) <= 2 and abs(y - y0) <= 4 and abs(xx - x0) <= 4:
                f += A * gk[y0 - y + 4] * gk[x0 - xx + 4] * gz[2 - dz]
            elif abs(y - y0) <= 4 and abs(xx - x0) <= 4:
                f += 0.0
            np.testing.assert_allclose(out[z, y, xx], f * gz.sum(), rtol=2e-6)
    np.testing.assert_array_equal(out[0], out[1])
    np.testing.assert_array_equal(out[-1], out[-2])


def test_update_edge_slice_rules(orc):
    """Slice 1 and NZS-2 use 0.333 taps; x=y=0 of slice 2 keeps slice 1's rule
    (the reference's index shift, GFIRST_CPU_Kernels.cpp:178-182)."""
    g = SMALL
    rng = np.random.default_rng(7)
    zprof = rng.uniform(0.5, 2.0, g.nzs)
    num = np.broadcast_to(zprof[:, None, None], (g.nzs, g.res, g.res)).astype(np.float32)
    out = orc.update(g, np.ones(g.nvox, np.float32), num.reshape(-1).copy(),
                     np.ones(g.nvox, np.float32), 1.5).reshape(num.shape)
    gk = orc.psf_taps(1.5).astype(np.float64)
    gz = orc.gz_taps().astype(np.float64)
    sxy = gk.sum() ** 2  # interior columns; border columns keep the raw value (x1)
    ymid, xmid = g.res // 2, g.res // 2
    z1 = 0.333 * (zprof[0] + zprof[1] + zprof[2])
    np.testing.assert_allclose(out[1, ymid, xmid], sxy * z1, rtol=2e-6)
    zl = 0.333 * (zprof[-3] + zprof[-2] + zprof[-1])
    np.testing.assert_allclose(out[-2, ymid, xmid], sxy * zl, rtol=2e-6)
    z5 = sum(gz[j] * zprof[5 + j - 2] for j in range(5))
    np.testing.assert_allclose(out[5, ymid, xmid], sxy * z5, rtol=2e-6)
    # (x=0, y=0) voxel of slice 2: uses the slice-1 rule (3 taps of 0.333)
    z2s = 0.333 * (zprof[1] + zprof[2] + zprof[3])
    np.testing.assert_allclose(out[2, 0, 0], z2s, rtol=2e-6)
    # any other border voxel of slice 2 uses the 5-tap rule
    z2 = sum(gz[j] * zprof[2 + j - 2] for j in range(5))
    np.testing.assert_allclose(out[2, 0, 1], z2, rtol=2e-6)


# ---------------------------------------------------------------- JBF
def jbf_brute(inp, guide, mref, r, s3d, sI):
    D, H, W = inp.shape
    out = np.zeros_like(inp, dtype=np.float64)
    for z in range(r, D):
        for y in range(r, H):
            for x in range(r, W):
                n = s = 0.0
                for zz in range(z - r, min(z + r, D - 1) + 1):
                    for yy in range(y - r, min(y + r, H - 1) + 1):
                        for xx in range(x - r, min(x + r, W - 1) + 1):
                            di = ((guide[zz, yy, xx] - guide[z, y, x]) / mref) ** 2
                            ds = (xx - x) ** 2 + (yy - y) ** 2 + (zz - z) ** 2
                            w = math.exp(-di / (2 * sI) ** 2) * math.exp(-ds / (2 * s3d) ** 2)
                            n += w
                            s += w * inp[zz, yy, xx]
                out[z, y, x] = s / n if n > 1e-8 else 0.0
    return out


def test_jbf_brute_force(orc):
    g = Geometry(10, 8, 10, 4, 3, 1, 2.0, 26.5, 10.0)
    rng = np.random.default_rng(8)
    inp = rng.uniform(0, 2, (g.nzs, g.res, g.res)).astype(np.float32)
    guide = rng.uniform(0, 1, (g.nzs, g.res, g.res)).astype(np.float32)
    mref = float(guide.max())
    got = orc.jbf(g, inp.reshape(-1), guide.reshape(-1), mref, r=2, sigma3d=0.75, sigmai=0.25)
    exp = jbf_brute(inp.astype(np.float64), guide.astype(np.float64), mref, 2, 0.75, 0.25)
    np.testing.assert_allclose(got.reshape(exp.shape), exp, rtol=2e-5, atol=1e-6)


def test_jbf_constant_and_low_border(orc):
    g = Geometry(12, 10, 12, 4, 3, 1, 2.0, 26.5, 12.0)
    rng = np.random.default_rng(9)
    guide = rng.uniform(0, 1, g.nvox).astype(np.float32)
    out = orc.jbf(g, np.full(g.nvox, 3.0, np.float32), guide, float(guide.max()), r=4)
    out = out.reshape(g.nzs, g.res, g.res)
    assert np.all(out[:4] == 0) and np.all(out[:, :4] == 0) and np.all(out[:, :, :4] == 0)
    np.testing.assert_allclose(out[4:, 4:, 4:], 3.0, rtol=1e-6)


def test_blend(orc):
    g = TINY
    x = np.linspace(0, 5, g.nvox).astype(np.float32)
    f = np.linspace(5, 0, g.nvox).astype(np.float32)
    got = orc.blend(g, x, f, 0.1)
    c = np.float32(0.1)
    exp = ((1.0 - float(c)) * x.astype(np.float64) + (c * f).astype(np.float64)).astype(np.float32)
    np.testing.assert_array_equal(got, exp)


# ---------------------------------------------------------------- VLOR
def test_vlor_records_follow_feof_loop(orc):
    w = np.array([5, 6, 7, 8, 9, 10], np.int32)
    r = orc.vlor_records(w.tobytes())
    np.testing.assert_array_equal(r, [[5, 6], [7, 8], [9, 10], [9, 10]])
    r = orc.vlor_records(np.array([5, 6, 7], np.int32).tobytes())
    np.testing.assert_array_equal(r, [[5, 6], [7, 6]])
    assert len(orc.vlor_records(b"")) == 0


def test_vlor_mass_conservation(orc):
    g = TINY
    nd = orc.ndata(g)
    rng = np.random.default_rng(10)
    rec = rng.integers(-2, nd + 3, size=(500, 2)).astype(np.int32)
    rec[:5] = [[1, 5], [0, 7], [nd + 1, 3], [4, 4], [9, nd]]
    yT = (rng.integers(0, 4, nd) * g.res).astype(np.float32)
    proj = rng.uniform(0, 2, nd).astype(np.float32)
    proj[::7] = 0.0
    v0 = orc.vlor_precount(g, rec)
    i1 = rec[:, 0].astype(np.int64) - 1
    i2 = rec[:, 1].astype(np.int64) - 1
    valid = (i1 > 0) & (i2 > 0) & (i1 < nd) & (i2 < nd)
    cnt = np.bincount(np.concatenate([i1[valid], i2[valid]]), minlength=nd)
    np.testing.assert_array_equal(v0, cnt.astype(np.float32))
    bt = orc.vlor_blankt(g, yT, v0)
    exp_bt = np.where(cnt > 0, yT / (g.res * np.maximum(cnt, 1)), 0)
    np.testing.assert_allclose(bt, exp_bt, rtol=1e-6)
    dv = orc.vlor_assign(g, rec, bt, proj)
    s = proj[i1[valid]].astype(np.float64) + proj[i2[valid]]
    val = bt[i1[valid]].astype(np.float64) + bt[i2[valid]]
    assert abs(dv.sum() - val[s > 0.01].sum()) <= 1e-4 * max(1.0, val.sum())
    # the pair [4, 4] puts its whole weight on bin 3 when s > 0.01


def test_vlor_subset_keeps_other_bins(orc):
    g = TINY
    nd = orc.ndata(g)
    rng = np.random.default_rng(11)
    rec = rng.integers(1, nd, size=(300, 2)).astype(np.int32)
    bt = rng.uniform(0, 1, nd).astype(np.float32)
    proj = rng.uniform(0.5, 2, nd).astype(np.float32)
    full = orc.vlor_assign(g, rec, bt, proj)
    part = np.full(nd, 7.0, np.float32)
    orc.vlor_assign(g, rec, bt, proj, subset=1, nsub=2, dv=part)
    it = (np.arange(nd) // g.nrad) % g.nang
    np.testing.assert_array_equal(part[it % 2 == 1], full[it % 2 == 1])
    assert np.all(part[it % 2 == 0] == 7.0)



def vlor_split_f32(rec, bt, proj, nd, swap=False):
    """The reference's per-record statements (GFIRST.cu:378-387) emulated one
    float32 operation at a time: valor = BLANKT[i1] + BLANKT[i2]; s = s1 + s2;
    s_inv = (float)(1.0/s) if s > 0.01f else 0; datos_v[i1] += s1*s_inv*valor,
    then datos_v[i2] += s2*s_inv*valor. swap=True gives each bin the other
    bin's share (a plausible mistake the oracle must not make)."""
    f = np.float32
    dv = np.zeros(nd, np.float32)
    for a, b in rec:
        i1, i2 = int(a) - 1, int(b) - 1
        if i1 <= 0 or i2 <= 0 or i1 >= nd or i2 >= nd:
            continue
        valor = f(bt[i1] + bt[i2])
        s1, s2 = proj[i1], proj[i2]
        s = f(s1 + s2)
        sinv = f(1.0 / float(s)) if s > f(0.01) else f(0.0)
        w1, w2 = (s2, s1) if swap else (s1, s2)
        dv[i1] = f(dv[i1] + f(f(w1 * sinv) * valor))
        dv[i2] = f(dv[i2] + f(f(w2 * sinv) * valor))
    return dv


def test_vlor_assign_split_brute_force(orc):
    """Hand-built events with unequal projections: every bin's value against
    the reference's statements (float32 emulation, exact) and a float64 loop
    (1e-6). Covers the s1:s2 split, an i1 == i2 record, s <= 0.01 (and s
    exactly 0.01f, which the strict test excludes), index 0, negative and
    >= NDATA indices, and bins shared by several events in record order."""
    g = TINY
    nd = orc.ndata(g)
    f = np.float32
    proj = np.zeros(nd, np.float32)
    bt = np.zeros(nd, np.float32)
    A, B, Cb, D, E, F, H, K = 110, 125, 140, 177, 200, 300, 400, 501
    proj[[A, B, Cb, D, E, F, H, K]] = [3.0, 1.0, 0.004, 0.005, 0.5, f(0.01), 0.0, 2.5]
    bt[[A, B, Cb, D, E, F, H, K]] = [0.75, 2.0, 5.0, 7.0, 1.25, 3.0, 4.0, 0.1]
    rec = [[A + 1, B + 1],       # s = 4: A takes 3/4, B takes 1/4 of 2.75
           [Cb + 1, D + 1],      # s = 0.009 <= 0.01: nothing
           [E + 1, E + 1],       # same bin twice: both halves land on E
           [F + 1, H + 1],       # s == 0.01f exactly: nothing (strict >)
           [1, A + 1],           # i1 = 0: skipped (bin 0 is excluded)
           [0, B + 1],           # i1 = -1: skipped
           [nd + 1, A + 1],      # i1 = NDATA: skipped
           [K + 1, A + 1],       # K shares with A: 2.5/5.5 and 3/5.5 of 0.85
           [B + 1, K + 1]]       # then B and K again, accumulated in record order
    rng = np.random.default_rng(23)
    extra = rng.integers(2, 60, size=(40, 2))
    proj[2:60] = rng.uniform(0.001, 3.0, 58).astype(np.float32)
    bt[2:60] = rng.uniform(0.0, 2.0, 58).astype(np.float32)
    rec = np.concatenate([np.array(rec), extra]).astype(np.int32)
    got = orc.vlor_assign(g, rec, bt, proj)
    exp = vlor_split_f32(rec, bt, proj, nd)
    np.testing.assert_array_equal(got, exp)
    # closed values of the hand-built events
    assert got[Cb] == 0 and got[D] == 0 and got[F] == 0 and got[H] == 0 and got[0] == 0
    assert got[E] == f(2.5)
    # float64 restatement of the same loop
    d64 = np.zeros(nd)
    for a, b in rec:
        i1, i2 = int(a) - 1, int(b) - 1
        if i1 <= 0 or i2 <= 0 or i1 >= nd or i2 >= nd:
            continue
        s = float(proj[i1]) + float(proj[i2])
        if s > 0.01:
            v = float(bt[i1]) + float(bt[i2])
            d64[i1] += proj[i1] / s * v
            d64[i2] += proj[i2] / s * v
    np.testing.assert_allclose(got, d64, rtol=1e-6, atol=1e-7)
    assert abs(got[A] - (0.75 * 2.75 + 3.0 / 5.5 * 0.85)) < 1e-6
    # the test discriminates: swapped shares and a non-strict 0.01 test differ
    assert not np.array_equal(vlor_split_f32(rec, bt, proj, nd, swap=True), got)
    assert got[F] == 0 and float(proj[F]) + float(proj[H]) == float(f(0.01))


# ---------------------------------------------------------------- output
def test_output_scale_and_circular_mask(orc):
    """Output image (GFIRST.cu:530-549): x * scale_D (float 3.70e-4*162.0)
    with every voxel zeroed where (x-RES/2)^2 + (y-RES/2)^2 > (RES/2-1)^2;
    the ring at exactly radius RES/2-1 is kept, the next lattice points go."""
    for g in (C0, TINY):
        rng = np.random.default_rng(29)
        x = rng.uniform(0.1, 9.0, g.nvox).astype(np.float32)
        scale = np.float32(3.70e-4 * 162.0)
        out = orc.output(g, x, scale).reshape(g.nzs, g.res, g.res)
        h = g.res // 2
        rr = h - 1
        yy, xx = np.mgrid[0:g.res, 0:g.res]
        keep = (xx - h) ** 2 + (yy - h) ** 2 <= rr * rr
        exp = np.where(keep[None], (x * scale).reshape(g.nzs, g.res, g.res), np.float32(0))
        np.testing.assert_array_equal(out, exp)
        # on the circle: kept; one step further out: zero
        for (px, py) in ((h + rr, h), (h - rr, h), (h, h + rr), (h, h - rr)):
            assert out[1, py, px] == np.float32(x.reshape(out.shape)[1, py, px] * scale)
        for (px, py) in ((h + rr, h + 1), (h - rr - 1, h), (h, h - rr - 1), (0, 0)):
            assert out[1, py, px] == 0
    assert abs(float(scale) - 0.05994) < 1e-8


# ---------------------------------------------------------------- log-likelihood
def test_loglik_closed_form_at_the_data(orc):
    """Model mean equal to the counts (den = RES*y): LL = sum y ln y - y over
    the subset's bins with y > 0 (zero bins have den < 0.01 and add 0)."""
    g = SMALL
    nd = orc.ndata(g)
    rng = np.random.default_rng(41)
    yc = rng.integers(0, 9, nd).astype(np.float32)
    res = np.float32(g.res)
    y = res * yc
    cal = np.ones(nd, np.float32)
    for s, S in ((0, 1), (2, 3)):
        ll, la = orc.loglik(g, y, y, None, cal, s, S)
        it = (np.arange(nd) // g.nrad) % g.nang
        sel = (it % S == s) & (yc > 0)
        v = yc[sel].astype(np.float64)
        exp = math.fsum(v * np.log(v) - v)
        assert abs(ll - exp) <= 1e-13 * la, (ll, exp)


def test_loglik_maximum_exclusions_and_background(orc):
    """The Poisson LL is largest where the mean equals the counts (any other
    mean lowers every term); masked bins, bins with den < 0.01 and bins of
    other subsets do not count; negative background counts as 0 and a
    positive one enters the mean (y=2, p=RES, bg=RES: 2 ln 2 - 2)."""
    g = TINY
    nd = orc.ndata(g)
    rng = np.random.default_rng(43)
    res = np.float32(g.res)
    yc = rng.integers(1, 6, nd).astype(np.float32)
    y = res * yc
    cal = np.ones(nd, np.float32)
    ll0, _ = orc.loglik(g, y, y, None, cal)
    for k in range(5):
        p = (y * rng.uniform(0.7, 1.3, nd)).astype(np.float32)
        assert orc.loglik(g, p, y, None, cal)[0] < ll0
    # exclusions
    y2 = y.copy()
    cal2 = cal.copy()
    cal2[5] = -1000.0
    y2[5] = 1e6
    p2 = y.copy()
    p2[9] = np.float32(0.005)
    y2[9] = 1e6
    a, _ = orc.loglik(g, p2, y2, None, cal2)
    base, _ = orc.loglik(g, p2, y, None, cal2)
    assert a == base
    it = (np.arange(nd) // g.nrad) % g.nang
    y3 = y.copy()
    y3[np.flatnonzero(it % 2 == 0)[:7]] = 1e6
    assert orc.loglik(g, y, y3, None, cal, 1, 2)[0] == orc.loglik(g, y, y, None, cal, 1, 2)[0]
    # background
    assert orc.loglik(g, y, y, np.full(nd, -5.0, np.float32), cal)[0] == ll0
    one = np.zeros(nd, np.float32)
    one[3] = res
    cal1 = np.zeros(nd, np.float32) - 1000.0
    cal1[3] = 1.0
    ll, _ = orc.loglik(g, one, 2 * one, one, cal1)
    assert abs(ll - (2 * math.log(2.0) - 2.0)) < 1e-15


# ---------------------------------------------------------------- files
def test_decode_sinogram_int_float_and_truncation(orc):
    nd = 10
    ints = np.arange(2 * nd, dtype=np.int32)
    out, kind = orc.decode_sinogram(ints.tobytes(), nd)
    assert kind == 0 and np.array_equal(out, ints.astype(np.float32))
    fl = np.linspace(0.5, 3, nd).astype(np.float32)  # prompts only, float file
    out, kind = orc.decode_sinogram(fl.tobytes(), nd)
    assert kind == 1 and np.array_equal(out[:nd], fl) and np.all(out[nd:] == 0)
    big = np.arange(3 * nd, dtype=np.int32)
    out, kind = orc.decode_sinogram(big.tobytes(), nd)
    assert kind == 0 and np.array_equal(out, big[:2 * nd].astype(np.float32))
    out, kind = orc.decode_sinogram(np.array([2000000], np.int32).tobytes(), nd)
    assert kind == 1  # max of the int view > 1e6 -> reinterpreted as float


# ---------------------------------------------------------------- driver
def test_osem_driver_composition(orc):
    """nsub = 1, one iteration of mode 1 equals the reference's sequence of
    operators (GFIRST.cu:150-189, 324-502)."""
    g = TINY
    nd, nv = orc.ndata(g), orc.nvox(g)
    rng = np.random.default_rng(12)
    y = rng.poisson(3.0, nd).astype(np.float32)
    d2 = np.concatenate([y, np.zeros(nd, np.float32)])
    norm = rng.uniform(0.8, 1.2, nd).astype(np.float32)
    norm[::11] = 0.0
    xD, _, pb, _ = orc.osem(g, orc.params(mode=1, niter=1, nsub=1), d2, None, norm)
    cyl = orc.cylinder(g, 1.0)
    pb2 = orc.forward(g, cyl)
    np.testing.assert_array_equal(pb, pb2)
    cal = orc.calib(g, pb2, norm)
    xb = orc.blur_xy(g, cyl, 1.5)
    proj = orc.forward(g, xb)
    corr = orc.ratio(g, proj, np.float32(g.res) * y, None)
    num, den = orc.back(g, corr, cal)
    x = orc.update(g, cyl, num, den, 1.5)
    np.testing.assert_array_equal(xD, x)


def test_osem_fp32_vs_fp64_drift(orc):
    g = SMALL
    nd = orc.ndata(g)
    rng = np.random.default_rng(13)
    y = rng.poisson(4.0, nd).astype(np.float32)
    d2 = np.concatenate([y, np.zeros(nd, np.float32)])
    norm = np.ones(nd, np.float32)
    p = orc.params(mode=1, niter=2, nsub=3)
    x32, _, _, _ = orc.osem(g, p, d2, None, norm)
    x64, _, _, _ = orc.osem(g, p, d2, None, norm, fp64=True)
    rel = np.linalg.norm(x32 - x64) / np.linalg.norm(x64)
    assert rel < 1e-4, rel
