"""Serving loop (gfr_server_*): requests of different geometries, modes,
storage formats and count levels submitted together come back bit-identical
to the same request reconstructed alone; admission respects slots and the
HBM budget; a bad or oversized request fails with its own status code while
the others complete."""
import time

import numpy as np
import pytest

import gfirst_amd as ga
from gfirst_amd.geometry import C0, C1, SMALL, TINY

pytestmark = pytest.mark.gpu


def request(g, seed, mode, counts):
    rng = np.random.default_rng(seed)
    nd = g.ndata
    lam = rng.uniform(0.2, 3.0, nd) * (counts / nd)
    d2 = np.zeros(2 * nd, np.float32)
    d2[:nd] = rng.poisson(lam)
    if seed % 2:
        d2[nd:] = rng.uniform(0, 0.3, nd)
    t2 = rec = None
    if mode >= 2:
        t2 = np.zeros(2 * nd, np.float32)
        t2[:nd] = rng.poisson(lam * 0.1)
    if mode == 3:
        rec = rng.integers(1, nd, size=(2000, 2)).astype(np.int32)
    norm = rng.uniform(0.8, 1.2, nd).astype(np.float32)
    norm[rng.uniform(size=nd) < 0.02] = 0.0
    return d2, t2, norm, rec


def alone(ctx, g, p, d2, t2, norm, rec):
    r = ctx.recon(g, p, d2, t2, norm, rec)
    r.run()
    xD = r.get(ga.OUT_XD)
    xT = r.get(ga.OUT_XT) if p.mode >= 2 else None
    r.close()
    return xD, xT


MIX = [  # geometry, mode, nsub, niter, storage, counts
    (C0, 1, 4, 2, ga.STORE_F16, 1e6),
    (SMALL, 2, 3, 2, ga.STORE_F32, 2e5),
    (TINY, 1, 2, 3, ga.STORE_BF16, 5e4),
    (SMALL, 3, 4, 1, ga.STORE_F16, 2e5),
    (C0, 1, 8, 1, ga.STORE_F32, 3e5),
    (SMALL, 1, 1, 2, ga.STORE_F16, 1e5),
    (C0, 2, 4, 1, ga.STORE_F16, 5e5),
    (TINY, 3, 2, 2, ga.STORE_F32, 2e4),
]


def test_results_alone_equal_results_in_the_mix(gctx):
    srv = gctx.server(max_active=4, inflight=2)
    ids, reqs = [], []
    for k, (g, mode, S, ni, st, counts) in enumerate(MIX):
        d2, t2, norm, rec = request(g, 500 + k, mode, counts)
        p = ga.recon_params(mode=mode, niter=ni, nsub=S, storage=st)
        reqs.append((g, p, d2, t2, norm, rec))
        if k % 3 == 0:  # some requests hand over device-resident inputs
            d2 = gctx.to_device(d2)
            norm = gctx.to_device(norm)
        ids.append(srv.submit(g, p, d2, t2, norm, rec))
    for rid in ids:
        assert srv.wait(rid, 120.0) == ga.REQ_DONE, srv.info(rid)
    infos = [srv.info(rid) for rid in ids]
    assert all(i["t_submit"] <= i["t_admit"] <= i["t_done"] for i in infos)
    st = srv.stats()
    assert st["done"] == len(MIX) and st["failed"] == 0 and st["pending"] == 0
    got = [srv.result(rid) for rid in ids]
    srv.close()
    for k, ((g, p, d2, t2, norm, rec), (xD, xT)) in enumerate(zip(reqs, got)):
        eD, eT = alone(gctx, g, p, d2, t2, norm, rec)
        n = int(np.sum(xD == eD))
        print("request %d (mode %d, S=%d): xD %d/%d bit-exact" % (k, p.mode, p.nsub, n, xD.size))
        assert np.array_equal(xD, eD)
        if p.mode >= 2:
            assert np.array_equal(xT, eT)


def test_inveon_size_requests_in_the_mix(gctx, orc):
    """The served stream's large classes beside small requests: a c1-size
    doubles request (Inveon geometry, FP16, 16 subsets) and a c2-size joint
    request on the same geometry (mode 3, doubles + triples sinograms and
    4000 VLOR records, FP16), submitted with three c0-size requests. Every
    result equals the same request alone, bit for bit, and the c1-size
    doubles image equals the FP32 oracle on the FP16-rounded inputs (one
    iteration, about 15 s on 16 host threads)."""
    mix = [(C0, 1, 4, 2, ga.STORE_F16, 1e6), (C1, 1, 16, 1, ga.STORE_F16, 5e7),
           (C0, 1, 8, 1, ga.STORE_F16, 3e5), (C1, 3, 16, 1, ga.STORE_F16, 2e7),
           (C0, 1, 4, 1, ga.STORE_F16, 1e5)]
    srv = gctx.server(max_active=4, inflight=2)
    ids, reqs = [], []
    for k, (g, mode, S, ni, st, counts) in enumerate(mix):
        d2, t2, norm, rec = request(g, 800 + k, mode, counts)
        if rec is not None:
            rec = np.random.default_rng(880 + k).integers(1, g.ndata, size=(4000, 2)).astype(
                np.int32)
        p = ga.recon_params(mode=mode, niter=ni, nsub=S, storage=st)
        reqs.append((g, p, d2, t2, norm, rec))
        ids.append(srv.submit(g, p, d2, t2, norm, rec))
    for rid in ids:
        assert srv.wait(rid, 300.0) == ga.REQ_DONE, srv.info(rid)
    got = [srv.result(rid) for rid in ids]
    assert srv.stats()["failed"] == 0
    srv.close()
    for k, ((g, p, d2, t2, norm, rec), (xD, xT)) in enumerate(zip(reqs, got)):
        eD, eT = alone(gctx, g, p, d2, t2, norm, rec)
        n = int(np.sum(xD == eD))
        print("request %d (%d voxels, mode %d, S=%d): xD %d/%d bit-exact" % (
            k, g.nvox, p.mode, p.nsub, n, xD.size))
        assert np.array_equal(xD, eD)
        if p.mode >= 2:
            assert np.array_equal(xT, eT)
    g, p, d2, t2, norm, rec = reqs[1]
    f16 = lambda a: a.astype(np.float16).astype(np.float32)
    ex, _, _, _ = orc.osem(g, orc.params(mode=1, niter=1, nsub=16), f16(d2), None, norm)
    assert np.array_equal(got[1][0], ex)


def test_kept_blocks_reused_by_later_sessions(gctx):
    """Blocks of finished sessions are kept by size and handed to the next
    session of the same class with the previous contents: a request run
    again after the same request, after a joint request of another geometry
    and after a doubles request with other data gives the same bits."""
    reqs = [(C0, 1, 4, 2, ga.STORE_F16, 1e6), (SMALL, 3, 4, 1, ga.STORE_F32, 2e5)]
    data = [request(g, 950 + k, mode, counts) for k, (g, mode, _, _, _, counts) in
            enumerate(reqs)]
    params = [ga.recon_params(mode=mode, niter=ni, nsub=S, storage=st)
              for (_, mode, S, ni, st, _) in reqs]
    first = [alone(gctx, g, p, *d) for (g, *_), p, d in zip(reqs, params, data)]
    other = request(C0, 960, 1, 3e5)
    for _ in range(2):
        alone(gctx, C0, params[0], *other)
        for k, ((g, *_), p, d) in enumerate(zip(reqs, params, data)):
            xD, xT = alone(gctx, g, p, *d)
            assert np.array_equal(xD, first[k][0]), k
            if p.mode >= 2:
                assert np.array_equal(xT, first[k][1]), k


def test_admission_waits_for_budget_and_isolates_failures(gctx):
    g = C0
    p = ga.recon_params(mode=1, niter=1, nsub=4, storage=ga.STORE_F32)
    need = ga.footprint(g, p)
    ctx = ga.Context(0, hbm_budget=int(need * 1.5))  # one session at a time
    srv = ctx.server(max_active=8)
    data = [request(g, 600 + k, 1, 3e5) for k in range(3)]
    ids = [srv.submit(g, p, d2, None, norm) for d2, _, norm, _ in data]
    # a request that cannot fit even alone, and one with invalid parameters
    big = ga.recon_params(mode=1, niter=1, nsub=1, storage=ga.STORE_F32)
    d2b, _, nb, _ = request(SMALL, 610, 1, 1e5)
    huge_geom = g.with_(nzs=g.nzs * 4)
    bad = srv.submit(g, ga.recon_params(mode=1, niter=1, nsub=g.nang + 1), data[0][0], None,
                     data[0][2])
    for rid in ids + [bad]:
        srv.wait(rid, 120.0)
    infos = [srv.info(rid) for rid in ids]
    assert all(i["state"] == ga.REQ_DONE for i in infos)
    # admitted one after another: each waited for the previous to retire
    for a, b in zip(infos, infos[1:]):
        assert b["t_admit"] >= a["t_done"] - 1e-3
    ib = srv.info(bad)
    assert ib["state"] == ga.REQ_FAILED and ib["error"] == ga.GFR_EINVAL
    with pytest.raises(ga.GfrError) as e:
        srv.result(bad)
    assert e.value.code == ga.GFR_EINVAL
    nd4 = huge_geom.ndata
    rid = srv.submit(huge_geom, big, np.zeros(2 * nd4, np.float32), None,
                     np.ones(nd4, np.float32))
    assert srv.wait(rid, 120.0) == ga.REQ_FAILED and srv.info(rid)["error"] == ga.GFR_ENOMEM
    for k, rid in enumerate(ids):
        xD, _ = srv.result(rid)
        eD, _ = alone(gctx, g, p, data[k][0], None, data[k][2], None)
        assert np.array_equal(xD, eD)
        srv.release(rid)
    srv.close()
    ctx.close()
    del d2b, nb


def test_nan_inputs_and_cancellation_are_isolated(gctx):
    """Faults forced from the host into a running server: a request with a
    NaN prompt and one with an infinite background fail with GFR_EINVAL; a
    request whose normalization holds NaN completes with those bins masked
    (as the reference's > 0.01 test masks them: the same image as with 0
    there); a running request and a pending one are cancelled
    (GFR_ECANCELED). Every other request comes back with its bits alone."""
    g = C0
    srv = gctx.server(max_active=2, inflight=2)
    good = [request(g, 900 + k, 1, 3e5) for k in range(3)]
    p = ga.recon_params(mode=1, niter=2, nsub=4)
    long_p = ga.recon_params(mode=1, niter=40, nsub=16)
    nan_d2 = good[0][0].copy()
    nan_d2[12345] = np.nan
    inf_d2 = good[1][0].copy()
    inf_d2[g.ndata + 77] = np.inf
    nan_norm = good[2][2].copy()
    nan_norm[::997] = np.nan
    ids = {}
    ids["long"] = srv.submit(g, long_p, good[0][0], None, good[0][2])
    ids["g0"] = srv.submit(g, p, good[0][0], None, good[0][2])
    ids["nan"] = srv.submit(g, p, nan_d2, None, good[0][2])
    ids["inf"] = srv.submit(g, p, inf_d2, None, good[1][2])
    ids["g1"] = srv.submit(g, p, good[1][0], None, good[1][2])
    ids["pending"] = srv.submit(g, long_p, good[2][0], None, good[2][2])
    ids["nannorm"] = srv.submit(g, p, good[2][0], None, nan_norm)
    srv.cancel(ids["pending"])
    t0 = time.time()
    while srv.info(ids["long"])["state"] == ga.REQ_PENDING and time.time() - t0 < 60:
        time.sleep(0.001)
    srv.cancel(ids["long"])
    for rid in ids.values():
        assert srv.wait(rid, 120.0) in (ga.REQ_DONE, ga.REQ_FAILED)
    info = {k: srv.info(v) for k, v in ids.items()}
    for k in ("long", "pending"):
        assert info[k]["state"] == ga.REQ_FAILED and info[k]["error"] == ga.GFR_ECANCELED, k
    assert info["pending"]["passes"] == 0
    assert info["long"]["passes"] < 40 * 16
    for k in ("nan", "inf"):
        assert info[k]["state"] == ga.REQ_FAILED and info[k]["error"] == ga.GFR_EINVAL, k
    with pytest.raises(ga.GfrError) as e:
        srv.cancel(ids["g0"])  # already complete
    assert e.value.code == ga.GFR_ESTATE
    res = {k: srv.result(ids[k])[0] for k in ("g0", "g1", "nannorm")}
    srv.close()
    assert np.array_equal(res["g0"], alone(gctx, g, p, good[0][0], None, good[0][2], None)[0])
    assert np.array_equal(res["g1"], alone(gctx, g, p, good[1][0], None, good[1][2], None)[0])
    zero_norm = np.where(np.isnan(nan_norm), np.float32(0), nan_norm)
    exp = alone(gctx, g, p, good[2][0], None, zero_norm, None)[0]
    assert np.array_equal(res["nannorm"], exp) and np.all(np.isfinite(exp))


def test_cancel_running_joint_request(gctx):
    """A joint request (mode 3: triples passes on the session's second
    stream) cancelled while both of its streams have work queued stops
    between passes with GFR_ECANCELED; the requests beside it and a joint
    request admitted afterwards come back with their bits alone."""
    g = C0
    srv = gctx.server(max_active=3, inflight=2)
    long_p = ga.recon_params(mode=3, niter=30, nsub=16)
    p3 = ga.recon_params(mode=3, niter=2, nsub=4)
    p2 = ga.recon_params(mode=2, niter=2, nsub=4)
    jl = request(g, 951, 3, 3e5)
    j3 = request(g, 953, 3, 3e5)
    j2 = request(g, 955, 2, 3e5)
    rid_long = srv.submit(g, long_p, jl[0], jl[1], jl[2], jl[3])
    rid2 = srv.submit(g, p2, j2[0], j2[1], j2[2], None)
    t0 = time.time()
    # past the first iteration, so triples passes of one iteration and
    # doubles passes of the next are both in flight
    while srv.info(rid_long)["passes"] < 40 and time.time() - t0 < 60:
        time.sleep(0.001)
    assert srv.info(rid_long)["state"] == ga.REQ_ACTIVE
    srv.cancel(rid_long)
    rid3 = srv.submit(g, p3, j3[0], j3[1], j3[2], j3[3])
    for rid in (rid_long, rid2, rid3):
        assert srv.wait(rid, 120.0) in (ga.REQ_DONE, ga.REQ_FAILED)
    info = srv.info(rid_long)
    assert info["state"] == ga.REQ_FAILED and info["error"] == ga.GFR_ECANCELED
    assert 40 <= info["passes"] < 30 * 2 * 16
    got2, got3 = srv.result(rid2), srv.result(rid3)
    srv.close()
    exp2 = alone(gctx, g, p2, j2[0], j2[1], j2[2], None)
    exp3 = alone(gctx, g, p3, j3[0], j3[1], j3[2], j3[3])
    for got, exp, what in ((got2, exp2, "mode 2"), (got3, exp3, "mode 3")):
        assert np.array_equal(got[0], exp[0]) and np.array_equal(got[1], exp[1]), what


def test_destroy_with_work_in_flight(gctx):
    srv = gctx.server(max_active=2)
    ids = []
    for k in range(6):
        d2, _, norm, _ = request(C0, 700 + k, 1, 2e5)
        ids.append(srv.submit(C0, ga.recon_params(mode=1, niter=2, nsub=4), d2, None, norm))
    time.sleep(0.05)
    t0 = time.time()
    srv.close()
    assert time.time() - t0 < 60
