"""Serving loop (gfr_server_*): requests of different geometries, modes,
storage formats and count levels submitted together come back bit-identical
to the same request reconstructed alone; admission respects slots and the
HBM budget; a bad or oversized request fails with its own status code while
the others complete."""
import time

import numpy as np
import pytest

import gfirst_amd as ga
from gfirst_amd.geometry import C0, SMALL, TINY

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


def test_destroy_with_work_in_flight(gctx):
    srv = gctx.server(max_active=2)
    for k in range(6):
        d2, _, norm, _ = request(C0, 700 + k, 1, 2e5)
        srv.submit(C0, ga.recon_params(mode=1, niter=2, nsub=4), d2, None, norm)
    time.sleep(0.05)
    t0 = time.time()
    srv.close()
    assert time.time() - t0 < 60
