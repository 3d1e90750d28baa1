"""Batched frames (gfr_batch_*): every frame's result must equal, bit for bit,
a gfr_recon session (mode 1) with the same data and parameters, whether the
steps replay the captured graph or launch the kernels one by one, with frames
of uneven subset counts, frames that join mid-run, reduced-precision storage,
and eager work and synchronisations between graph replays. Sessions are
themselves pinned to the oracle (test_gpu_parity.py)."""
import numpy as np
import pytest

import gfirst_amd as ga
from gfirst_amd import synth
from gfirst_amd.geometry import C0, C1, REF, SMALL

pytestmark = pytest.mark.gpu


def frame_data(g, seed, counts=1e5, bg=True):
    rng = np.random.default_rng(seed)
    nd = g.ndata
    lam = rng.uniform(0.2, 3.0, nd) * (counts / nd)
    y = rng.poisson(lam).astype(np.float32)
    b = rng.uniform(0, 0.3, nd).astype(np.float32) if bg else None
    return y, b


def norm_of(g, seed=99):
    rng = np.random.default_rng(seed)
    n = rng.uniform(0.8, 1.2, g.ndata).astype(np.float32)
    n[rng.uniform(size=g.ndata) < 0.02] = 0.0
    return n


def session_result(ctx, g, norm, y, b, niter, nsub, storage):
    d2 = np.zeros(2 * g.ndata, np.float32)
    d2[:g.ndata] = y
    if b is not None:
        d2[g.ndata:] = b
    r = ctx.recon(g, ga.recon_params(mode=1, niter=niter, nsub=nsub, storage=storage), d2, None,
                  norm)
    r.run()
    x = r.get(ga.OUT_XD)
    ll = r.loglik()
    r.close()
    return x, ll


def same(a, b, what):
    n = int(np.sum(a == b))
    print("%s: %d/%d bit-exact" % (what, n, a.size))
    assert np.array_equal(a, b), what


@pytest.mark.parametrize("storage", [ga.STORE_F32, ga.STORE_F16, ga.STORE_BF16],
                         ids=["f32", "f16", "bf16"])
@pytest.mark.parametrize("graph", [True, False], ids=["graph", "eager"])
def test_batch_frames_equal_sessions(gctx, storage, graph):
    g = SMALL
    norm = norm_of(g)
    spec = [(2, 2, True), (1, 3, False), (2, 4, True), (1, 8, True)]  # (niter, nsub, bg)
    bp = ga.batch_params(storage=storage, max_frames=8, min_nsub=2, lanes=2)
    bt = gctx.batch(g, bp, norm)
    bt.use_graph(graph)
    data = []
    for k, (ni, ns, hb) in enumerate(spec):
        y, b = frame_data(g, 100 + k, bg=hb)
        data.append((y, b))
        assert bt.add(ni, ns, y, b) == k
    left = []
    while True:
        n = bt.step()
        left.append(n)
        if n == 0:
            break
    bt.sync()
    # frames finish after niter*nsub steps: 4, 3, 8, 8
    assert left == [4, 4, 3, 2, 2, 2, 2, 0]
    for k, (ni, ns, _) in enumerate(spec):
        c = bt.cursor(k)
        assert c["done"] and c["passes"] == ni * ns
        x_ref, ll_ref = session_result(gctx, g, norm, *data[k], ni, ns, storage)
        same(bt.get(k), x_ref, "frame %d (S=%d, %s)" % (k, ns, "graph" if graph else "eager"))
        assert bt.loglik(k) == ll_ref
    st = bt.stats()
    if graph:
        # one build per distinct membership: 4 frames, then 3 after frame 1, 2 after frame 0
        assert st["builds"] == 3 and st["launches"] == 8
    else:
        assert st["launches"] == 0
    bt.close()


@pytest.mark.parametrize("storage,graph", [(ga.STORE_F32, True), (ga.STORE_F16, False),
                                           (ga.STORE_BF16, True)],
                         ids=["f32-graph", "f16-eager", "bf16-graph"])
def test_frame_groups_equal_sessions(gctx, storage, graph):
    """Frames at the same subset of the same subset count share one
    backprojection launch (groups of up to three over interleaved rows). Seven
    S=4 frames (groups 3+3+1, regrouped as members finish), two S=2 frames (a
    group of two), one S=3 frame alone, and an S=4 frame that joins two steps
    late (another subset phase, so never grouped with the first seven): every
    frame equals its own session bit for bit, image and FP64 log-likelihood,
    and a batch without groups gives the same bits."""
    g = SMALL
    norm = norm_of(g, 11)
    spec = [(2, 4, True), (2, 4, False), (1, 4, True), (1, 4, True), (2, 4, False),
            (1, 4, True), (1, 4, True), (1, 2, True), (2, 2, False), (1, 3, True)]
    late = (1, 4, True)
    data = [frame_data(g, 500 + k, bg=hb) for k, (_, _, hb) in enumerate(spec + [late])]
    out = {}
    for grp in (0, 1):
        bp = ga.batch_params(storage=storage, max_frames=12, min_nsub=2, lanes=2, group=grp)
        bt = gctx.batch(g, bp, norm)
        bt.use_graph(graph)
        for k, (ni, ns, _) in enumerate(spec):
            assert bt.add(ni, ns, *data[k]) == k
        bt.step()
        bt.step()
        assert bt.add(late[0], late[1], *data[-1]) == len(spec)
        bt.run()
        out[grp] = [(bt.get(k), bt.loglik(k)) for k in range(len(spec) + 1)]
        bt.close()
    for k, (ni, ns, _) in enumerate(spec + [late]):
        x_ref, ll_ref = session_result(gctx, g, norm, *data[k], ni, ns, storage)
        for grp in (0, 1):
            same(out[grp][k][0], x_ref, "frame %d (S=%d) group=%d" % (k, ns, grp))
            assert out[grp][k][1] == ll_ref


def test_frame_groups_at_reference_geometry(gctx):
    """The reference's compiled geometry (GFIRST.h:4-16: NRAD 147, two
    z-sum radial blocks, odd-length interleaved rows; 168 angles, span 11):
    four S=12 frames (a group of three and a lone frame) and two S=6 frames (a
    group of two), FP16 storage, graph steps; each equals its own session bit
    for bit, image and FP64 log-likelihood."""
    g = REF
    norm = norm_of(g, 13)
    spec = [(1, 12)] * 4 + [(1, 6)] * 2
    data = [frame_data(g, 700 + k, counts=2e6, bg=(k % 2 == 0)) for k in range(len(spec))]
    bp = ga.batch_params(storage=ga.STORE_F16, max_frames=len(spec), min_nsub=6, lanes=2)
    bt = gctx.batch(g, bp, norm)
    for k, (ni, ns) in enumerate(spec):
        assert bt.add(ni, ns, *data[k]) == k
    bt.run()
    got = [(bt.get(k), bt.loglik(k)) for k in range(len(spec))]
    bt.close()
    for k, (ni, ns) in enumerate(spec):
        x_ref, ll_ref = session_result(gctx, g, norm, *data[k], ni, ns, ga.STORE_F16)
        same(got[k][0], x_ref, "REF frame %d (S=%d)" % (k, ns))
        assert got[k][1] == ll_ref


def test_graph_replays_match_eager_each_step(gctx):
    """Every replay of the captured step against the same step launched
    eagerly, with eager library work and host synchronisation in between;
    three of the frames share their backprojection (one group)."""
    g = C0
    norm = norm_of(g, 5)
    bp = ga.batch_params(storage=ga.STORE_F16, max_frames=6, min_nsub=4, lanes=2)
    bg_ = gctx.batch(g, bp, norm)
    be = gctx.batch(g, bp, norm)
    be.use_graph(False)
    spec = [(1, 4), (1, 8), (1, 16), (1, 4), (1, 4)]
    for k, (ni, ns) in enumerate(spec):
        y, b = frame_data(g, 200 + k, counts=3e5)
        bg_.add(ni, ns, y, b)
        be.add(ni, ns, y, b)
    junk = gctx.to_device(np.ones(g.nvox, np.float32))
    for step in range(16):
        a = bg_.step()
        e = be.step()
        assert a == e
        # eager work on the context's stream between replays
        gctx.forward_project(g, junk, step % 4, 4)
        gctx.sync()
        for k in range(len(spec)):
            same(bg_.get(k), be.get(k), "step %d frame %d" % (step, k))
            assert bg_.loglik(k) == be.loglik(k)
        if a == 0:
            break
    assert bg_.stats()["launches"] == 16
    bg_.close()
    be.close()


def test_frame_joining_mid_run_and_slot_reuse(gctx):
    g = SMALL
    norm = norm_of(g, 7)
    bp = ga.batch_params(storage=ga.STORE_F32, max_frames=2, min_nsub=2, lanes=1)
    bt = gctx.batch(g, bp, norm)
    y0, b0 = frame_data(g, 300)
    y1, b1 = frame_data(g, 301)
    y2, b2 = frame_data(g, 302, bg=False)
    f0 = bt.add(2, 3, y0, b0)
    bt.step()
    bt.step()
    f1 = bt.add(1, 4, y1, b1)  # joins while frame 0 is mid-iteration
    with pytest.raises(ga.GfrError) as e:
        bt.add(1, 2, y2, b2)  # no free slot
    assert e.value.code == ga.GFR_ESTATE
    while bt.cursor(f0)["passes"] < 6:
        bt.step()
    same(bt.get(f0), session_result(gctx, g, norm, y0, b0, 2, 3, ga.STORE_F32)[0], "frame 0")
    bt.remove(f0)
    f2 = bt.add(1, 2, y2, b2)  # reuses the slot, joins mid-run of frame 1
    assert f2 == f0
    bt.run()
    same(bt.get(f1), session_result(gctx, g, norm, y1, b1, 1, 4, ga.STORE_F32)[0], "frame 1")
    same(bt.get(f2), session_result(gctx, g, norm, y2, None, 1, 2, ga.STORE_F32)[0], "frame 2")
    bt.close()


def test_device_resident_frame_data_and_budget(gctx):
    g = SMALL
    norm = norm_of(g, 8)
    bp = ga.batch_params(max_frames=4, min_nsub=2, lanes=1)
    y, b = frame_data(g, 400)
    # room for the batch, one frame with background, and the uploaded prompts
    ctx = ga.Context(0, hbm_budget=ga.batch_footprint(g, bp, 1, has_bg=True) + 4 * g.ndata)
    bt = ctx.batch(g, bp, norm)
    dy = ctx.to_device(y)
    f = bt.add(1, 2, dy, None)  # device pointer, no background
    with pytest.raises(ga.GfrError) as e:
        bt.add(1, 2, y, b)  # over the budget
    assert e.value.code == ga.GFR_ENOMEM
    bt.run()
    same(bt.get(f), session_result(gctx, g, norm, y, None, 1, 2, ga.STORE_F32)[0], "device frame")
    with pytest.raises(ga.GfrError):
        bt.add(1, 1, y, b)  # nsub below min_nsub
    bt.close()
    ctx.close()


def test_inveon_size_batch_fp16_uneven_subsets(gctx, orc):
    """The batched dynamic study at its size: three frames on the Inveon
    geometry (image 128x128x159, sinogram 128x160x4319) with 8, 16 and 24
    subsets and uneven counts, FP16 storage, one iteration each, advanced
    together by the captured graph over three lanes, plus a second S=8 frame
    that shares the first one's backprojection launches. Every frame equals
    its own session bit for bit (image and FP64 log-likelihood), and the first
    S=8 frame (with a non-integer background) equals the FP32 oracle run on
    the FP16-rounded inputs: one S=8 iteration, about 15 s on 16 host
    threads."""
    g = C1
    nd = g.ndata
    norm = synth.norm_uniform(nd, seed=61)
    # (nsub, counts, bg)
    spec = [(8, 3e6, True), (16, 1e7, False), (24, 1e5, True), (8, 1e6, False)]
    bp = ga.batch_params(storage=ga.STORE_F16, max_frames=4, min_nsub=8, lanes=3)
    bt = gctx.batch(g, bp, norm)
    data = []
    for k, (S, counts, hb) in enumerate(spec):
        y, b = frame_data(g, 900 + k, counts=counts, bg=hb)
        data.append((y, b))
        assert bt.add(1, S, y, b) == k
    bt.run()
    assert bt.stats()["launches"] == 24
    for k, (S, _, _) in enumerate(spec):
        c = bt.cursor(k)
        assert c["done"] and c["passes"] == S
        x_ref, ll_ref = session_result(gctx, g, norm, *data[k], 1, S, ga.STORE_F16)
        same(bt.get(k), x_ref, "c1 frame %d (S=%d)" % (k, S))
        assert bt.loglik(k) == ll_ref
    x0 = bt.get(0)
    bt.close()
    y, b = data[0]
    f16 = lambda a: a.astype(np.float16).astype(np.float32)
    assert not np.array_equal(f16(b), b)
    d2 = np.concatenate([f16(y), f16(b)])
    ex, _, _, _ = orc.osem(g, orc.params(mode=1, niter=1, nsub=8), d2, None, norm)
    same(x0, ex, "c1 frame 0 vs the oracle on FP16-rounded inputs")


@pytest.mark.parametrize("lam", [0.05, 0.7, 4.0, 9.99, 10.0, 37.0, 2500.0])
def test_poisson_moments_and_determinism(gctx, lam):
    n = 1 << 20
    mean = gctx.to_device(np.full(n, lam, np.float32))
    a = gctx.poisson(mean, 1.0, 1234)
    b = gctx.poisson(mean, 1.0, 1234)
    c = gctx.poisson(mean, 1.0, 1235)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert np.all(a >= 0) and np.all(a == np.round(a))
    m, v = a.mean(dtype=np.float64), a.var(dtype=np.float64)
    se_m = np.sqrt(lam / n)
    print("lam %g: mean %.5g (se %.2g) var %.5g" % (lam, m, se_m, v))
    assert abs(m - lam) < 6 * se_m
    # var of the sample variance of a Poisson: (lam + 2 lam^2) / n
    assert abs(v - lam) < 6 * np.sqrt((lam + 2 * lam * lam) / n)
    if lam < 1:
        p0 = np.mean(a == 0)
        assert abs(p0 - np.exp(-lam)) < 6 * np.sqrt(np.exp(-lam) * (1 - np.exp(-lam)) / n)
