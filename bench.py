"""Benchmark: Inveon-scale doubles OSEM on one MI355X.

Workload: 128x128x159 image (0.776 mm), 128 radial x 160 angles x 4319
oblique sinograms (80 rings, span 3, max ring difference 79), 16 subsets,
FP32, synthetic mouse-like phantom, 5e7 Poisson counts, normalization
U[0.8, 1.2]. A step is one subset-iteration: PSF blur, forward projection +
ratio (+ FP64 log-likelihood), backprojection and image update of one
subset, all on the device with the data resident in HBM. The same line
carries the metric's latency half ("serving"): --serve-s seconds (default
60) of the served stream below (the c4 mix at 4 requests/s, FP16 storage),
with p50/p99 from submission to the image on the host and the alone-vs-mix
check of the first request of each class.

--workload c0: the small plain study (64x64x32 image, 64 bins x 64 angles
x 32 rings, span 3, max ring difference 15; uniform 25 mm cylinder with three
hot 4 mm spheres, 1e6 Poisson counts, normalization 1.0): the whole
4-subset x 2-iteration reconstruction and the projection-only run timed on
the GPU with the data resident; the oracle times the same reconstruction on
the host and its image is compared with the GPU's bit for bit.

--workload c2: the joint doubles + triples study (mode 3) at the same
Inveon size: doubles Poisson 5e7 counts from phantom channel 0, triples
Poisson 1e6 counts from channel 1, 2e6 triple events binned to VLOR records;
a step is one pass (one subset of one type: doubles, or triples with VLOR
assignment, JBF and blend), timed over whole iterations. The oracle times one
pass of each type. --c2-mode 2 runs the same study in mode 2 (triples from
the triples sinogram, no VLOR records).

--workload c3: the batched dynamic study. 64 frames of the same geometry,
counts log-uniform in [1e5, 1e7], subset counts drawn from {8, 16, 24},
activity = phantom channels weighted by a seeded two-compartment time curve,
FP16 storage by default; every frame advances one subset per launch of one
captured hipGraph until all are complete. Compared with the same frames run
one study at a time as single sessions.

--workload c4: the served stream. Requests arrive as a Poisson process (4/s
for 300 s), 60% c0-size, 30% c1-size, 10% joint doubles+triples (VLOR) at
c1 size, counts log-uniform in [1e5, 5e7], 2-4 iterations, FP16 storage,
200 GB HBM budget; p50/p99 latency from submission to the image on the host.
The process runs with --hw-queues hardware queues (GPU_MAX_HW_QUEUES, default
16: one per admitted session; at most 32).

--copy-probe: a plain 1 GiB device copy (torch) repeated, for scaling the
profiler's FETCH_SIZE / WRITE_SIZE counters to bytes.

Prints one JSON line (stdout). Progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0      # MI355X HBM3E
PEAK_FP32_TFLOPS = 157.3   # MI355X FP32 vector
STORAGE = {"f32": 0, "f16": 1, "bf16": 2}  # GFR_STORE_*


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def make_inputs(ctx, g, counts, seed):
    import gfirst_amd as ga
    from gfirst_amd import synth
    img = synth.phantom_mouse(g, seed=seed, channel=0)
    d_img = ctx.to_device(img)
    d_blur = ctx.array(g.nvox)
    ctx.blur_xy(g, d_img, 1.5, out=d_blur)
    ybar = ctx.forward_project(g, d_blur)
    y = synth.poisson_sinogram(ybar, counts, seed=seed + 1).astype(np.float32)
    d2 = synth.as_file_pair(y)
    norm = synth.norm_uniform(g.ndata, seed=seed + 2)
    return d2, norm


def alg_bytes(g, nsub, storage_bytes, has_bg):
    """Algorithmic HBM bytes of one subset-iteration and of its two projector
    kernels. Per LOR of the subset: measured counts (+ background) read once,
    the ratio written once by the forward kernel and read once (with the
    1-byte calibration mask) by the backprojector. Per voxel: the blurred
    image read once by the forward kernel, NUMER and DENOM written once."""
    L = g.ndata // nsub
    nv = g.nvox
    fwd = L * (storage_bytes * (2 if has_bg else 1) + 4) + 4 * nv
    back = L * (4 + 1) + 8 * nv
    blur = 8 * nv
    update = 4 * nv * 2 + 4 * nv * 2 + 4 * nv * 2  # num+den in, work out; work in, x rw
    return dict(forward=fwd, back=back, blur=blur, update=update,
                total=fwd + back + blur + update)


def step_alg_bytes(g, nsub, storage_bytes, has_bg):
    """Algorithmic HBM bytes of one subset-iteration by the benchmark plan's
    formula B_alg = L_s*(b_y + b_bg + b_n + 8) + 16*NVOX: per LOR of the
    subset the counts, background and normalization read once (b_n = 4: an
    FP32 normalization value) and the FP32 ratio written and read once; per
    voxel the image read twice, written once and one intermediate."""
    L = g.ndata // nsub
    return L * (storage_bytes * (2 if has_bg else 1) + 4 + 8) + 16 * g.nvox


def alg_flops(g, nsub):
    """FP32 flops of the two projectors per subset-iteration: 22 per
    trilinear sample (7 lerps x 3 + 1 add), 10 per bilinear voxel-view fetch
    of the ratio image (3 lerps x 3 + 1 add)."""
    L = g.ndata // nsub
    nviews = g.nseg * (g.nang // nsub)
    return dict(forward=22.0 * L * (g.res + 1), back=10.0 * g.nvox * nviews)


def measured_traffic(dom, back_zb, storage, back_lds, back_bw=4):
    """HBM bytes per launch of the backprojector from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE runs (scaled by a plain copy's counts),
    for the tile shape in use; None when no measurement matches."""
    path = os.path.join(ROOT, "profiles", "k_back_c1_traffic.json")
    if dom != "back" or storage != 0 or back_bw != 4 or not os.path.exists(path):
        return None
    with open(path) as f:
        m = json.load(f)
    kind = {0: "direct:", 1: "lds:"}.get(back_lds)
    if kind is None:
        return None
    key = kind + str(back_zb)
    e = m.get("by_kind_slices", {}).get(key)
    return None if e is None else e["fetch_bytes"] + e["write_bytes"]


def measured_pipes(dom, tune):
    """Busy fractions of the CU's vector-memory units (texture addresser TA,
    texture data TD) for the backprojector shape in use, from the committed
    rocprofv3 counter runs; None when no measurement matches."""
    path = os.path.join(ROOT, "profiles", "k_back_c1_pipes.json")
    if dom != "back" or tune.get("back_lds") or tune.get("back_bw", 4) != 4 or \
            not os.path.exists(path):
        return None
    with open(path) as f:
        m = json.load(f)
    e = m.get("by_kernel", {}).get("k_back_direct<%d, %d, float>" % (tune["back_zb"],
                                                                    tune["back_waves"]))
    if e is None:
        return None
    # VALU issue: a CU's four 16-lane SIMDs issue about one wave64 VALU
    # instruction per cycle between them (4 cycles each)
    return {"unit": "VALU issue and the vector-memory path (TA address, TD data)",
            "valu_wave_instr_per_cu_cycle": e["valu_wave_instr_per_cu_cycle"],
            "valu_issue_limit_per_cu_cycle": 1.0, "ta_busy": e["ta_busy"],
            "td_busy": e["td_busy"], "source": "profiles/k_back_c1_pipes.json"}


def cpu_baseline(g, d2, norm, pb, nsub, nsteps):
    """The CPU oracle (FP32, OpenMP) on a bounded sample: nsteps
    subset-iterations of the same study (the untimed mask set-up reuses the
    device's PROJ_BLANK)."""
    from oracle import oracle as orc
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    orc.set_threads(threads)
    nd = g.ndata
    x = orc.cylinder(g, 1.0)
    cal = orc.calib(g, pb, norm)
    y = np.float32(g.res) * d2[:nd]
    t0 = time.perf_counter()
    for s in range(nsteps):
        xb = orc.blur_xy(g, x, 1.5)
        proj = orc.forward(g, xb, s, nsub)
        corr = orc.ratio(g, proj, y, None, s, nsub)
        num, den = orc.back(g, corr, cal, s, nsub)
        x = orc.update(g, x, num, den, 1.5)
    dt = time.perf_counter() - t0
    cores = orc.get_threads()
    # the same work on one thread: one subset-iteration
    orc.set_threads(1)
    t0 = time.perf_counter()
    xb = orc.blur_xy(g, x, 1.5)
    proj = orc.forward(g, xb, 0, nsub)
    corr = orc.ratio(g, proj, y, None, 0, nsub)
    num, den = orc.back(g, corr, cal, 0, nsub)
    orc.update(g, x, num, den, 1.5)
    d1 = time.perf_counter() - t0
    orc.set_threads(threads)
    return dict(value=nsteps / dt, unit="subset-iters/s", cores=cores, kind="port",
                sample="%d subset-iterations (S=%d) of the same Inveon-scale study, "
                       "oracle FP32 restatement with OpenMP" % (nsteps, nsub),
                seconds=dt,
                single_thread={"value": 1.0 / d1, "cores": 1, "seconds": d1,
                               "sample": "1 subset-iteration of the same study"})


def cpu_baseline_c0(g, d2, norm, nsub, niter, x_gpu):
    """The oracle runs the same c0 reconstruction on the host (timed), and
    its image is compared with the GPU's."""
    sys.path.insert(0, ROOT)
    from oracle import oracle as orc
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
    orc.set_threads(threads)
    t0 = time.perf_counter()
    xD, _, _, _ = orc.osem(g, orc.params(mode=1, niter=niter, nsub=nsub), d2, None, norm)
    dt = time.perf_counter() - t0
    base = dict(value=niter * nsub / dt, unit="subset-iters/s", cores=orc.get_threads(),
                kind="port", sample="the whole c0 reconstruction (%d subsets x %d iterations), "
                "oracle FP32 restatement with OpenMP" % (nsub, niter), seconds=dt)
    orc.set_threads(1)
    t0 = time.perf_counter()
    orc.osem(g, orc.params(mode=1, niter=niter, nsub=nsub), d2, None, norm)
    d1 = time.perf_counter() - t0
    orc.set_threads(threads)
    base["single_thread"] = {"value": niter * nsub / d1, "cores": 1, "seconds": d1,
                             "sample": "the same reconstruction on one thread"}
    exact = int(np.sum(xD == x_gpu))
    return base, {"bit_exact_voxels": exact, "voxels": int(xD.size)}


def run_c0(args):
    import gfirst_amd as ga
    from gfirst_amd import synth
    from gfirst_amd.geometry import C0
    g = C0
    nsub, niter = 4, 2
    ctx = ga.Context(0)
    img = synth.phantom_c0(g, seed=args.seed)
    d_img = ctx.to_device(img)
    d_blur = ctx.array(g.nvox)
    ctx.blur_xy(g, d_img, 1.5, out=d_blur)
    ybar = ctx.forward_project(g, d_blur)
    y = synth.poisson_sinogram(ybar, 1e6, seed=args.seed + 1).astype(np.float32)
    d2 = synth.as_file_pair(y)
    norm = np.ones(g.ndata, np.float32)
    # projection-only run (mode 1 of the reference): the phantom's sinogram
    d_sino = ctx.array(g.ndata)
    for _ in range(3):
        ctx.forward_project(g, d_img, out=d_sino)
    ctx.sync()
    reps = 20
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.forward_project(g, d_img, out=d_sino)
    ctx.sync()
    ms_proj = 1e3 * (time.perf_counter() - t0) / reps
    # the reconstruction: session set-up (uploads, mask, autotuning) untimed
    p = ga.recon_params(mode=1, niter=niter, nsub=nsub)
    runs = []
    for k in range(4):
        r = ctx.recon(g, p, d2, None, norm)
        r.sync()
        t0 = time.perf_counter()
        r.run()
        r.sync()
        runs.append(time.perf_counter() - t0)
        if k == 3:
            x = r.get(ga.OUT_XD)
        r.close()
    dt = min(runs[1:])
    log("c0: reconstruction %.3f ms (%d subset-iterations), projection %.3f ms" % (
        1e3 * dt, niter * nsub, ms_proj))
    res = {
        "metric": "c0 plain OSEM subset-iters/s",
        "value": round(niter * nsub / dt, 2),
        "unit": "subset-iters/s",
        "n_gpus": 1,
        "steps": niter * nsub,
        "warmup": 1,
        "ms_per_step": round(1e3 * dt / (niter * nsub), 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded cylinder + 3 hot spheres, Poisson 1e6 counts, norm 1.0)",
        "config": {"workload": "c0 plain OSEM: image 64x64x32, sinogram 64x64x%d, span 3, "
                               "MRD 15, 4 subsets x 2 iterations" % g.nsinogs},
        "reconstruction_ms": round(1e3 * dt, 3),
        "projection_ms": round(ms_proj, 4),
    }
    if not args.no_cpu:
        log("timing the oracle on the same reconstruction ...")
        res["cpu_baseline"], res["parity"] = cpu_baseline_c0(g, d2, norm, nsub, niter, x)
        log("cpu: %.2f subset-iters/s; %d/%d voxels bit-exact" % (
            res["cpu_baseline"]["value"], res["parity"]["bit_exact_voxels"], res["parity"]["voxels"]))
    ctx.close()
    return res


def run_c1(args):
    import gfirst_amd as ga
    from gfirst_amd.geometry import C1
    g = C1
    storage = STORAGE[args.storage or "f32"]
    log("geometry c1: nsinogs=%d ndata=%d nvox=%d" % (g.nsinogs, g.ndata, g.nvox))
    ctx = ga.Context(0)
    t = time.perf_counter()
    d2, norm = make_inputs(ctx, g, args.counts, seed=1)
    log("synthetic study ready in %.1fs (%.3g counts)" % (time.perf_counter() - t, d2.sum()))
    total = args.warmup + args.steps
    niter = (total + args.nsub - 1) // args.nsub + 1
    p = ga.recon_params(mode=1, niter=niter, nsub=args.nsub, storage=storage)
    r = ctx.recon(g, p, d2, None, norm)
    pb = r.get(ga.OUT_PROJ_BLANK)
    tune = ctx.tuning(g, args.nsub)
    log("autotuned launch shapes: %s" % tune)
    log("reconstruction resident on device (%.2f GB)" % (ctx.bytes_in_use() / 1e9))
    for _ in range(args.warmup):
        r.step()
    r.sync()
    log("warmup %d steps done" % args.warmup)
    r.profile(True)
    r.phase_times()
    r.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r.step()
    r.sync()
    t1 = time.perf_counter()
    ph = r.phase_times()
    r.profile(False)
    ll = r.loglik()
    dt = t1 - t0
    ms_step = 1e3 * dt / args.steps
    log("timed %d steps: %.3f ms/step; phases %s" % (
        args.steps, ms_step, {k: round(v[0] / max(v[1], 1), 4) for k, v in ph.items()}))
    sb = 4 if storage == ga.STORE_F32 else 2
    ab = alg_bytes(g, args.nsub, sb, has_bg=False)
    af = alg_flops(g, args.nsub)
    # dominant kernel: the projector with the larger device time
    dom = max(("forward", "back"), key=lambda k: ph[k][0])
    avg_ms = ph[dom][0] / max(ph[dom][1], 1)
    achieved = ab[dom] / (avg_ms * 1e-3) / 1e9
    flops = af[dom] / (avg_ms * 1e-3) / 1e12
    step_flops = (af["forward"] + af["back"]) / (ms_step * 1e-3) / 1e12
    b_step = step_alg_bytes(g, args.nsub, sb, has_bg=False)
    res = {
        "metric": "Inveon-size OSEM subset-iters/s + HBM GB/s; p50/p99 recon latency under load",
        "value": round(args.steps / dt, 3),
        "unit": "subset-iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if storage == ga.STORE_F32 else args.storage + "-storage/f32",
        "data": "synthetic (seeded mouse phantom, Poisson 5e7 counts, norm U[0.8,1.2])",
        "config": {"workload": "c1 Inveon-scale doubles OSEM: image 128x128x159, sinogram "
                               "128x160x4319, span 3, MRD 79, %d subsets" % args.nsub,
                   "subsets": args.nsub, "storage": args.storage or "f32"},
        "hbm_gbps_alg": round(b_step / (ms_step * 1e-3) / 1e9, 2),
        "hbm_gbps_alg_formula": "B_alg = L_s*(b_y+b_bg+b_n+8) + 16*NVOX = %d bytes per "
                                "subset-iteration, over ms_per_step" % b_step,
        "fp32_tflops_step": round(step_flops, 2),
        "phase_ms": {k: round(v[0] / max(v[1], 1), 4) for k, v in ph.items() if v[1]},
        # The projectors are interpolation gathers: the binding roofline is the
        # FP32 vector path fed by the CU's vector-memory units (F_alg/t against
        # 157.3 TFLOP/s); the HBM fraction of the same launch is reported beside.
        "roofline": {"kernel": "k_zsum+k_rotate_tile" if dom == "forward" else "k_back_direct",
                     "bound": "fp32-valu (vector-memory fed)",
                     "achieved": round(flops, 2), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(flops / PEAK_FP32_TFLOPS, 4),
                     "flops_per_launch": af[dom], "avg_launch_ms": round(avg_ms, 4),
                     "traffic": measured_traffic(dom, tune["back_zb"], storage, tune["back_lds"],
                                                 tune["back_bw"]),
                     "hbm": {"achieved": round(achieved, 2), "peak": PEAK_HBM_GBS,
                             "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 5),
                             "alg_bytes_per_launch": ab[dom]},
                     "binding": measured_pipes(dom, tune) if storage == 0 else None},
        "loglik_last_pass": ll,
        "tiles": tune,
    }
    r.close()
    ctx.close()
    if args.serve_s > 0:
        # the metric's latency half: a served stream of the c4 mix beside the headline
        sctx = ga.Context(0, hbm_budget=int(args.budget_gb * 1e9))
        t = time.perf_counter()
        cls = c4_setup(sctx, args.seed)
        log("serve: classes ready in %.1fs" % (time.perf_counter() - t))
        res["serving"] = served_stream(sctx, cls, STORAGE["f16"], "f16", args.serve_s, args.rate,
                                       args.max_active, args.seed, args.budget_gb)
        sctx.close()
    if not args.no_cpu and args.cpu_steps > 0:
        log("timing the CPU oracle on %d subset-iterations ..." % args.cpu_steps)
        res["cpu_baseline"] = cpu_baseline(g, d2, norm, pb, args.nsub, args.cpu_steps)
        log("cpu: %.3f subset-iters/s on %d threads" % (res["cpu_baseline"]["value"],
                                                       res["cpu_baseline"]["cores"]))
    return res



def c2_study(ctx, counts, nevents, seed):
    """The joint study at Inveon size: doubles from phantom channel 0 and
    triples from channel 1 (counts/50), both Poisson on the device from the
    blurred phantoms' projections, and nevents triple events binned to VLOR
    records on the host."""
    from gfirst_amd import listmode, synth
    from gfirst_amd.geometry import C1
    g = C1

    def expected(img):
        d = ctx.to_device(img)
        b = ctx.array(g.nvox)
        ctx.blur_xy(g, d, 1.5, out=b)
        p = ctx.array(g.ndata)
        ctx.forward_project(g, b, out=p)
        ctx.sync()
        return p, float(p.download().sum(dtype=np.float64))

    chA = synth.phantom_mouse(g, seed=seed, channel=0)
    chB = synth.phantom_mouse(g, seed=seed, channel=1)
    yD, sD = expected(chA)
    yT, sT = expected(chB)
    d2 = ctx.array(2 * g.ndata).zero()
    ctx.poisson(yD, counts / sD, seed + 1, out=d2)
    t2 = ctx.array(2 * g.ndata).zero()
    ctx.poisson(yT, counts / 50.0 / sT, seed + 2, out=t2)
    norm = synth.norm_uniform(g.ndata, seed=seed + 3)
    hits = listmode.triple_events(g, chB, nevents, seed=seed + 5)
    rec = listmode.bin_hits(g, hits, 80.5)
    rec = rec[(rec[:, 0] > 0) & (rec[:, 1] > 0)]
    ctx.sync()
    return g, d2, t2, norm, rec


def cpu_baseline_c2(g, d2, t2, norm, rec, pb, nsub):
    """The oracle on a bounded sample of the joint study: one doubles pass
    and one triples pass (subset 0 of the first iteration, as in the
    reference's loop GFIRST.cu:324-527: blur, forward, VLOR assignment for
    the triples, ratio, back, update, then JBF and blend). The set-up (mask
    from the device's PROJ_BLANK, VLOR precount and blank, the triples
    projection cache) is untimed."""
    from oracle import oracle as orc
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
    orc.set_threads(threads)
    nd, nv = g.ndata, g.nvox
    res = np.float32(g.res)
    cal = orc.calib(g, pb, norm)
    yD, bD = res * d2[:nd], res * d2[nd:]
    yT, bT = res * t2[:nd], res * t2[nd:]
    if rec is not None:
        blankt = orc.vlor_blankt(g, yT, orc.vlor_precount(g, rec))
        ptc = orc.forward(g, orc.blur_xy(g, orc.cylinder(g, 1.0), 1.5), 0, 1)
    xD = orc.cylinder(g, 1.0)
    xT = orc.cylinder(g, 1.0)
    dv = np.zeros(nd, np.float32)
    sub = np.zeros((g.nsinogs, g.nang), bool)
    sub[:, 0::nsub] = True
    sub = np.repeat(sub.reshape(-1), g.nrad)
    t0 = time.perf_counter()
    xb = orc.blur_xy(g, xD, 1.5)
    proj = orc.forward(g, xb, 0, nsub)
    corr = orc.ratio(g, proj, yD, bD, 0, nsub)
    num, den = orc.back(g, corr, cal, 0, nsub)
    xD = orc.update(g, xD, num, den, 1.5)
    t1 = time.perf_counter()
    xb = orc.blur_xy(g, xT, 1.5)
    proj = orc.forward(g, xb, 0, nsub)
    if rec is not None:
        ptc[sub] = proj[sub]
        dv = orc.vlor_assign(g, rec, blankt, ptc, 0, nsub, dv)
        corr = orc.ratio(g, proj, dv * res, bT, 0, nsub)
    else:
        corr = orc.ratio(g, proj, yT, bT, 0, nsub)
    num, den = orc.back(g, corr, cal, 0, nsub)
    xT = orc.update(g, xT, num, den, 1.5)
    filt = orc.jbf(g, xT, xD, orc.vmax(xD))
    xT = orc.blend(g, xT, filt)
    t2_ = time.perf_counter()
    dt = t2_ - t0
    return dict(value=2 / dt, unit="subset-iters/s", cores=orc.get_threads(), kind="port",
                sample="one doubles pass (%.1f s) and one triples pass with %sJBF and blend "
                       "(%.1f s), subset 0 of %d of the same joint study, oracle FP32 "
                       "restatement with OpenMP" % (t1 - t0, "VLOR assignment, "
                                                    if rec is not None else "", t2_ - t1, nsub),
                seconds=dt)


def run_c2(args):
    import gfirst_amd as ga
    storage = STORAGE[args.storage or "f32"]
    nsub = args.nsub
    ctx = ga.Context(0)
    t = time.perf_counter()
    g, d2, t2, norm, rec = c2_study(ctx, args.counts, args.events, args.seed)
    log("c2: joint study ready in %.1fs (%d VLOR records)" % (time.perf_counter() - t, len(rec)))
    per_iter = 2 * nsub  # a doubles and a triples pass per subset
    warm = -(-args.warmup // per_iter) * per_iter
    steps = -(-args.steps // per_iter) * per_iter
    niter = (warm + steps) // per_iter
    mode = args.c2_mode
    p = ga.recon_params(mode=mode, niter=niter, nsub=nsub, storage=storage)
    r = ctx.recon(g, p, d2, t2, norm, rec if mode == 3 else None)
    tune = ctx.tuning(g, nsub)
    log("c2: %d iterations, %d passes each; resident %.2f GB" % (
        niter, per_iter, ctx.bytes_in_use() / 1e9))
    for _ in range(warm):
        r.step()
    r.sync()
    r.profile(True)
    r.phase_times()
    r.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        r.step()
    r.sync()
    dt = time.perf_counter() - t0
    ph = r.phase_times()
    r.profile(False)
    cur = r.cursor()
    ms_step = 1e3 * dt / steps
    log("c2: timed %d passes: %.3f ms/pass; phases %s" % (
        steps, ms_step, {k: round(v[0], 2) for k, v in ph.items()}))
    sb = 4 if storage == ga.STORE_F32 else 2
    ab = alg_bytes(g, nsub, sb, has_bg=False)
    af = alg_flops(g, nsub)
    dom = max(("forward", "back"), key=lambda k: ph[k][0])
    avg_ms = ph[dom][0] / max(ph[dom][1], 1)
    achieved = ab[dom] / (avg_ms * 1e-3) / 1e9
    flops = af[dom] / (avg_ms * 1e-3) / 1e12
    res = {
        "metric": "joint doubles+triples OSEM subset-iters/s" + (
            "" if mode == 3 else " (mode 2, no VLOR)"),
        "value": round(steps / dt, 3),
        "unit": "subset-iters/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warm,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if storage == ga.STORE_F32 else args.storage + "-storage/f32",
        "data": "synthetic (seeded mouse phantom channels, device Poisson %.3g doubles / %.3g "
                "triples counts, %d triple events binned to %d VLOR records, norm U[0.8,1.2])" % (
                    args.counts, args.counts / 50.0, args.events, len(rec)),
        "config": {"workload": ("c2 joint doubles+triples OSEM (mode 3) with VLOR assignment: "
                                if mode == 3 else
                                "c2 doubles+triples OSEM (mode 2, triples from the sinogram): ") +
                               "image 128x128x159, sinogram 128x160x4319, span 3, MRD 79, "
                               "%d subsets; a step is one pass (one subset of one type)" % nsub,
                   "subsets": nsub, "storage": args.storage or "f32", "mode": mode},
        "phase_ms_per_iteration": {k: round(v[0] * per_iter / steps, 3) for k, v in ph.items()
                                   if v[1]},
        # phases of the two chains overlap on the device: a phase's event
        # time includes the other chain's kernels running beside it
        "roofline": {"kernel": "k_zsum+k_rotate_tile" if dom == "forward" else "k_back_direct",
                     "bound": "fp32-valu (vector-memory fed)",
                     "achieved": round(flops, 2), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(flops / PEAK_FP32_TFLOPS, 4), "traffic": None,
                     "flops_per_launch": af[dom], "avg_launch_ms": round(avg_ms, 4),
                     "hbm": {"achieved": round(achieved, 2), "peak": PEAK_HBM_GBS,
                             "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 5),
                             "alg_bytes_per_launch": ab[dom]}},
        "cursor_after": cur,
        "tiles": tune,
    }
    pb = r.get(ga.OUT_PROJ_BLANK)
    r.close()
    if not args.no_cpu:
        log("timing the oracle on one doubles and one triples pass ...")
        res["cpu_baseline"] = cpu_baseline_c2(g, d2.download(), t2.download(), norm,
                                              rec if mode == 3 else None, pb, nsub)
        log("cpu: %.3f subset-iters/s (%s)" % (res["cpu_baseline"]["value"],
                                              res["cpu_baseline"]["sample"]))
    ctx.close()
    return res


def c3_inputs(ctx, g, seed):
    """Expected sinograms of the two phantom channels (PSF-blurred, projected
    on the device) and their totals."""
    import gfirst_amd as ga  # noqa: F401
    from gfirst_amd import synth
    P, tot = [], []
    for ch in (0, 1):
        img = ctx.to_device(synth.phantom_mouse(g, seed=seed, channel=ch))
        blur = ctx.array(g.nvox)
        ctx.blur_xy(g, img, 1.5, out=blur)
        p = ctx.array(g.ndata)
        ctx.forward_project(g, blur, out=p)
        ctx.sync()
        tot.append(float(p.download().sum(dtype=np.float64)))
        P.append(p)
    return P, tot


def frame_scales(f, tot):
    """Scales of the two channel projections giving the frame's activity mix
    and total expected counts."""
    mix = f["cp"] * tot[0] + f["ct"] * tot[1]
    return f["counts"] * f["cp"] / mix, f["counts"] * f["ct"] / mix


def c3_fill(ctx, bt, frames, P, tot, niter, ybuf, seed):
    for k, f in enumerate(frames):
        a, b = frame_scales(f, tot)
        ctx.poisson(P[0], a, seed * 1000003 + k, out=ybuf, mean2=P[1], scale2=b)
        ctx.sync()  # the batch copies on its own stream
        bt.add(niter, f["nsub"], ybuf)


def equiv_passes(g, nsub, passes):
    """Subset passes expressed in passes of a 16-subset split (work scales
    with the angles per subset)."""
    return passes * ((g.nang + nsub - 1) // nsub) / ((g.nang + 15) // 16)


def run_c3(args):
    import gfirst_amd as ga
    from gfirst_amd import synth
    from gfirst_amd.geometry import C1
    g = C1
    sname = args.storage or "f16"
    storage = STORAGE[sname]
    frames = synth.dynamic_study(args.frames, seed=args.seed)
    niter = args.c3_niter
    ctx = ga.Context(0)
    t = time.perf_counter()
    P, tot = c3_inputs(ctx, g, args.seed)
    norm = synth.norm_uniform(g.ndata, seed=args.seed + 2)
    ybuf = ctx.array(g.ndata)
    log("c3: %d frames, subsets %s, counts %.3g..%.3g, channel projections in %.1fs" % (
        len(frames), sorted(set(f["nsub"] for f in frames)), min(f["counts"] for f in frames),
        max(f["counts"] for f in frames), time.perf_counter() - t))

    def make_batch(lanes, sub):
        bp = ga.batch_params(storage=storage, max_frames=len(sub), min_nsub=min(f["nsub"] for f in sub),
                             lanes=lanes, group=args.group)
        bt = ctx.batch(g, bp, norm)
        c3_fill(ctx, bt, sub, P, tot, niter, ybuf, args.seed)
        return bt

    # lanes: measured on a short run of the first frames unless given
    lanes = args.lanes
    probe = {}
    if lanes <= 0:
        sub = frames[:min(16, len(frames))]
        for L in (1, 2, 3, 4):
            bt = make_batch(L, sub)
            bt.step()
            bt.sync()
            t0 = time.perf_counter()
            n = 0
            for _ in range(6):
                n += sum(1 for k in range(len(sub)) if not bt.cursor(k)["done"])
                bt.step()
            bt.sync()
            probe[L] = n / (time.perf_counter() - t0)
            bt.close()
        lanes = max(probe, key=probe.get)
        log("lanes probe (passes/s): %s -> %d" % ({k: round(v, 1) for k, v in probe.items()}, lanes))
    bt = make_batch(lanes, frames)
    log("batch resident: %d frames, %.2f GB in use" % (len(frames), ctx.bytes_in_use() / 1e9))
    ctx.sync()
    t0 = time.perf_counter()
    nsteps = 0
    while bt.step() > 0:
        nsteps += 1
    nsteps += 1
    bt.sync()
    dt = time.perf_counter() - t0
    passes = sum(bt.cursor(k)["passes"] for k in range(len(frames)))
    eq = sum(equiv_passes(g, f["nsub"], niter * f["nsub"]) for f in frames)
    st = bt.stats()
    lls = [bt.loglik(k) for k in range(len(frames))]
    bt.close()
    log("batched: %d frames, %d passes in %.3fs over %d graph launches (%d builds): "
        "%.1f passes/s, %.1f S=16-equivalent subset-iters/s" % (
            len(frames), passes, dt, st["launches"], st["builds"], passes / dt, eq / dt))
    # plain single-study path: the first frames as separate sessions, one at a time
    pl_pass = pl_eq = pl_t = 0.0
    nplain = min(args.plain_frames, len(frames))
    for k in range(nplain):
        f = frames[k]
        a, b = frame_scales(f, tot)
        ctx.poisson(P[0], a, args.seed * 1000003 + k, out=ybuf, mean2=P[1], scale2=b)
        ctx.sync()
        d2 = synth.as_file_pair(ybuf.download())
        r = ctx.recon(g, ga.recon_params(mode=1, niter=niter, nsub=f["nsub"], storage=storage),
                      d2, None, norm)
        r.sync()
        t1 = time.perf_counter()
        r.run()
        pl_t += time.perf_counter() - t1
        if r.loglik() != lls[k]:
            raise RuntimeError("frame %d: batched and single-session results differ" % k)
        r.close()
        pl_pass += niter * f["nsub"]
        pl_eq += equiv_passes(g, f["nsub"], niter * f["nsub"])
    res = {
        "metric": "Inveon-size OSEM subset-iters/s (batched dynamic study)",
        "value": round(eq / dt, 3),
        "unit": "subset-iters/s (S=16-equivalent, all frames)",
        "n_gpus": 1,
        "steps": nsteps,
        "warmup": 0,
        "ms_per_step": round(1e3 * dt / nsteps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "%s-storage/f32" % sname if sname != "f32" else "f32",
        "data": "synthetic (two seeded mouse-phantom channels x two-compartment curves, "
                "device Poisson, norm U[0.8,1.2])",
        "config": {"workload": "c3 batched dynamic study: %d frames of the Inveon-scale "
                               "geometry, subsets {8,16,24}, %d iterations, counts "
                               "log-U[1e5,1e7]" % (len(frames), niter),
                   "frames": len(frames), "lanes": lanes, "storage": sname,
                   "group": args.group or 3},
        "passes": passes,
        "passes_per_s": round(passes / dt, 2),
        "seconds": round(dt, 3),
        "graph": st,
        "lanes_probe": {str(k): round(v, 2) for k, v in probe.items()},
        "plain_sessions": {"frames": nplain, "seconds": round(pl_t, 3),
                           "subset_iters_per_s_equiv": round(pl_eq / pl_t, 3) if pl_t else None,
                           "same_loglik_as_batched": True},
        "batched_over_plain": round((eq / dt) / (pl_eq / pl_t), 3) if pl_t else None,
    }
    if args.error_report:
        res["error_vs_fp64"] = c3_error_report(ctx, g, frames[0], P, tot, norm, ybuf, args.seed)
    ctx.close()
    return res


def c3_error_report(ctx, g, f, P, tot, norm, ybuf, seed):
    """Image error of one frame (one iteration) in each storage format against
    the FP64 oracle on the same data."""
    import gfirst_amd as ga
    from gfirst_amd import synth
    from oracle import oracle as orc
    a, b = frame_scales(f, tot)
    ctx.poisson(P[0], a, seed * 1000003, out=ybuf, mean2=P[1], scale2=b)
    ctx.sync()
    d2 = synth.as_file_pair(ybuf.download())
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    orc.set_threads(threads)
    t0 = time.perf_counter()
    x64 = oracle_fp64_osem1(orc, g, d2, norm, f["nsub"])
    log("FP64 oracle: 1 iteration S=%d in %.1fs" % (f["nsub"], time.perf_counter() - t0))
    x64 = np.asarray(x64, np.float64)
    out = {"frame": 0, "nsub": f["nsub"], "iterations": 1, "counts": round(f["counts"]),
           "max_count": float(d2[:g.ndata].max())}
    for name, st in STORAGE.items():
        r = ctx.recon(g, ga.recon_params(mode=1, niter=1, nsub=f["nsub"], storage=st), d2, None,
                      norm)
        r.run()
        x = r.get(ga.OUT_XD).astype(np.float64)
        r.close()
        out[name] = float(np.linalg.norm(x - x64) / np.linalg.norm(x64))
        log("rel L2 vs FP64 oracle, %s storage: %.3g" % (name, out[name]))
    return out


def oracle_fp64_osem1(orc, g, d2, norm, S, psf=1.5):
    """One doubles iteration (S subsets) of the oracle in FP64, pass by pass
    with progress lines: mask from the projected cylinder and the norm
    (GFIRST.cu:155-181, 453-458), then blur, forward, ratio, back, update per
    subset (GFIRST.cu:324-502)."""
    nd = g.ndata
    cyl = orc.cylinder(g, 1.0, fp64=True)
    pb = np.zeros(nd, np.float64)
    for s in range(16):
        orc.forward(g, cyl, s, 16, out=pb, fp64=True)
    log("FP64 oracle: PROJ_BLANK done")
    cal = orc.calib(g, pb, norm, fp64=True)
    del pb
    y = float(g.res) * d2[:nd].astype(np.float64)
    x = cyl
    for s in range(S):
        xb = orc.blur_xy(g, x, psf, fp64=True)
        proj = orc.forward(g, xb, s, S, fp64=True)
        corr = orc.ratio(g, proj, y, None, s, S, fp64=True)
        del proj
        num, den = orc.back(g, corr, cal, s, S, fp64=True)
        del corr
        x = orc.update(g, x, num, den, psf, fp64=True)
        log("FP64 oracle: subset %d/%d" % (s + 1, S))
    return x


C4_CLASSES = (("c0", 0.6), ("c1", 0.3), ("c2", 0.1))


def c4_setup(ctx, seed):
    """Per request class: geometry, expected sinograms on the device, the
    normalization (device) and, for the joint class, VLOR records."""
    import gfirst_amd as ga  # noqa: F401
    from gfirst_amd import listmode, synth
    from gfirst_amd.geometry import C0, C1

    def project(g, img):
        d = ctx.to_device(img)
        b = ctx.array(g.nvox)
        ctx.blur_xy(g, d, 1.5, out=b)
        p = ctx.array(g.ndata)
        ctx.forward_project(g, b, out=p)
        ctx.sync()
        return p, float(p.download().sum(dtype=np.float64))

    cls = {}
    y0, s0 = project(C0, synth.phantom_c0(C0, seed=seed))
    cls["c0"] = dict(g=C0, nsub=4, mode=1, y=y0, ysum=s0,
                     norm=ctx.to_device(np.ones(C0.ndata, np.float32)))
    n1 = ctx.to_device(synth.norm_uniform(C1.ndata, seed=seed + 2))
    y1, s1 = project(C1, synth.phantom_mouse(C1, seed=seed, channel=0))
    cls["c1"] = dict(g=C1, nsub=16, mode=1, y=y1, ysum=s1, norm=n1)
    chB = synth.phantom_mouse(C1, seed=seed, channel=1)
    yT, sT = project(C1, chB)
    t0 = time.perf_counter()
    hits = listmode.triple_events(C1, chB, 2000000, seed=seed + 5)
    rec = listmode.bin_hits(C1, hits, 80.5)
    rec = rec[(rec[:, 0] > 0) & (rec[:, 1] > 0)]
    log("c4: %d VLOR records binned in %.1fs" % (len(rec), time.perf_counter() - t0))
    cls["c2"] = dict(g=C1, nsub=16, mode=3, y=y1, ysum=s1, norm=n1, yT=yT, yTsum=sT, rec=rec)
    return cls


def c4_inputs(ctx, c, counts, seed):
    g = c["g"]
    d2 = ctx.array(2 * g.ndata).zero()
    ctx.poisson(c["y"], counts / c["ysum"], seed, out=d2)
    t2 = None
    if c["mode"] >= 2:
        t2 = ctx.array(2 * g.ndata).zero()
        ctx.poisson(c["yT"], counts / 50.0 / c["yTsum"], seed + 1, out=t2)
    ctx.sync()
    return d2, t2


def served_stream(ctx, cls, storage, sname, duration, rate, max_active, seed, budget_gb,
                  error_report=False, inflight=2, admitters=4):
    """Plays the served stream: Poisson arrivals at `rate`/s for `duration` s
    of the c4 mix (classes from c4_setup), one untimed warm-up request per
    class, then every request waited for; latency = submission to the image
    on the host. The first request of each class is compared with the same
    request run alone as a single session (bit for bit)."""
    import gfirst_amd as ga
    rng = np.random.default_rng(seed + 11)
    arrivals = []
    ta = 0.0
    while True:
        ta += rng.exponential(1.0 / rate)
        if ta > duration:
            break
        name = rng.choice([c for c, _ in C4_CLASSES], p=[w for _, w in C4_CLASSES])
        counts = float(np.exp(rng.uniform(np.log(1e5), np.log(5e7))))
        niter = int(rng.integers(2, 5))
        arrivals.append((ta, str(name), counts, niter))
    log("serve: %d arrivals over %.0fs (rate %.1f/s)" % (len(arrivals), duration, rate))
    # untimed warm-up: one request of each class (tile shapes, code paths)
    srv = ctx.server(max_active=max_active, inflight=inflight, admitters=admitters)
    for k, name in enumerate(cls):
        c = cls[name]
        d2, t2 = c4_inputs(ctx, c, 1e6, 10 ** 6 + k)
        p = ga.recon_params(mode=c["mode"], niter=1, nsub=c["nsub"], storage=storage)
        rid = srv.submit(c["g"], p, d2, t2, c["norm"], c.get("rec"))
        srv.wait(rid)
        srv.release(rid)
    ids = []
    t0 = time.perf_counter()
    for k, (ta, name, counts, niter) in enumerate(arrivals):
        dt = t0 + ta - time.perf_counter()
        if dt > 0:
            time.sleep(dt)
        c = cls[name]
        d2, t2 = c4_inputs(ctx, c, counts, seed * 7919 + k)
        p = ga.recon_params(mode=c["mode"], niter=niter, nsub=c["nsub"], storage=storage)
        ids.append((srv.submit(c["g"], p, d2, t2, c["norm"], c.get("rec")), name, niter, counts))
        if k % 40 == 0:
            st = srv.stats()
            log("serve: t=%.1fs submitted %d, active %d, pending %d, done %d" % (
                time.perf_counter() - t0, k + 1, st["active"], st["pending"], st["done"]))
    for rid, _, _, _ in ids:
        srv.wait(rid)
    t_end = time.perf_counter()
    infos = [(srv.info(rid), name, niter, counts) for rid, name, niter, counts in ids]
    failed = [i for i, _, _, _ in infos if i["state"] != ga.REQ_DONE]
    lat, adm, run = {}, {}, {}
    passes = eq = 0.0
    for i, name, niter, counts in infos:
        if i["state"] != ga.REQ_DONE:
            continue
        lat.setdefault(name, []).append(i["t_done"] - i["t_submit"])
        # admitted = session created (after any wait for a slot or memory)
        adm.setdefault(name, []).append(i["t_admit"] - i["t_submit"])
        run.setdefault(name, []).append(i["t_done"] - i["t_admit"])
        c = cls[name]
        npass = niter * c["nsub"] * (2 if c["mode"] >= 2 else 1)
        passes += npass
        eq += equiv_passes(c["g"], c["nsub"], npass) * (c["g"].ndata / cls["c1"]["g"].ndata)
    allat = np.concatenate([np.asarray(v) for v in lat.values()])
    span = max(i["t_done"] for i, _, _, _ in infos) - min(i["t_submit"] for i, _, _, _ in infos)
    q = lambda a, x: round(float(np.percentile(a, x)) * 1e3, 2)  # noqa: E731
    # bit-exact check of the first request of each class against a lone session
    same = {}
    for name in cls:
        k = next((j for j, (_, nm, _, _) in enumerate(ids) if nm == name), None)
        if k is None:
            continue
        rid, _, niter, counts = ids[k]
        c = cls[name]
        d2, t2 = c4_inputs(ctx, c, counts, seed * 7919 + k)
        p = ga.recon_params(mode=c["mode"], niter=niter, nsub=c["nsub"], storage=storage)
        r = ctx.recon(c["g"], p, d2, t2, c["norm"], c.get("rec"))
        r.run()
        xD, _ = srv.result(rid)
        same[name] = bool(np.array_equal(r.get(ga.OUT_XD), xD))
        r.close()
    err = c4_error_report(ctx, srv, cls, storage, seed) if error_report else None
    srv.close()
    out = {
        "workload": "Poisson arrivals %.1f/s for %.0fs, mix 60%% c0 / 30%% c1 / 10%% c2 "
                    "(mode 3, 2e6 VLOR events), counts log-U[1e5,5e7], 2-4 iterations, "
                    "%s storage, %.0f GB budget" % (rate, duration, sname, budget_gb),
        "p50_ms": q(allat, 50), "p99_ms": q(allat, 99), "max_ms": q(allat, 100),
        "per_class": {k: {"n": len(v), "p50_ms": q(v, 50), "p99_ms": q(v, 99),
                          "admit_p50_ms": q(adm[k], 50), "run_p50_ms": q(run[k], 50)}
                      for k, v in lat.items()},
        "requests": len(ids), "failed": len(failed),
        "subset_iters_per_s": round(passes / span, 2),
        "c1_equiv_subset_iters_per_s": round(eq / span, 2),
        "makespan_s": round(span, 2),
        "drain_s": round(t_end - t0 - duration, 2),
        "alone_equals_mix": same,
        "max_active": max_active,
        "inflight": inflight,
        "admitters": admitters,
        "hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "0") or 0),
    }
    if err is not None:
        out["error_vs_fp64_oracle"] = err
    log("serve: p50 %.2f ms, p99 %.2f ms, %d requests, %d failed, alone == mix %s" % (
        out["p50_ms"], out["p99_ms"], out["requests"], out["failed"], same))
    return out


def run_c4(args):
    import gfirst_amd as ga  # noqa: F401
    sname = args.storage or "f16"
    storage = STORAGE[sname]
    ctx = ga.Context(0, hbm_budget=int(args.budget_gb * 1e9))
    t = time.perf_counter()
    cls = c4_setup(ctx, args.seed)
    log("c4: classes ready in %.1fs" % (time.perf_counter() - t))
    sv = served_stream(ctx, cls, storage, sname, args.duration, args.rate, args.max_active,
                       args.seed, args.budget_gb, args.error_report, args.inflight,
                       args.admitters)
    res = {
        "metric": "p50/p99 recon latency under load",
        "value": sv["p50_ms"],
        "unit": "ms (p50 request latency)",
        "n_gpus": 1,
        "steps": sv["requests"],
        "warmup": len(cls),
        "ms_per_step": round(1e3 * sv["makespan_s"] / max(sv["requests"], 1), 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "%s-storage/f32" % sname,
        "data": "synthetic (seeded phantoms, device Poisson, Poisson arrivals)",
        "config": {"workload": "c4 served stream: " + sv["workload"],
                   "max_active": args.max_active, "storage": sname,
                   "hw_queues": sv["hw_queues"]},
        "latency_ms": {"p50": sv["p50_ms"], "p99": sv["p99_ms"], "max": sv["max_ms"],
                       "per_class": {k: {"n": v["n"], "p50": v["p50_ms"], "p99": v["p99_ms"]}
                                     for k, v in sv["per_class"].items()}},
        "requests": sv["requests"], "failed": sv["failed"],
        "subset_iters_per_s": sv["subset_iters_per_s"],
        "c1_equiv_subset_iters_per_s": sv["c1_equiv_subset_iters_per_s"],
        "makespan_s": sv["makespan_s"],
        "drain_s": sv["drain_s"],
        "alone_equals_mix": sv["alone_equals_mix"],
    }
    if "error_vs_fp64_oracle" in sv:
        res["error_vs_fp64_oracle"] = sv["error_vs_fp64_oracle"]
    ctx.close()
    return res


def c4_error_report(ctx, srv, cls, storage, seed):
    """Per request class, one served request (1 iteration, 1e7 counts, the
    stream's storage format) against the FP64 oracle's reconstruction of the
    same data: relative L2 image error of the doubles (and triples) image,
    beside the FP32 oracle's own distance from the FP64 oracle."""
    import gfirst_amd as ga
    from oracle import oracle as orc
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
    orc.set_threads(threads)
    out = {}
    for k, name in enumerate(cls):
        c = cls[name]
        d2, t2 = c4_inputs(ctx, c, 1e7, 10 ** 7 + k)
        p = ga.recon_params(mode=c["mode"], niter=1, nsub=c["nsub"], storage=storage)
        rid = srv.submit(c["g"], p, d2, t2, c["norm"], c.get("rec"))
        srv.wait(rid)
        xD, xT = srv.result(rid)
        srv.release(rid)
        log("c4 error report: %s served; FP64 oracle (mode %d, S=%d, 1 iteration) ..." % (
            name, c["mode"], c["nsub"]))
        t0 = time.perf_counter()
        oD, oT, _, _ = orc.osem(c["g"], orc.params(mode=c["mode"], niter=1, nsub=c["nsub"]),
                                d2.download(), None if t2 is None else t2.download(),
                                c["norm"].download(), c.get("rec"), fp64=True)
        e = {"mode": c["mode"], "subsets": c["nsub"], "iterations": 1, "counts": 1e7,
             "doubles": float(np.linalg.norm(xD - oD) / np.linalg.norm(oD)),
             "oracle_s": round(time.perf_counter() - t0, 1)}
        if xT is not None:
            e["triples"] = float(np.linalg.norm(xT - oT) / np.linalg.norm(oT))
        # the reference's own FP32 arithmetic (the FP32 oracle) on the same data
        fD, fT, _, _ = orc.osem(c["g"], orc.params(mode=c["mode"], niter=1, nsub=c["nsub"]),
                                d2.download(), None if t2 is None else t2.download(),
                                c["norm"].download(), c.get("rec"))
        e["fp32_oracle_doubles"] = float(np.linalg.norm(fD - oD) / np.linalg.norm(oD))
        if xT is not None:
            e["fp32_oracle_triples"] = float(np.linalg.norm(fT - oT) / np.linalg.norm(oT))
        out[name] = e
        log("c4 error report: %s %s" % (name, e))
    return out


def copy_probe(nbytes=1 << 30, reps=5):
    """A plain device copy of known size, for scaling FETCH_SIZE/WRITE_SIZE."""
    import torch
    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    return {"copy_bytes": nbytes, "reps": reps}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="c1", choices=["c0", "c1", "c2", "c3", "c4"])
    ap.add_argument("--steps", type=int, default=48)
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--nsub", type=int, default=16)
    ap.add_argument("--counts", type=float, default=5e7)
    ap.add_argument("--storage", default=None, choices=["f32", "f16", "bf16"])
    ap.add_argument("--cpu-steps", type=int, default=12)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--frames", type=int, default=64)
    ap.add_argument("--c3-niter", type=int, default=4)
    ap.add_argument("--lanes", type=int, default=0, help="0: measure 1..4")
    ap.add_argument("--group", type=int, default=0, choices=[0, 1, 2, 3],
                    help="c3: frames per shared backprojection (1: none; 0: 3)")
    ap.add_argument("--plain-frames", type=int, default=6)
    ap.add_argument("--error-report", action="store_true")
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--copy-probe", action="store_true")
    ap.add_argument("--duration", type=float, default=300.0, help="c4: stream length, s")
    ap.add_argument("--serve-s", type=float, default=60.0,
                    help="c1: length of the served stream reported beside the headline "
                         "(0: none)")
    ap.add_argument("--rate", type=float, default=4.0)
    ap.add_argument("--inflight", type=int, default=2,
                    help="served stream: passes of a session queued ahead of the device")
    ap.add_argument("--admitters", type=int, default=4,
                    help="served stream: threads creating sessions")
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--max-active", type=int, default=16)
    ap.add_argument("--hw-queues", type=int, default=16, help="c4: GPU_MAX_HW_QUEUES (<= 32)")
    ap.add_argument("--events", type=int, default=2000000, help="c2: triple events")
    ap.add_argument("--c2-mode", type=int, default=3, choices=[2, 3],
                    help="c2: 3 = triples from VLOR records, 2 = from the triples sinogram")
    args = ap.parse_args()
    if args.workload in ("c1", "c2", "c4"):
        # One hardware queue per stream in use: per admitted session of the
        # served stream, and the two chains of a joint session (HIP's default
        # is 4 per process, and streams sharing a queue run in FIFO order);
        # read by the runtime when the first context is created.
        os.environ["GPU_MAX_HW_QUEUES"] = str(max(1, min(args.hw_queues, 32)))
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        res = copy_probe()
    elif args.workload == "c0":
        res = run_c0(args)
    elif args.workload == "c2":
        res = run_c2(args)
    elif args.workload == "c3":
        res = run_c3(args)
    elif args.workload == "c4":
        res = run_c4(args)
    else:
        res = run_c1(args)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
