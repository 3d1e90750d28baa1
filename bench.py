"""Benchmark: Inveon-scale doubles OSEM on one MI355X.

Workload: 128x128x159 image (0.776 mm), 128 radial x 160 angles x 4319
oblique sinograms (80 rings, span 3, max ring difference 79), 16 subsets,
FP32, synthetic mouse-like phantom, 5e7 Poisson counts, normalization
U[0.8, 1.2]. A step is one subset-iteration: PSF blur, forward projection +
ratio (+ FP64 log-likelihood), backprojection and image update of one
subset, all on the device with the data resident in HBM.

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


def alg_flops(g, nsub):
    """FP32 flops of the two projectors per subset-iteration: 22 per
    trilinear sample (7 lerps x 3 + 1 add), 10 per bilinear voxel-view fetch
    of the ratio image (3 lerps x 3 + 1 add)."""
    L = g.ndata // nsub
    nviews = g.nseg * (g.nang // nsub)
    return dict(forward=22.0 * L * (g.res + 1), back=10.0 * g.nvox * nviews)


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
    return dict(value=nsteps / dt, unit="subset-iters/s", cores=orc.get_threads(), kind="port",
                sample="%d subset-iterations (S=%d) of the same Inveon-scale study, "
                       "oracle FP32 restatement with OpenMP" % (nsteps, nsub),
                seconds=dt)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=48)
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--nsub", type=int, default=16)
    ap.add_argument("--counts", type=float, default=5e7)
    ap.add_argument("--storage", default="f32", choices=["f32", "f16", "bf16"])
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import gfirst_amd as ga
    from gfirst_amd.geometry import C1
    g = C1
    storage = {"f32": ga.STORE_F32, "f16": ga.STORE_F16, "bf16": ga.STORE_BF16}[args.storage]
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
    log("autotuned tiles: back slices/thread=%d, z-sum sinograms/wave=%d (measured=%s)"
        % tune[:3])
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
    res = {
        "metric": "Inveon-size OSEM subset-iters/s",
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
                   "subsets": args.nsub, "storage": args.storage},
        "hbm_gbps_alg": round(ab["total"] / (ms_step * 1e-3) / 1e9, 2),
        "phase_ms": {k: round(v[0] / max(v[1], 1), 4) for k, v in ph.items() if v[1]},
        "roofline": {"kernel": "k_forward" if dom == "forward" else "k_back", "bound": "hbm",
                     "achieved": round(achieved, 2), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 5), "traffic": None,
                     "alg_bytes_per_launch": ab[dom], "avg_launch_ms": round(avg_ms, 4),
                     "fp32_tflops": round(flops, 2),
                     "fp32_frac": round(flops / PEAK_FP32_TFLOPS, 4)},
        "loglik_last_pass": ll,
        "tiles": {"back_slices_per_thread": tune[0], "zsum_sinograms_per_wave": tune[1],
                  "autotuned": tune[2]},
    }
    r.close()
    ctx.close()
    if not args.no_cpu and args.cpu_steps > 0:
        log("timing the CPU oracle on %d subset-iterations ..." % args.cpu_steps)
        res["cpu_baseline"] = cpu_baseline(g, d2, norm, pb, args.nsub, args.cpu_steps)
        log("cpu: %.3f subset-iters/s on %d threads" % (res["cpu_baseline"]["value"],
                                                       res["cpu_baseline"]["cores"]))
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
