"""Writes the regression fixtures of tests/golden/ (run from the repository
root: python tests/golden/make_golden.py).

The reference ships no tests, fixtures or golden vectors for this path, and
its CUDA sources cannot be built here, so these vectors come from the CPU
oracle (oracle/, the FP32 restatement of the reference arithmetic, itself
pinned by tests/test_oracle_kat.py). They freeze its outputs on seeded
inputs at the 16x16x8 test geometry: the CPU suite checks that the oracle
still reproduces them, the GPU suite checks the HIP kernels against them
directly. Arrays only (numpy .npz, no pickles).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from gfirst_amd.geometry import TINY  # noqa: E402
from oracle import oracle as orc  # noqa: E402


def inputs(seed=2024):
    g = TINY
    rng = np.random.default_rng(seed)
    nd, nv = g.ndata, g.nvox
    img = rng.uniform(0.0, 2.0, nv).astype(np.float32)
    lam = rng.uniform(0.2, 3.0, nd) * 20.0
    d2 = np.zeros(2 * nd, np.float32)
    d2[:nd] = rng.poisson(lam)
    d2[nd:] = rng.uniform(0, 0.3, nd)
    t2 = np.zeros(2 * nd, np.float32)
    t2[:nd] = rng.poisson(lam * 0.2)
    norm = rng.uniform(0.8, 1.2, nd).astype(np.float32)
    norm[rng.uniform(size=nd) < 0.03] = 0.0
    corr = rng.uniform(0, 3, nd).astype(np.float32)
    corr[rng.uniform(size=nd) < 0.05] = -1000.0
    cal = np.where(rng.uniform(size=nd) < 0.93, 1.0, -1000.0).astype(np.float32)
    num = rng.uniform(0, 50, nv).astype(np.float32)
    den = rng.uniform(0, 50, nv).astype(np.float32)
    den[::13] = 0.0
    rec = rng.integers(1, nd, size=(300, 2)).astype(np.int32)
    rec = np.concatenate([rec, rec[-1:]])
    return dict(img=img, d2=d2, t2=t2, norm=norm, corr=corr, cal=cal, num=num, den=den, rec=rec)


def outputs(x):
    g = TINY
    out = {}
    out["blur"] = orc.blur_xy(g, x["img"], 1.5)
    out["fwd_s1"] = orc.forward(g, x["img"])
    out["fwd_1of4"] = orc.forward(g, x["img"], 1, 4)
    out["back_num"], out["back_den"] = orc.back(g, x["corr"], x["cal"], 1, 2)
    out["update"] = orc.update(g, x["img"], x["num"], x["den"], 1.5)
    out["jbf"] = orc.jbf(g, x["img"], x["num"], float(orc.vmax(x["num"])))
    for mode in (1, 2, 3):
        xD, xT, pb, _ = orc.osem(g, orc.params(mode=mode, niter=1, nsub=2), x["d2"],
                                 x["t2"] if mode >= 2 else None, x["norm"],
                                 x["rec"] if mode == 3 else None)
        out["osem%d_xD" % mode] = xD
        if mode >= 2:
            out["osem%d_xT" % mode] = xT
        out["proj_blank"] = pb
    return out


def main():
    x = inputs()
    np.savez_compressed(os.path.join(HERE, "tiny_inputs.npz"), **x)
    np.savez_compressed(os.path.join(HERE, "tiny_outputs.npz"), **outputs(x))
    print("wrote tiny_inputs.npz, tiny_outputs.npz")


if __name__ == "__main__":
    main()
