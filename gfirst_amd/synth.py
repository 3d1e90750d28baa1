"""Seeded synthetic inputs for the OSEM path (phantoms, Poisson sinograms).

Everything derives from numpy's PCG64 generator with an explicit seed, so any
machine with the same numpy regenerates identical inputs.
"""
import numpy as np

from .geometry import Geometry


def voxel_centres_mm(g: Geometry):
    """(z, y, x) voxel-centre coordinates in mm; the z voxel is pitch/2
    (planes index slices directly, GFIRST_GPU_Kernels.cu:88)."""
    vx = g.fov / g.res
    vz = g.pitch / 2.0
    x = (np.arange(g.res) - (g.res - 1) / 2.0) * vx
    z = (np.arange(g.nzs) - (g.nzs - 1) / 2.0) * vz
    return z, x, x


def _ellipsoid(g, centre, semi, value, img):
    z, y, x = voxel_centres_mm(g)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    m = (((X - centre[0]) / semi[0]) ** 2 + ((Y - centre[1]) / semi[1]) ** 2 +
         ((Z - centre[2]) / semi[2]) ** 2) <= 1.0
    img[m] = value
    return img


def phantom_c0(g: Geometry, seed=0, radius=25.0, sphere_r=4.0, hot=4.0):
    """Uniform cylinder (radius 25 mm, value 1, all slices) with 3 hot spheres
    (radius 4 mm, value 4) whose centres are drawn inside r <= 17 mm."""
    rng = np.random.default_rng(seed)
    z, y, x = voxel_centres_mm(g)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    img = np.where(X ** 2 + Y ** 2 <= radius ** 2, 1.0, 0.0).astype(np.float32)
    zr = max(abs(z[0]) - sphere_r, 0.0)
    for _ in range(3):
        rr = 17.0 * np.sqrt(rng.uniform())
        ph = rng.uniform(0, 2 * np.pi)
        cz = rng.uniform(-zr, zr)
        m = (X - rr * np.cos(ph)) ** 2 + (Y - rr * np.sin(ph)) ** 2 + (Z - cz) ** 2 <= sphere_r ** 2
        img[m] = hot
    return img.reshape(-1)


def phantom_mouse(g: Geometry, seed=0, channel=0):
    """Mouse-like union of ellipsoids: body 25x25x90 mm (full axes), value 1,
    and 6 organ ellipsoids with values in [0.5, 8] at seeded positions.
    channel selects an independent organ draw (tracer A = 0, B = 1)."""
    rng = np.random.default_rng(seed * 7919 + channel)
    img = np.zeros((g.nzs, g.res, g.res), np.float32)
    zhalf = min(45.0, (g.nzs - 1) / 2.0 * g.pitch / 2.0)
    _ellipsoid(g, (0.0, 0.0, 0.0), (12.5, 12.5, zhalf), 1.0, img)
    vals = rng.uniform(0.5, 8.0, size=6)
    for k in range(6):
        semi = rng.uniform(2.0, 5.0, size=3)
        semi[2] = rng.uniform(3.0, 10.0)
        r = rng.uniform(0, 12.5 - semi[:2].max())
        ph = rng.uniform(0, 2 * np.pi)
        cz = rng.uniform(-(zhalf - semi[2]), zhalf - semi[2])
        _ellipsoid(g, (r * np.cos(ph), r * np.sin(ph), cz), semi, vals[k], img)
    return img.reshape(-1)


def poisson_sinogram(ybar, total_counts, seed=0):
    """Scale the expected sinogram to total_counts and draw Poisson counts
    (int32, as a prompts file the reader takes the integer path for)."""
    ybar = np.asarray(ybar, np.float64)
    s = ybar.sum()
    lam = ybar * (total_counts / s) if s > 0 else ybar
    rng = np.random.default_rng(seed)
    return rng.poisson(lam).astype(np.int32)


def norm_uniform(ndata, lo=0.8, hi=1.2, seed=0):
    rng = np.random.default_rng(seed)
    return rng.uniform(lo, hi, size=ndata).astype(np.float32)


def as_file_pair(prompts, background=None):
    """The 2*NDATA float array the sinogram reader returns (prompts, background)."""
    nd = prompts.size
    out = np.zeros(2 * nd, np.float32)
    out[:nd] = prompts
    if background is not None:
        out[nd:] = background
    return out


def frame_schedule(nframes):
    """Frame start/end times (min) of a dynamic acquisition: short frames
    first, then longer ones (10 s, 30 s, 1 min, 2 min, 5 min blocks)."""
    blocks = [(10 / 60.0, 12), (0.5, 12), (1.0, 12), (2.0, 14), (5.0, 14)]
    t, out = 0.0, []
    while len(out) < nframes:
        for d, n in blocks:
            for _ in range(n):
                if len(out) < nframes:
                    out.append((t, t + d))
                    t += d
        blocks = [(5.0, nframes)]
    return np.array(out, np.float64)


def two_compartment(t, K1, k2, A=(8.0, 1.5), lam=(2.0, 0.05)):
    """Plasma input Cp(t) = sum_i A_i exp(-lam_i t) and tissue response of a
    one-tissue (two-compartment) model, Ct = K1 Cp (*) exp(-k2 t), in closed
    form: Ct(t) = K1 sum_i A_i (exp(-lam_i t) - exp(-k2 t)) / (k2 - lam_i)."""
    t = np.asarray(t, np.float64)
    cp = sum(a * np.exp(-l * t) for a, l in zip(A, lam))
    ct = K1 * sum(a * (np.exp(-l * t) - np.exp(-k2 * t)) / (k2 - l) for a, l in zip(A, lam))
    return cp, ct


def dynamic_study(nframes, seed=0, counts=(1e5, 1e7), subsets=(8, 16, 24)):
    """Per-frame parameters of the batched dynamic study: mid-frame time,
    seeded subset count, log-uniform total counts and the two compartment
    weights (plasma, tissue) of that frame."""
    rng = np.random.default_rng(seed)
    fr = frame_schedule(nframes)
    tm = fr.mean(axis=1)
    K1, k2 = rng.uniform(0.1, 0.6), rng.uniform(0.05, 0.3)
    cp, ct = two_compartment(tm, K1, k2)
    nsub = rng.choice(np.asarray(subsets), size=nframes)
    cts = np.exp(rng.uniform(np.log(counts[0]), np.log(counts[1]), size=nframes))
    return [dict(t=float(tm[k]), nsub=int(nsub[k]), counts=float(cts[k]), cp=float(cp[k]),
                 ct=float(ct[k])) for k in range(nframes)]
