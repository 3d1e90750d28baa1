"""List-mode triples: synthetic detector hits and their binning to VLOR
records (gfr_lor_bin / gfr_vlor_bin_hits of the C ABI)."""
import ctypes as C

import numpy as np

from ._lib import GfrError, lib
from .synth import voxel_centres_mm


def lor_bin(g, a, b):
    """0-based sinogram bin of the LOR through a and b (mm), -1 if outside."""
    pa = (C.c_double * 3)(*[float(v) for v in a])
    pb = (C.c_double * 3)(*[float(v) for v in b])
    return lib().gfr_lor_bin(C.byref(g.c()), pa, pb)


def bin_hits(g, hits, radius):
    """VLOR records (N x 2 int32, 1-based bins, 0 = outside) of triple events
    given as N x 6 float32 (phi0, z0, phi1, z1, phi2, z2) on a cylinder."""
    h = np.ascontiguousarray(hits, np.float32).reshape(-1, 6)
    rec = np.zeros((len(h), 2), np.int32)
    n = lib().gfr_vlor_bin_hits(C.byref(g.c()), h.ctypes.data, len(h), float(radius),
                                rec.ctypes.data)
    if n < 0:
        raise GfrError(int(n), "gfr_vlor_bin_hits")
    return rec


def _cyl_hits(E, u, radius):
    """Both intersections of the lines E + t u with the cylinder x^2+y^2=R^2."""
    a = u[:, 0] ** 2 + u[:, 1] ** 2
    b = 2 * (E[:, 0] * u[:, 0] + E[:, 1] * u[:, 1])
    c = E[:, 0] ** 2 + E[:, 1] ** 2 - radius ** 2
    disc = np.sqrt(np.maximum(b * b - 4 * a * c, 0.0))
    a = np.maximum(a, 1e-12)
    tp = (-b + disc) / (2 * a)
    tm = (-b - disc) / (2 * a)
    return E + tp[:, None] * u, E + tm[:, None] * u


def _dirs(rng, n):
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def triple_events(g, activity, nev, radius=80.5, axial=127.0, seed=0):
    """nev triple events drawn from `activity` (NVOX image): an annihilation
    pair (two back-to-back photons) and a prompt gamma from the same point,
    all three detected on the cylinder (|z| <= axial/2). Hit 0 is one of the
    annihilation hits; hits 1 and 2 are the other annihilation hit and the
    prompt gamma in random order. Returns N x 6 float32 (phi, z) x 3."""
    rng = np.random.default_rng(seed)
    act = np.asarray(activity, np.float64).reshape(-1)
    p = act / act.sum()
    zc, yc, xc = voxel_centres_mm(g)
    vx = g.fov / g.res
    vz = g.pitch / 2.0
    out = np.zeros((0, 6), np.float32)
    while len(out) < nev:
        m = int((nev - len(out)) * 1.6) + 64
        idx = rng.choice(act.size, size=m, p=p)
        iz, rem = np.divmod(idx, g.res * g.res)
        iy, ix = np.divmod(rem, g.res)
        E = np.stack([xc[ix] + (rng.uniform(size=m) - 0.5) * vx,
                      yc[iy] + (rng.uniform(size=m) - 0.5) * vx,
                      zc[iz] + (rng.uniform(size=m) - 0.5) * vz], axis=1)
        A, B = _cyl_hits(E, _dirs(rng, m), radius)
        P, _ = _cyl_hits(E, _dirs(rng, m), radius)
        ok = ((np.abs(A[:, 2]) <= axial / 2) & (np.abs(B[:, 2]) <= axial / 2) &
              (np.abs(P[:, 2]) <= axial / 2))
        A, B, P = A[ok], B[ok], P[ok]
        swap = rng.uniform(size=len(A)) < 0.5
        h0 = np.where(swap[:, None], B, A)
        ha = np.where(swap[:, None], A, B)
        order = rng.uniform(size=len(A)) < 0.5
        h1 = np.where(order[:, None], ha, P)
        h2 = np.where(order[:, None], P, ha)
        ev = np.stack([np.arctan2(h0[:, 1], h0[:, 0]), h0[:, 2],
                       np.arctan2(h1[:, 1], h1[:, 0]), h1[:, 2],
                       np.arctan2(h2[:, 1], h2[:, 0]), h2[:, 2]], axis=1).astype(np.float32)
        out = np.concatenate([out, ev])
    return out[:nev]
