"""The drop-in executable (gfirst_amd/bin/gfirst, C ABI only) run like the
reference binary on raw files in a scratch directory: its output files equal
the oracle's computation of the same run bit for bit (projection mode, and
reconstruction modes 1-3 with the reference's file layouts: int32 prompts
auto-detected, float normalization, 1-based VLOR records)."""
import subprocess

import numpy as np
import pytest

from gfirst_amd import _lib
from gfirst_amd.geometry import REF, SMALL

pytestmark = pytest.mark.gpu

G = SMALL
FLAGS = ["--res=%d" % G.res, "--nzs=%d" % G.nzs, "--nrad=%d" % G.nrad, "--nang=%d" % G.nang,
         "--span=%d" % G.span, "--nseg=%d" % G.nseg, "--pitch=%r" % G.pitch,
         "--diam=%r" % G.diam_det, "--fov=%r" % G.fov]
SCALE = np.float32(3.70e-4 * 162.0)  # GFIRST.h:29-30


def run(cwd, *args, flags=FLAGS):
    out = subprocess.run([_lib.CLI_PATH] + flags + list(args), cwd=cwd, capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    return out.stdout


def raw(path, dtype=np.float32):
    return np.fromfile(path, dtype=dtype)


def test_projection_mode(tmp_path, orc):
    img = np.random.default_rng(51).uniform(0, 2, G.nvox).astype(np.float32)
    img.tofile(tmp_path / "image_volume.raw")
    stdout = run(tmp_path, "image_volume.raw")
    assert "Finish After Projection" in stdout
    assert np.array_equal(raw(tmp_path / "proj_ref.raw"), orc.forward(G, img))


@pytest.mark.parametrize("mode", [1, 2, 3])
def test_reconstruction_modes(tmp_path, orc, mode):
    rng = np.random.default_rng(60 + mode)
    nd = G.ndata
    prompts = rng.poisson(rng.uniform(0.2, 3.0, nd) * 10).astype(np.int32)
    prompts.tofile(tmp_path / "sinogramD.raw")          # int32: the reader's integer path
    norm = rng.uniform(0.8, 1.2, nd).astype(np.float32)
    norm[rng.uniform(size=nd) < 0.02] = 0.0
    norm.tofile(tmp_path / "normalization.raw")
    args = ["--niter=2", "--nsub=3", "sinogramD.raw"]
    if mode >= 2:
        trip = rng.poisson(rng.uniform(0.2, 3.0, nd)).astype(np.int32)
        trip.tofile(tmp_path / "sinogramT.raw")
        args.append("sinogramT.raw")
    args.append("normalization.raw")
    if mode == 3:
        rng.integers(1, nd, size=(2500, 2)).astype(np.int32).tofile(tmp_path / "VLORs.raw")
        args.append("VLORs.raw")
    run(tmp_path, *args)

    def decoded(name):
        out, _ = orc.decode_sinogram((tmp_path / name).read_bytes(), nd)
        return out

    d2 = decoded("sinogramD.raw")
    assert np.array_equal(d2[:nd], prompts.astype(np.float32)) and not d2[nd:].any()
    nrm = decoded("normalization.raw")[:nd]
    assert np.array_equal(nrm, norm)
    t2 = decoded("sinogramT.raw") if mode >= 2 else None
    rec = orc.vlor_records((tmp_path / "VLORs.raw").read_bytes()) if mode == 3 else None
    if rec is not None:
        assert len(rec) == 2501  # the last record once more, as the reference's loop reads it
    xD, xT, pb, vo = orc.osem(G, orc.params(mode=mode, niter=2, nsub=3), d2, t2, nrm, rec,
                              want_vlor=mode == 3)
    assert np.array_equal(raw(tmp_path / "proj_ref.raw"), pb)
    assert np.array_equal(raw(tmp_path / "img_GPU_D_ITER2.raw"), orc.output(G, xD, SCALE))
    if mode >= 2:
        assert np.array_equal(raw(tmp_path / "img_GPU_T_ITER2.raw"), orc.output(G, xT, SCALE))
    if mode == 3:
        assert np.array_equal(raw(tmp_path / "SINOG_VLOR.raw"), vo)


def test_default_geometry_doubles_reconstruction(tmp_path, orc):
    """The executable with no geometry flags runs the reference's compiled
    geometry (GFIRST.h:4-16: 147 x 168 x 1293 sinogram, 128 x 128 x 159 image):
    one doubles iteration (README mode 2) of 84 two-angle subsets on int32
    prompts, against the oracle's run of the same files bit for bit. About 10 s
    of oracle work on 16 host threads."""
    g = REF
    nd = g.ndata
    rng = np.random.default_rng(81)
    prompts = rng.poisson(rng.uniform(0.05, 0.6, nd)).astype(np.int32)
    prompts.tofile(tmp_path / "sinogram.raw")
    norm = rng.uniform(0.8, 1.2, nd).astype(np.float32)
    norm[rng.uniform(size=nd) < 0.02] = 0.0
    norm.tofile(tmp_path / "normalization.raw")
    run(tmp_path, "--niter=1", "--nsub=84", "sinogram.raw", "normalization.raw", flags=[])
    d2 = np.zeros(2 * nd, np.float32)
    d2[:nd] = prompts
    xD, _, pb, _ = orc.osem(g, orc.params(mode=1, niter=1, nsub=84), d2, None, norm)
    assert np.array_equal(raw(tmp_path / "proj_ref.raw"), pb)
    got = raw(tmp_path / "img_GPU_D_ITER1.raw")
    exp = orc.output(g, xD, SCALE)
    print("ref geometry CLI: %d/%d voxels bit-exact" % (int(np.sum(got == exp)), got.size))
    assert np.array_equal(got, exp)
