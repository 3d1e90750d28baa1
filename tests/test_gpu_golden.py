"""The HIP path against the committed oracle fixtures (tests/golden/) at the
16x16x8 test geometry: every operator and 1-iteration reconstructions in
modes 1-3, compared with the stored arrays directly (no oracle call)."""
import os

import numpy as np
import pytest

import gfirst_amd as ga
from gfirst_amd.geometry import TINY

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def fx():
    x = dict(np.load(os.path.join(HERE, "golden", "tiny_inputs.npz")))
    y = dict(np.load(os.path.join(HERE, "golden", "tiny_outputs.npz")))
    return x, y


def test_operators_match_fixtures(gctx, fx):
    x, y = fx
    g = TINY
    assert np.array_equal(gctx.blur_xy(g, x["img"], 1.5), y["blur"])
    assert np.array_equal(gctx.forward_project(g, x["img"]), y["fwd_s1"])
    assert np.array_equal(gctx.forward_project(g, x["img"], 1, 4), y["fwd_1of4"])
    n, d = gctx.back_project(g, x["corr"], (x["cal"] > 0).astype(np.uint8), 1, 2)
    assert np.array_equal(n, y["back_num"]) and np.array_equal(d, y["back_den"])
    assert np.array_equal(gctx.update(g, x["img"], x["num"], x["den"], 1.5), y["update"])
    j = gctx.jbf3d(g, x["img"], x["num"], float(x["num"].max()))
    assert np.array_equal(j, y["jbf"])


@pytest.mark.parametrize("mode", [1, 2, 3])
def test_reconstruction_matches_fixtures(gctx, fx, mode):
    x, y = fx
    g = TINY
    r = gctx.recon(g, ga.recon_params(mode=mode, niter=1, nsub=2), x["d2"],
                   x["t2"] if mode >= 2 else None, x["norm"], x["rec"] if mode == 3 else None)
    assert np.array_equal(r.get(ga.OUT_PROJ_BLANK), y["proj_blank"])
    r.run()
    assert np.array_equal(r.get(ga.OUT_XD), y["osem%d_xD" % mode])
    if mode >= 2:
        assert np.array_equal(r.get(ga.OUT_XT), y["osem%d_xT" % mode])
    r.close()
