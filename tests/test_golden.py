"""The oracle still reproduces the committed regression fixtures
(tests/golden/, written by tests/golden/make_golden.py from the oracle at the
16x16x8 test geometry), bit for bit."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))


def test_oracle_reproduces_fixtures(orc):
    import make_golden
    x = dict(np.load(os.path.join(HERE, "golden", "tiny_inputs.npz")))
    ref = dict(np.load(os.path.join(HERE, "golden", "tiny_outputs.npz")))
    assert set(x) == set(make_golden.inputs())
    for k, v in make_golden.inputs().items():
        assert np.array_equal(v, x[k]), "input %s changed" % k
    got = make_golden.outputs(x)
    assert set(got) == set(ref)
    for k in ref:
        assert np.array_equal(got[k], ref[k]), "oracle output %s changed" % k
