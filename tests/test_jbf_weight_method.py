"""The fast intensity weight of the radius-4 joint bilateral filter, checked
on the CPU (tests/jbf_weight_check.c restates its arithmetic): the FMA
quotient equals IEEE division, and every weight the fast path keeps equals
the reference expression (float)exp(-(double)di2 / kI)
(jointbilateralfilter3D_K, GFIRST_GPU_Kernels.cu:508-539). Sampled here;
the same program run with larger arguments covers every divisor significand
and every float di2 (0 mismatches over ~1e9 values each). The GPU parity
tests compare the kernel itself with the oracle bit for bit."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("jbf") / "jbf_weight_check")
    subprocess.run(["gcc", "-O2", "-mfma", "-o", exe, os.path.join(HERE, "jbf_weight_check.c"),
                    "-lm"], check=True)
    return exe


def run(exe, *args):
    out = subprocess.run([exe, *map(str, args)], check=True, capture_output=True, text=True)
    name, cases, slow, bad = out.stdout.split()
    return int(cases), int(slow), int(bad)


def test_fma_quotient_is_ieee_division(checker):
    cases, _, bad = run(checker, "div", 2)
    assert cases == 2 * (1 << 23) and bad == 0


@pytest.mark.parametrize("sigma", [0.025, 0.3])
def test_kept_weights_equal_reference(checker, sigma):
    cases, slow, bad = run(checker, "exp", 61, sigma)
    assert cases > 10_000_000 and bad == 0
    assert slow < cases // 100


@pytest.mark.parametrize("sigma", [0.025, 1e-3, 100.0])
def test_whole_weight_all_regimes(checker, sigma):
    cases, slow, bad = run(checker, "whole", 3, sigma)
    assert cases > 2_000_000 and bad == 0
