"""The backprojector's shared z axis (one floor/fraction per voxel column and
view instead of one per slice) reproduces the reference's per-slice plane
positions bit for bit under the binade condition the kernel tests; checked by
sampling 2e7 (yr, fzb, zb, ZB) cases, half of them aimed at binade
boundaries, compiled with the same no-contraction rule as the kernels."""
import os
import subprocess

ROOT = os.path.dirname(os.path.abspath(__file__))


def test_shared_z_axis_identity(tmp_path):
    exe = tmp_path / "zaxis"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", str(exe),
                    os.path.join(ROOT, "zaxis_identity.c"), "-lm"], check=True)
    out = subprocess.run([str(exe), "20000000"], capture_output=True, text=True, timeout=300)
    n, fast, bad = (int(v) for v in out.stdout.split())
    print("cases %d, fast-path cases %d, mismatches %d" % (n, fast, bad))
    assert out.returncode == 0 and bad == 0
    assert fast > n // 3  # the fast path is the common case
