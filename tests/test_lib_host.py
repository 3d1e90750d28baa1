"""CPU checks of libgfirst.so: it loads, exports every entry point declared in
include/gfirst.h, and its host-only geometry/tap tables equal the oracle's bit
for bit (no kernel is launched here)."""
import ctypes as C
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import gfirst_amd as ga
from gfirst_amd import _lib
from gfirst_amd.geometry import C0, C1, REF, SMALL, TINY

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "gfirst.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(gfr_[a-z0-9_]+)\s*\(", src)))


def test_header_and_bindings_agree():
    names = declared_functions()
    assert len(names) >= 40
    assert set(names) == set(_lib.SIGNATURES), set(names) ^ set(_lib.SIGNATURES)


def declared_struct(name):
    """(C type, field name) of `typedef struct name { ... } name;` in the header."""
    src = open(os.path.join(ROOT, "include", "gfirst.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    body = re.search(r"typedef struct %s \{(.*?)\} %s;" % (name, name), src, flags=re.S).group(1)
    out = []
    for decl in body.split(";"):
        decl = " ".join(decl.split())
        if not decl:
            continue
        m = re.match(r"(.+?)\s+(\w+(?:\s*,\s*\w+)*)$", decl)
        out += [(m.group(1), n.strip()) for n in m.group(2).split(",")]
    return out


CTYPE = {"int": C.c_int, "float": C.c_float, "double": C.c_double, "long long": C.c_longlong,
         "size_t": C.c_size_t}


@pytest.mark.parametrize("cname,py", [("gfr_geometry", "GfrGeometry"),
                                      ("gfr_geometry_info", "GfrGeometryInfo"),
                                      ("gfr_recon_params", "GfrReconParams"),
                                      ("gfr_tuning", "GfrTuning"),
                                      ("gfr_batch_params", "GfrBatchParams"),
                                      ("gfr_request_info", "GfrRequestInfo"),
                                      ("gfr_server_totals", "GfrServerTotals")])
def test_struct_mirrors_match_header(cname, py):
    """The ctypes mirrors have the header's fields, in order, with its types."""
    fields = getattr(_lib, py)._fields_
    decl = declared_struct(cname)
    assert [n for n, _ in fields] == [n for _, n in decl]
    for (n, t), (ct, _) in zip(fields, decl):
        if ct in CTYPE:
            assert t == CTYPE[ct], (n, ct, t)


def test_library_exports_every_declared_symbol():
    L = C.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (gfr_[a-z0-9_]+)$", out, flags=re.M))
    assert set(declared_functions()) <= exported


def test_default_geometry_is_reference_constants():
    g = _lib.GfrGeometry()
    assert ga.lib().gfr_geometry_default(C.byref(g)) == 0
    assert (g.res, g.nzs, g.nrad, g.nang, g.span, g.nseg) == (128, 159, 147, 168, 11, 15)
    assert abs(g.pitch - 1.63) < 1e-6 and abs(g.diam_det - 165.7758) < 1e-4
    assert REF.nsinogs == 1293


@pytest.mark.parametrize("g", [REF, C0, C1, TINY, SMALL], ids=["ref", "c0", "c1", "tiny", "small"])
def test_geometry_tables_equal_oracle(orc, g):
    L = ga.lib()
    inf = g.info()
    assert inf.nsinogs == orc.nsinogs(g)
    assert inf.ndata == orc.ndata(g)
    assert (inf.lambda0, inf.lambda1) == (-(g.res // 2), g.res // 2)
    zmin = np.zeros(g.nseg, np.int32)
    sgn = np.zeros(g.nseg, np.int32)
    nza = np.zeros(g.nseg + 1, np.int32)
    assert L.gfr_seg_tables(C.byref(g.c()), zmin.ctypes.data, sgn.ctypes.data,
                            nza.ctypes.data) == inf.nsinogs
    _, z2, s2, n2 = orc.seg_tables(g)
    assert np.array_equal(zmin, z2) and np.array_equal(sgn, s2) and np.array_equal(nza, n2)
    st = np.zeros(g.nang, np.float32)
    ct = np.zeros(g.nang, np.float32)
    assert L.gfr_angle_tables(C.byref(g.c()), st.ctypes.data, ct.ctypes.data) == 0
    so, co = orc.angle_tables(g)
    assert np.array_equal(st, so) and np.array_equal(ct, co)


def test_taps_equal_oracle(orc):
    L = ga.lib()
    for psf in (0.5, 1.0, 1.5, 2.25):
        t = np.zeros(9, np.float32)
        L.gfr_psf_taps(psf, t.ctypes.data)
        assert np.array_equal(t, orc.psf_taps(psf))
    t5 = np.zeros(5, np.float32)
    L.gfr_gz_taps(t5.ctypes.data)
    assert np.array_equal(t5, orc.gz_taps())


def test_invalid_geometry_rejected():
    bad = [TINY.with_(ntbins=2), TINY.with_(nseg=99), TINY.with_(res=0), TINY.with_(pitch=0.0)]
    for g in bad:
        with pytest.raises(ValueError):
            g.info()


def test_footprint_scales_with_storage():
    p32 = ga.recon_params(mode=1, nsub=16, storage=ga.STORE_F32)
    p16 = ga.recon_params(mode=1, nsub=16, storage=ga.STORE_F16)
    f32, f16 = ga.footprint(C1, p32), ga.footprint(C1, p16)
    assert f32 > f16 > 0
    assert f32 - f16 == 2 * 2 * C1.ndata  # prompts + background halve


def test_batch_footprint():
    bp = ga.batch_params(storage=ga.STORE_F16, max_frames=64, min_nsub=8, lanes=2)
    base = ga.batch_footprint(C1, bp)
    per_bg = ga.batch_footprint(C1, bp, 1, has_bg=True) - base
    per = ga.batch_footprint(C1, bp, 1) - base
    assert per == 4 * C1.nvox + 2 * C1.ndata and per_bg == per + 2 * C1.ndata
    # 64 FP16 frames of the Inveon-size study fit comfortably in one GPU
    assert ga.batch_footprint(C1, bp, 64, has_bg=True) < 32 << 30
    # one more lane adds one scratch set; a larger min_nsub shrinks it
    bp3 = ga.batch_params(storage=ga.STORE_F16, max_frames=64, min_nsub=8, lanes=3)
    bp16 = ga.batch_params(storage=ga.STORE_F16, max_frames=64, min_nsub=16, lanes=2)
    assert ga.batch_footprint(C1, bp3) > base > ga.batch_footprint(C1, bp16)
    for bad in (ga.batch_params(lanes=0), ga.batch_params(min_nsub=0),
                ga.batch_params(max_frames=0), ga.batch_params(group=4),
                ga.batch_params(group=-1)):
        assert ga.batch_footprint(C1, bad) == ga.GFR_EINVAL
    # frame groups: per lane, interleaved rows (twice the packed rows) and a
    # numerator and denominator per extra member; group 0 means 3
    bp1 = ga.batch_params(storage=ga.STORE_F16, max_frames=64, min_nsub=8, lanes=2, group=1)
    bp2 = ga.batch_params(storage=ga.STORE_F16, max_frames=64, min_nsub=8, lanes=2, group=2)
    bp3g = ga.batch_params(storage=ga.STORE_F16, max_frames=64, min_nsub=8, lanes=2, group=3)
    rows = 8 * 20 * (C1.nsinogs + 2 * C1.nseg) * (C1.nrad + 2)  # {ratio, calib}, 20 angles
    f1, f2 = ga.batch_footprint(C1, bp1), ga.batch_footprint(C1, bp2)
    assert f2 - f1 == 2 * (2 * rows + 2 * 4 * C1.nvox)
    assert ga.batch_footprint(C1, bp3g) - f2 == 2 * (2 * 4 * C1.nvox) == base - f2


def test_read_sinogram_file_semantics(tmp_path, orc):
    nd = 12
    p = tmp_path / "s.raw"
    ints = np.arange(2 * nd, dtype=np.int32) * 3
    ints.tofile(p)
    from gfirst_amd import rawio
    out, found, isf, tr = rawio.read_sinogram(p, nd)
    exp, kind = orc.decode_sinogram(ints.tobytes(), nd)
    assert found and not isf and not tr and np.array_equal(out, exp)
    fl = np.linspace(0.25, 9, 2 * nd + 5).astype(np.float32)
    fl.tofile(p)
    out, found, isf, tr = rawio.read_sinogram(p, nd)
    exp, kind = orc.decode_sinogram(fl.tobytes(), nd)
    assert isf and tr and kind == 1 and np.array_equal(out, exp)
    out, found, isf, tr = rawio.read_sinogram(tmp_path / "missing.raw", nd)
    assert not found and np.all(out == 0)


def test_read_vlor_file_semantics(tmp_path, orc):
    from gfirst_amd import rawio
    p = tmp_path / "v.raw"
    for words in ([3, 4, 5, 6], [3, 4, 5], [], [7, 8]):
        w = np.array(words, np.int32)
        w.tofile(p)
        got = rawio.read_vlor(p)
        exp = orc.vlor_records(w.tobytes())
        assert np.array_equal(got.reshape(-1, 2), exp.reshape(-1, 2))


def test_cli_usage_without_gpu():
    out = subprocess.run([_lib.CLI_PATH], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0 and "1 Input  --> Projection" in out.stdout


def test_bench_refuses_more_than_one_gpu():
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    out = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"],
                         capture_output=True, text=True, timeout=300, cwd=root)
    assert out.returncode == 2 and out.stdout == ""
    assert "one GPU" in out.stderr and "--gpus 2" in out.stderr
