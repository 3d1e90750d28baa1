"""Runtime geometry record (the reference compiles it in: GFIRST.h:4-16, 31-34).

Presets:
  REF   the compiled constants of the reference (Inveon, span 11, 147x168x1293)
  C0    64x64x32 image, 64 radial x 64 angles, span 3, max ring difference 15
  C1    Inveon-scale span 3: 128x128x159 image, 128x160x4319 sinogram
"""
import ctypes as C
from dataclasses import dataclass, replace

from . import _lib


@dataclass(frozen=True)
class Geometry:
    res: int
    nzs: int
    nrad: int
    nang: int
    span: int
    nseg: int
    pitch: float
    diam_det: float
    fov: float
    ntbins: int = 1

    def c(self):
        return _lib.GfrGeometry(self.res, self.nzs, self.nrad, self.nang, self.span, self.nseg,
                                self.ntbins, self.pitch, self.diam_det, self.fov)

    def info(self):
        inf = _lib.GfrGeometryInfo()
        g = self.c()
        rc = _lib.lib().gfr_geometry_check(C.byref(g), C.byref(inf))
        if rc != 0:
            raise ValueError("invalid geometry %r" % (self,))
        return inf

    @property
    def nsinogs(self):
        return self.info().nsinogs

    @property
    def ndata(self):
        return self.info().ndata

    @property
    def nvox(self):
        return self.res * self.res * self.nzs

    @property
    def nviews(self):
        return self.nseg * self.nang

    def with_(self, **kw):
        return replace(self, **kw)


def nseg_for(span, mrd):
    """Segments whose minimum ring difference does not exceed mrd."""
    k = 0
    while (k + 1 - 1) * span + span // 2 + 1 <= mrd:
        k += 1
    return 2 * k + 1


REF = Geometry(128, 159, 147, 168, 11, 15, 1.63, 165.7758, 100.0)
C0 = Geometry(64, 32, 64, 64, 3, nseg_for(3, 15), 2.0, 106.1, 64.0)
C1 = Geometry(128, 159, 128, 160, 3, nseg_for(3, 79), 1.63, 165.7758, 99.328)
# Small geometries for parity tests (seconds on the CPU oracle).
TINY = Geometry(16, 8, 16, 8, 3, 3, 2.0, 26.5, 16.0)
SMALL = Geometry(32, 16, 36, 24, 3, 5, 2.0, 53.0, 32.0)

PRESETS = {"ref": REF, "c0": C0, "c1": C1, "tiny": TINY, "small": SMALL}
