// gfr_lor.cpp — binning of detected LORs into the sinogram (host only).
//
// The reference consumes VLOR records as pairs of 1-based sinogram bin
// indices (GFIRST.cu:213-236); list-mode triples come as detector hits. A LOR
// between points A and B (mm, scanner frame, z axial, origin at the centre)
// is binned with the parametrisation of the forward projector
// (GFIRST_GPU_Kernels.cu:75-96): for angle index it, theta = PI - it*PI/NANG,
// the LOR of radial bin r is {s n + t d} with n = (cos theta, sin theta),
// d = (sin theta, -cos theta), s = (pitch/2)(r - RADM); along it the plane
// coordinate advances by dz = vx*sgn*zmin/(DIAM/2) per voxel step, i.e. the
// LOR's axial slope is sgn*zmin*pitch/DIAM mm per mm: ring difference
// RD = slope*DIAM/pitch. Plane p is the LOR's axial position at its closest
// point to the axis in half-ring units (p = z/(pitch/2) + (NZS-1)/2).
#include <math.h>
#include <stdlib.h>

#include "gfr_internal.h"

namespace {

// Returns the 0-based bin index or -1 if the LOR is outside the sinogram.
long long bin_lor(const gfr::GeomHost& h, double xa, double ya, double za, double xb, double yb,
                  double zb) {
  const gfr_geometry& g = h.g;
  const double PI = (double)3.1415926f;  // GFIRST.h:1
  double dx = xb - xa, dy = yb - ya;
  const double L = sqrt(dx * dx + dy * dy);
  if (!(L > 1e-9)) return -1;
  dx /= L;
  dy /= L;
  // orient d so that sin(theta) = dx >= 0 (theta in (0, PI])
  if (dx < 0.0 || (dx == 0.0 && dy < 0.0)) {
    dx = -dx;
    dy = -dy;
  }
  double theta = atan2(dx, -dy);  // in [0, PI]
  long long it = llround((PI - theta) / PI * g.nang);
  if (it >= g.nang) it = 0;  // theta ~ 0 is the line of theta = PI
  if (it < 0) it = 0;
  const double th = PI - (double)it * PI / g.nang;
  const double nx = cos(th), ny = sin(th);
  const double ddx = sin(th), ddy = -cos(th);  // the bin's LOR direction
  const double mx = 0.5 * (xa + xb), my = 0.5 * (ya + yb);
  const double s = mx * nx + my * ny;  // signed distance along the bin's normal
  const double px = 0.5 * (double)g.pitch;
  const double RADM = 0.5 * (g.nrad - 1);
  const long long r = llround(RADM + s / px);
  if (r < 0 || r >= g.nrad) return -1;
  // axial slope along the bin's direction (sinth, -costh)
  const double ta = xa * ddx + ya * ddy, tb = xb * ddx + yb * ddy;
  if (!(fabs(tb - ta) > 1e-9)) return -1;
  const double slope = (zb - za) / (tb - ta);
  const double z0 = za - slope * ta;
  const double rd = slope * (double)g.diam_det / (double)g.pitch;
  const long long q = llround(fabs(rd));
  int k;
  if (q <= g.span / 2) {
    k = 0;
  } else {
    k = (int)((q - g.span / 2 - 1) / g.span) + 1;
  }
  int idx = (k == 0) ? 0 : (rd < 0 ? 2 * k - 1 : 2 * k);
  if (idx >= g.nseg) return -1;
  const int zm = h.zmin[idx];
  const long long p = llround(z0 / px + 0.5 * (g.nzs - 1));
  if (p < zm || p >= g.nzs - zm) return -1;
  const long long sino = h.nza[idx] + (p - zm);
  return (sino * g.nang + it) * g.nrad + r;
}

}  // namespace

extern "C" {

long long gfr_lor_bin(const gfr_geometry* g, const double* a, const double* b) {
  gfr::GeomHost h;
  if (!a || !b || gfr::derive_geometry(g, &h, nullptr) != GFR_OK) return GFR_EINVAL;
  return bin_lor(h, a[0], a[1], a[2], b[0], b[1], b[2]);
}

long long gfr_vlor_bin_hits(const gfr_geometry* g, const float* hits, long long nev,
                            float radius, int32_t* rec) {
  gfr::GeomHost h;
  if (!hits || !rec || nev < 0 || !(radius > 0.f) || gfr::derive_geometry(g, &h, nullptr) != GFR_OK)
    return GFR_EINVAL;
  if (h.info.ndata >= (1LL << 31) - 1) return GFR_EINVAL;
  long long valid = 0;
  for (long long e = 0; e < nev; e++) {
    const float* q = hits + 6 * e;
    double P[3][3];
    for (int k = 0; k < 3; k++) {
      P[k][0] = radius * cos((double)q[2 * k]);
      P[k][1] = radius * sin((double)q[2 * k]);
      P[k][2] = (double)q[2 * k + 1];
    }
    const long long b1 = bin_lor(h, P[0][0], P[0][1], P[0][2], P[1][0], P[1][1], P[1][2]);
    const long long b2 = bin_lor(h, P[0][0], P[0][1], P[0][2], P[2][0], P[2][1], P[2][2]);
    rec[2 * e] = b1 >= 0 ? (int32_t)(b1 + 1) : 0;
    rec[2 * e + 1] = b2 >= 0 ? (int32_t)(b2 + 1) : 0;
    if (b1 >= 0 && b2 >= 0) valid++;
  }
  return valid;
}

}  // extern "C"
