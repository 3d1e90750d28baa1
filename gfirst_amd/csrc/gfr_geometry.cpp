// gfr_geometry.cpp — runtime geometry record and the host tables derived from
// it. The reference fixes these at compile time (GFIRST.h:4-16, 31-34); here
// one library serves any scanner/sinogram/image geometry.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>

#include "gfr_internal.h"

namespace gfr {

namespace {
// PI of the reference (GFIRST.h:1).
constexpr float kPI = 3.1415926f;
}  // namespace

int derive_geometry(const gfr_geometry* g, GeomHost* out, std::string* err, bool with_ztab) {
  auto fail = [&](const char* m) {
    if (err) *err = m;
    return GFR_EINVAL;
  };
  if (!g) return fail("geometry is NULL");
  if (g->res < 1 || g->nzs < 1 || g->nrad < 1 || g->nang < 1 || g->nseg < 1 || g->span < 1)
    return fail("geometry: res, nzs, nrad, nang, nseg and span must be >= 1");
  if (g->ntbins != 1) return fail("geometry: ntbins must be 1 (no TOF on this path)");
  if (!(g->pitch > 0.f) || !(g->diam_det > 0.f) || !(g->fov > 0.f))
    return fail("geometry: pitch, diam_det and fov must be > 0");
  if (g->res > 4096 || g->nzs > 4096 || g->nrad > 8192 || g->nang > 8192 || g->nseg > 4096)
    return fail("geometry: dimension out of supported range");
  GeomHost& h = *out;
  h.g = *g;
  h.zmin.assign(g->nseg, 0);
  h.sgn.assign(g->nseg, 0);
  h.nza.assign(g->nseg + 1, 0);
  // Segment order 0,-1,+1,-2,+2,... (GFIRST.cu:163-168); zmin = minimum ring
  // difference of the segment (GFIRST_GPU_Kernels.cu:54); nz_acum
  // (GFIRST.cu:133-140).
  int acc = 0;
  for (int k = 0; k < g->nseg; k++) {
    const int isg = 1 - (k % 2) * 2;
    const int idz = ((k + 1) / 2) * isg;
    const int zm = (idz == 0) ? 0 : (abs(idz) - 1) * g->span + (g->span / 2 + 1);
    const int nz = g->nzs - 2 * zm;
    if (nz <= 0) return fail("geometry: a segment has no planes (nseg too large for nzs/span)");
    h.zmin[k] = zm;
    h.sgn[k] = (idz == 0) ? 0 : (idz > 0 ? 1 : -1);
    acc += nz;
    h.nza[k + 1] = acc;
  }
  h.info.nsinogs = acc;
  h.info.ndata = (long long)acc * g->nang * g->nrad;
  h.info.nvox = (long long)g->res * g->res * g->nzs;
  if (h.info.ndata >= (1LL << 40)) return fail("geometry: sinogram too large");
  // Sample range along a LOR (GFIRST_GPU_Kernels.cu:65-66, 90-91), NPT = RES.
  const float tof = (float)g->res / (float)g->ntbins;
  h.info.lambda0 = (int)floorf((0.f - 0.5f) * tof);
  h.info.lambda1 = (int)floorf((0.f + 0.5f) * tof);
  h.RADM = (float)(g->nrad - 1) * 0.5f;
  h.RESM = (float)(g->res - 1) * 0.5f;
  // Forward: vx = FOV/RES, px = pitch/2, fx = px/vx (GFIRST_GPU_Kernels.cu:58-60).
  const float vx = g->fov / (float)g->res;
  h.fxf = (0.5f * g->pitch) / vx;
  // Backward: px = pitch/2.0 (double), fx = dx/px (GFIRST_GPU_Kernels.cu:127-129).
  h.fxb = vx / (float)((double)g->pitch / 2.0);
  h.seg.resize(g->nseg);
  const float vz = g->fov / (float)g->res;  // FOV/NPT
  for (int k = 0; k < g->nseg; k++) {
    SegEntry& e = h.seg[k];
    const int sz = h.sgn[k] * h.zmin[k];
    e.zmin = h.zmin[k];
    e.nza = h.nza[k];
    e.dzf = (vz * (float)sz) / (0.5f * g->diam_det);                     // :83
    e.fzb = (float)((double)(vx * (float)sz) / ((double)g->diam_det / 2.0));  // :134
  }
  h.sin_t.resize(g->nang);
  h.cos_t.resize(g->nang);
  for (int it = 0; it < g->nang; it++) {
    const float ang = (-(float)it) / (float)g->nang * kPI + kPI;
    h.sin_t[it] = (float)sin((double)ang);
    h.cos_t[it] = (float)cos((double)ang);
  }
  h.sino.resize(acc);
  for (int k = 0; k < g->nseg; k++)
    for (int p = h.zmin[k]; p < g->nzs - h.zmin[k]; p++)
      h.sino[h.nza[k] + p - h.zmin[k]] = make_int2(k, p);
  // z interpolation of every LOR sample: z = zc + l*dz with zc = p + 0.5
  // (GFIRST_GPU_Kernels.cu:88, 96), through the clamp-to-edge linear filter.
  // Same float operations as the device code (fmaf is exactly rounded on both).
  if (g->nzs > 0xFFFF) return fail("geometry: nzs too large");
  if (!with_ztab) return GFR_OK;
  const int nl = h.info.lambda1 - h.info.lambda0 + 1;
  h.ztab.resize((size_t)acc * nl);
  for (int s = 0; s < acc; s++) {
    const float dz = h.seg[h.sino[s].x].dzf;
    const float zc = (float)h.sino[s].y + 0.5f;
    for (int li = 0; li < nl; li++) {
      const float z = fmaf((float)(h.info.lambda0 + li), dz, zc);
      const float t = z - 0.5f;
      const float f = floorf(t);
      const float a = t - f;
      const int i = (int)f;
      const int z0 = i < 0 ? 0 : (i > g->nzs - 1 ? g->nzs - 1 : i);
      const int z1 = i + 1 < 0 ? 0 : (i + 1 > g->nzs - 1 ? g->nzs - 1 : i + 1);
      uint32_t ab;
      memcpy(&ab, &a, 4);
      h.ztab[(size_t)li * acc + s] = make_uint2((uint32_t)z0 | ((uint32_t)z1 << 16), ab);
    }
  }
  return GFR_OK;
}

// calcGaussianCoefficients (GFIRST_GPU_Kernels.cu:322-330). The host exp of a
// float argument is the float overload, i.e. expf.
void psf_taps(float psf, float* g9) {
  float w[9], total = 0.f;
  for (int z = -4; z <= 4; ++z) {
    const float a = (float)z * 1.662f / (psf + 0.1f);
    w[z + 4] = expf(-(a * a));
    total += w[z + 4];
  }
  for (int k = 0; k < 9; k++) g9[k] = w[k] / total;
}

// gaussian_filterZ taps (GFIRST_CPU_Kernels.cpp:169-173).
void gz_taps(float* g5) {
  float w[5], total = 0.f;
  for (int z = -2; z <= 2; ++z) {
    const float a = (float)z * 1.0f;
    w[z + 2] = expf(-(a * a));
    total += w[z + 2];
  }
  for (int k = 0; k < 5; k++) g5[k] = w[k] / total;
}

}  // namespace gfr

extern "C" {

int gfr_geometry_default(gfr_geometry* g) {
  if (!g) return GFR_EINVAL;
  // GFIRST.h:4-16 and the __constant__ scanner values of GFIRST.h:32-34.
  g->res = 128;
  g->nzs = 2 * 80 - 1;
  g->nrad = 147;
  g->nang = 168;
  g->span = 11;
  g->nseg = 15;
  g->ntbins = 1;
  g->pitch = 1.63f;
  g->diam_det = 165.7758f;
  g->fov = 100.0f;
  return GFR_OK;
}

int gfr_geometry_check(const gfr_geometry* g, gfr_geometry_info* info) {
  gfr::GeomHost h;
  int rc = gfr::derive_geometry(g, &h, nullptr);
  if (rc == GFR_OK && info) *info = h.info;
  return rc;
}

int gfr_seg_tables(const gfr_geometry* g, int* zmin, int* sgn, int* nz_acum) {
  gfr::GeomHost h;
  int rc = gfr::derive_geometry(g, &h, nullptr);
  if (rc != GFR_OK) return rc;
  if (zmin) memcpy(zmin, h.zmin.data(), sizeof(int) * g->nseg);
  if (sgn) memcpy(sgn, h.sgn.data(), sizeof(int) * g->nseg);
  if (nz_acum) memcpy(nz_acum, h.nza.data(), sizeof(int) * (g->nseg + 1));
  return h.info.nsinogs;
}

int gfr_angle_tables(const gfr_geometry* g, float* sin_t, float* cos_t) {
  gfr::GeomHost h;
  int rc = gfr::derive_geometry(g, &h, nullptr);
  if (rc != GFR_OK) return rc;
  if (sin_t) memcpy(sin_t, h.sin_t.data(), sizeof(float) * g->nang);
  if (cos_t) memcpy(cos_t, h.cos_t.data(), sizeof(float) * g->nang);
  return GFR_OK;
}

int gfr_psf_taps(float psf, float* g9) {
  if (!g9) return GFR_EINVAL;
  gfr::psf_taps(psf, g9);
  return GFR_OK;
}

int gfr_gz_taps(float* g5) {
  if (!g5) return GFR_EINVAL;
  gfr::gz_taps(g5);
  return GFR_OK;
}

}  // extern "C"
