/*
 * gfirst_oracle.c — CPU oracle for GFIRST's fully-3D OSEM hot path.
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker, never by the product.
 *
 * Built with -ffp-contract=off so the only fused multiply-adds are the
 * explicit FMA() calls of oracle_core.inc.
 */
#include "gfirst_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* PI of the reference (GFIRST.h:1). */
#define ORC_PI 3.1415926f

void orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

int orc_get_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* Segment table (GFIRST.cu:133-140 for nz_acum; :163-168 for the signed
 * segment order 0,-1,+1,-2,+2,...; zmin as GFIRST_GPU_Kernels.cu:54). */
int orc_seg_tables(const orc_geom* g, int* zmin, int* sgn, int* nz_acum) {
  if (!g || g->nseg < 1 || g->nzs < 1) return -1;
  int acc = 0;
  if (nz_acum) nz_acum[0] = 0;
  for (int k = 0; k < g->nseg; k++) {
    int isg = 1 - (k % 2) * 2;
    int idz = ((k + 1) / 2) * isg;
    int zm = (idz == 0) ? 0 : (abs(idz) - 1) * g->span + (g->span / 2 + 1);
    int nz = g->nzs - 2 * zm;
    if (nz <= 0) return -1;
    if (zmin) zmin[k] = zm;
    if (sgn) sgn[k] = (idz == 0) ? 0 : (idz > 0 ? 1 : -1);
    acc += nz;
    if (nz_acum) nz_acum[k + 1] = acc;
  }
  return acc;
}

int orc_nsinogs(const orc_geom* g) { return orc_seg_tables(g, NULL, NULL, NULL); }

long long orc_ndata(const orc_geom* g) {
  int ns = orc_nsinogs(g);
  if (ns <= 0) return -1;
  return (long long)ns * g->nang * g->nrad;
}

long long orc_nvox(const orc_geom* g) { return (long long)g->res * g->res * g->nzs; }

/* ReadSinogram_Type (GFIRST_CPU_Kernels.cpp:104-142): the file is read as
 * int32 into a zeroed 2*NDATA buffer; if the max of that int view exceeds
 * 1e6 the words are reinterpreted as float32. Words beyond 2*NDATA are
 * dropped (the reference overflows its buffer there). */
int orc_decode_sinogram(const uint32_t* words, long long nwords, long long ndata, float* out) {
  long long n = 2 * ndata;
  long long m = nwords < n ? nwords : n;
  float mx = 0.f;
  for (long long i = 0; i < n; i++) {
    float v = (i < m) ? (float)(int32_t)words[i] : 0.f;
    out[i] = v;
    if (v > mx) mx = v;
  }
  if (mx > 1.0e6) {
    for (long long i = 0; i < m; i++) memcpy(&out[i], &words[i], 4);
    return 1;
  }
  return 0;
}

/* while(!feof) { fread(VLOR,4,2); ... } (GFIRST.cu:233-240, 376-388): every
 * full record once, then one more pass with whatever the last fread left in
 * VLOR: the last record again for an even word count, or (last word, previous
 * second word) for an odd one. An empty file yields no usable record. */
long long orc_vlor_records(const int32_t* words, long long nwords, int32_t* rec_out) {
  long long nfull = nwords / 2, k = 0;
  int32_t a = 0, b = 0;
  for (long long e = 0; e < nfull; e++) {
    a = words[2 * e];
    b = words[2 * e + 1];
    rec_out[2 * k] = a;
    rec_out[2 * k + 1] = b;
    k++;
  }
  if (nwords == 0) return 0;
  if (nwords % 2) a = words[nwords - 1];
  rec_out[2 * k] = a;
  rec_out[2 * k + 1] = b;
  return k + 1;
}

/* ---------------- FP32 reference-faithful variant ---------------- */
#define R float
#define F(n) orc_##n
#define FMA(a, b, c) fmaf((a), (b), (c))
#define FLOOR(x) floorf(x)
#define EXP(x) expf(x)
#define ORC_FP32 1
#include "oracle_core.inc"
#undef R
#undef F
#undef FMA
#undef FLOOR
#undef EXP
#undef ORC_FP32

void orc_tex3(const float* v, int W, int H, int D, const float* xyz, float* out, long long n) {
  for (long long i = 0; i < n; i++)
    out[i] = orc_tex3s(v, W, H, D, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
}

void orc_blend(const orc_geom* g, float coef, const float* filt, float* x) {
  orc_blend_r(orc_nvox(g), coef, filt, x);
}

float orc_max(const float* v, long long n) { return orc_max_r(v, n); }

void orc_vlor_precount(const orc_geom* g, const int32_t* rec, long long nrec, float* v0) {
  long long nd = orc_ndata(g);
  for (long long i = 0; i < nd; i++) v0[i] = 0.f;
  for (long long e = 0; e < nrec; e++) {
    long long i1 = (long long)rec[2 * e] - 1, i2 = (long long)rec[2 * e + 1] - 1;
    if (i1 <= 0 || i2 <= 0 || i1 >= nd || i2 >= nd) continue;
    v0[i1] += 1.f;
    v0[i2] += 1.f;
  }
}

void orc_vlor_blankt(const orc_geom* g, const float* yT, const float* v0, float* blankt) {
  long long nd = orc_ndata(g);
  float res = (float)g->res;
  for (long long i = 0; i < nd; i++)
    blankt[i] = ((double)v0[i] > 0.01) ? yT[i] / (res * v0[i]) : 0.f;
}

void orc_vlor_assign(const orc_geom* g, const int32_t* rec, long long nrec, const float* blankt,
                     const float* proj, float* dv, int subset, int nsub) {
  orc_vlor_assign_r(g, rec, nrec, blankt, proj, dv, subset, nsub);
}

/* Output image (GFIRST.cu:532-549): x*scale, zero outside the circle of
 * radius RES/2-1 around (RES/2, RES/2). */
void orc_output(const orc_geom* g, float scale, const float* x, float* out) {
  int res = g->res, h = res / 2;
  long long i = 0;
  for (int z = 0; z < g->nzs; z++)
    for (int y = 0; y < res; y++)
      for (int xx = 0; xx < res; xx++, i++) {
        float dist = (float)((xx - h) * (xx - h) + (y - h) * (y - h) - (h - 1) * (h - 1));
        out[i] = (dist > 0.f) ? 0.f : x[i] * scale;
      }
}

/* FP64 Poisson log-likelihood of one (type, subset) pass. The reference
 * computes none (it only sums prompts, projection and background in float,
 * GFIRST.cu:413-422); this is the build's definition, stated here as the
 * checker of the HIP path's reduction:
 *   LL = sum over the subset's bins b with calib[b] > 0 and
 *        den_b = p_b + max(bg_b, 0) >= 0.01f (the bins whose ratio enters the
 *        backprojection, GFIRST.cu:423-429, 453-458)
 *        of  (y_b/RES) * ln(den_b/RES) - den_b/RES,
 * i.e. the Poisson log-likelihood (without the constant -ln y!) of the
 * measured counts under the model mean (p + RES*bg)/RES, in count units.
 * y and bg are RES-scaled as the ratio sees them (GFIRST.cu:224-260); den is
 * the ratio's FP32 sum; every other operation is FP64. Bins are summed in
 * sinogram order; *abs_sum (if not NULL) receives the sum of |term|. */
double orc_loglik(const orc_geom* g, const float* proj, const float* y, const float* bg,
                  const float* calib, int subset, int nsub, double* abs_sum) {
  long long nsino = orc_nsinogs(g);
  int nang = g->nang, nrad = g->nrad;
  double res = (double)g->res, ll = 0.0, la = 0.0;
  for (long long s = 0; s < nsino; s++) {
    for (int it = subset; it < nang; it += nsub) {
      size_t b = ((size_t)s * nang + it) * nrad;
      for (int r = 0; r < nrad; r++) {
        float vb = bg ? bg[b + r] : 0.f;
        if (vb < 0.f) vb = 0.f;
        float den = proj[b + r] + vb;
        if (!(den >= 0.01f) || !(calib[b + r] > 0.f)) continue;
        double m = (double)den / res;
        double t = ((double)y[b + r] / res) * log(m) - m;
        ll += t;
        la += fabs(t);
      }
    }
  }
  if (abs_sum) *abs_sum = la;
  return ll;
}

/* ---------------- FP64 error-oracle variant ---------------- */
#define R double
#define F(n) orc64_##n
#define FMA(a, b, c) fma((a), (b), (c))
#define FLOOR(x) floor(x)
#define EXP(x) exp(x)
#define ORC_FP32 0
#include "oracle_core.inc"
#undef R
#undef F
#undef FMA
#undef FLOOR
#undef EXP
#undef ORC_FP32
