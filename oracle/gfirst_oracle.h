/*
 * gfirst_oracle.h — CPU oracle for the fully-3D OSEM hot path of GFIRST.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: it is loaded by
 * tests/, by __graft_entry__.smoke() and by the cpu_baseline leg of bench.py,
 * and never by the product library (gfirst_amd/) or its CLI.
 *
 * It restates, from scratch in plain C, the arithmetic of the reference's GPU
 * kernels and host loop (reference file:line cited at each function).  Two
 * variants are built from one source (oracle_core.inc):
 *   orc_*   : FP32, reference-faithful (same promotions to double as the
 *             reference, explicit fmaf where the reference's device code is
 *             contracted, exact FP32 interpolation weights);
 *   orc64_* : FP64 error oracle (same algorithm, every value in double).
 *
 * Layouts (native little-endian, as the reference's raw files):
 *   image    [z][y][x], x fastest, NVOX = res*res*nzs
 *   sinogram [s][theta][r], r fastest, NDATA = nsinogs*nang*nrad
 */
#ifndef GFIRST_ORACLE_H
#define GFIRST_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Runtime geometry; the reference compiles these in (GFIRST.h:4-36). */
typedef struct orc_geom {
  int res;      /* RES   image x/y size (GFIRST.h:14)            */
  int nzs;      /* NZS   image slices == sinogram planes (:15)   */
  int nrad;     /* NRAD  radial bins (:5)                        */
  int nang;     /* NANG  angles over 180 deg (:6)                */
  int span;     /* SPAN  (:8)                                    */
  int nseg;     /* NSEGMENTS (:10)                               */
  int ntbins;   /* NTBINS (:4), 1                                */
  float pitch;  /* mm (:32)                                      */
  float diam_det; /* mm (:33)                                    */
  float fov;    /* mm (:34)                                      */
} orc_geom;

/* Reconstruction parameters (GFIRST.h:20-30, GFIRST.cu:512-517). */
typedef struct orc_params {
  int mode;        /* 1 doubles, 2 doubles+triples sinograms, 3 + VLOR   */
  int niter;       /* iterations                                          */
  int nsub;        /* subsets (1 reproduces the reference)                */
  double psf_fw;   /* PSF_FW (GFIRST.h:26)                                */
  double psf_bw;   /* PSF_BW (GFIRST.h:27)                                */
  int jbf_r;       /* 4     (GFIRST.cu:512)                               */
  double jbf_sigma3d; /* 0.75 (GFIRST.cu:516)                             */
  double jbf_sigmai;  /* 0.025 (GFIRST.cu:517)                            */
  double jbf_coef;    /* 0.1  (GFIRST.cu:514)                             */
} orc_params;

/* ---- geometry (integer; shared by both variants) ---- */
int orc_nsinogs(const orc_geom* g);             /* -1 if invalid */
long long orc_ndata(const orc_geom* g);
long long orc_nvox(const orc_geom* g);
/* zmin[nseg], sgn[nseg], nz_acum[nseg+1] in reference segment order
 * 0,-1,+1,-2,+2,... (GFIRST.cu:133-140, 163-168). Returns nsinogs. */
int orc_seg_tables(const orc_geom* g, int* zmin, int* sgn, int* nz_acum);
void orc_set_threads(int n);
int orc_get_threads(void);

/* ---- FP32 reference-faithful ---- */
void orc_angle_tables(const orc_geom* g, float* sin_t, float* cos_t);
void orc_psf_taps(float psf, float* g9);
void orc_gz_taps(float* g5);
void orc_cylinder(const orc_geom* g, float val, float* img);
void orc_tex3(const float* v, int W, int H, int D, const float* xyz, float* out, long long n);
void orc_blur_xy(const orc_geom* g, float psf, const float* in, float* out);
void orc_forward(const orc_geom* g, const float* img, float* sino, int subset, int nsub);
void orc_calib(const orc_geom* g, const float* proj_blank, const float* blank, float* calib);
void orc_ratio(const orc_geom* g, const float* proj, const float* y, const float* bg,
               float* corr, int subset, int nsub);
void orc_back(const orc_geom* g, const float* corr, const float* calib,
              float* numer, float* denom, int subset, int nsub);
void orc_update(const orc_geom* g, float psf_bw, const float* numer, const float* denom,
                float* x);
void orc_jbf(const orc_geom* g, const float* in, const float* guide, float max_ref, int r,
             float sigma3d, float sigmai, float* out);
void orc_blend(const orc_geom* g, float coef, const float* filt, float* x);
float orc_max(const float* v, long long n);
void orc_vlor_precount(const orc_geom* g, const int32_t* rec, long long nrec, float* v0);
void orc_vlor_blankt(const orc_geom* g, const float* yT, const float* v0, float* blankt);
void orc_vlor_assign(const orc_geom* g, const int32_t* rec, long long nrec,
                     const float* blankt, const float* proj, float* dv, int subset, int nsub);
void orc_output(const orc_geom* g, float scale, const float* x, float* out);
/* FP64 Poisson log-likelihood of a pass over the subset's bins that enter the
 * update (calib > 0, p + max(bg,0) >= 0.01f): sum (y/RES) ln(den/RES) - den/RES
 * with y, bg RES-scaled. The reference computes no likelihood. */
double orc_loglik(const orc_geom* g, const float* proj, const float* y, const float* bg,
                  const float* calib, int subset, int nsub, double* abs_sum);
int orc_osem(const orc_geom* g, const orc_params* p, const float* d2, const float* t2,
             const float* norm, const int32_t* rec, long long nrec,
             float* xD, float* xT, float* proj_blank, float* vlor_out);

/* ---- FP64 error oracle ---- */
void orc64_angle_tables(const orc_geom* g, double* sin_t, double* cos_t);
void orc64_psf_taps(double psf, double* g9);
void orc64_gz_taps(double* g5);
void orc64_cylinder(const orc_geom* g, double val, double* img);
void orc64_blur_xy(const orc_geom* g, double psf, const double* in, double* out);
void orc64_forward(const orc_geom* g, const double* img, double* sino, int subset, int nsub);
void orc64_calib(const orc_geom* g, const double* proj_blank, const double* blank, double* calib);
void orc64_ratio(const orc_geom* g, const double* proj, const double* y, const double* bg,
                 double* corr, int subset, int nsub);
void orc64_back(const orc_geom* g, const double* corr, const double* calib,
                double* numer, double* denom, int subset, int nsub);
void orc64_update(const orc_geom* g, double psf_bw, const double* numer, const double* denom,
                  double* x);
void orc64_jbf(const orc_geom* g, const double* in, const double* guide, double max_ref, int r,
               double sigma3d, double sigmai, double* out);
int orc64_osem(const orc_geom* g, const orc_params* p, const float* d2, const float* t2,
               const float* norm, const int32_t* rec, long long nrec,
               double* xD, double* xT, double* proj_blank, double* vlor_out);

/* ---- raw-file semantics (GFIRST_CPU_Kernels.cpp:104-142, GFIRST.cu:229-240) ---- */
/* Decodes a sinogram file image of nwords 32-bit words into out[2*ndata]
 * (prompts, background). Returns 0 int path, 1 float path. */
int orc_decode_sinogram(const uint32_t* words, long long nwords, long long ndata, float* out);
/* Expands a VLOR file of nwords int32 into the record sequence the reference's
 * while(!feof) loop processes (last record twice). rec_out holds 2*(nwords/2+1)
 * ints. Returns the number of records. */
long long orc_vlor_records(const int32_t* words, long long nwords, int32_t* rec_out);

#ifdef __cplusplus
}
#endif
#endif
