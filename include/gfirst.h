/*
 * gfirst.h — C ABI of the MI355X-native fully-3D OSEM hot path (libgfirst.so).
 *
 * The reference (GFIRST) has no library boundary: its path is a set of CUDA
 * host wrappers that share global state (texture references, cudaArrays,
 * __constant__ taps) and exit the process on error.  Each entry point below
 * replaces one of those wrappers (reference file:line given per function);
 * GFIRST_GPU_Kernels.cu is abbreviated "GPU:", GFIRST.cu "DRV:" and
 * GFIRST_CPU_Kernels.cpp "CPU:".
 *
 * Conventions
 *  - Every entry returns an int status: 0 ok, <0 error (GFR_E*); nothing exits.
 *    The message of the last error on a context is gfr_last_error(ctx).
 *  - Geometry travels by value (pointer to a caller-owned record) in every
 *    call, so one context serves studies of different scanner geometries.
 *  - Device buffers are caller-owned (gfr_alloc/gfr_free); calls taking a
 *    `stream` argument are asynchronous and ordered on that HIP stream
 *    (hipStream_t passed as void*, NULL = the context's own stream).
 *  - A context may be used by one host thread at a time.
 *  - Layouts: image [z][y][x] (x fastest, NVOX = res*res*nzs floats);
 *    sinogram [s][theta][r] (r fastest, NDATA = nsinogs*nang*nrad), with
 *    sinograms ordered by segment 0,-1,+1,-2,+2,... then plane (DRV:133-140).
 *  - Subsets: subset s of S holds the angles theta with theta % S == s, all
 *    segments; S = 1 is the reference's full MLEM pass.
 */
#ifndef GFIRST_H
#define GFIRST_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GFR_OK 0
#define GFR_EINVAL -1    /* bad argument or geometry                */
#define GFR_EHIP -2      /* HIP runtime error                       */
#define GFR_ENOMEM -3    /* allocation failed or over the HBM budget */
#define GFR_EIO -4       /* file missing/unreadable                 */
#define GFR_ESTATE -5    /* call not valid in the object's state     */
#define GFR_ECANCELED -6 /* request cancelled (gfr_server_cancel)    */

/* Scanner / sinogram / image geometry. The reference compiles these in
 * (GFIRST.h:4-16, 31-34); gfr_geometry_default() returns those values. */
typedef struct gfr_geometry {
  int res;        /* RES: image x/y size                                  */
  int nzs;        /* NZS: image slices == direct sinogram planes          */
  int nrad;       /* NRAD: radial bins                                    */
  int nang;       /* NANG: angles over 180 degrees                        */
  int span;       /* SPAN                                                 */
  int nseg;       /* NSEGMENTS (odd: 0, -1, +1, ...)                      */
  int ntbins;     /* NTBINS: must be 1 (no TOF on this path)              */
  float pitch;    /* crystal pitch, mm                                    */
  float diam_det; /* detector diameter, mm                                */
  float fov;      /* transaxial field of view, mm                         */
} gfr_geometry;

typedef struct gfr_geometry_info {
  int nsinogs;      /* NSINOGS derived from the segment table             */
  long long ndata;  /* nsinogs*nang*nrad                                  */
  long long nvox;   /* res*res*nzs                                        */
  int lambda0;      /* first sample index along a LOR (GPU:90)            */
  int lambda1;      /* last sample index (inclusive, GPU:91)              */
} gfr_geometry_info;

/* Storage precision of measured sinograms inside a reconstruction. */
#define GFR_STORE_F32 0
#define GFR_STORE_F16 1
#define GFR_STORE_BF16 2

/* Reconstruction parameters (GFIRST.h:20-30; DRV:512-517). */
typedef struct gfr_recon_params {
  int mode;          /* 1 doubles; 2 doubles+triples sinograms; 3 triples from VLORs */
  int niter;         /* iterations (reference NITER = 60)                  */
  int nsub;          /* subsets S (1 = reference)                          */
  float psf_fw;      /* PSF_FW, voxels                                     */
  float psf_bw;      /* PSF_BW, voxels                                     */
  int jbf_r;         /* joint bilateral radius (4)                         */
  float jbf_sigma3d; /* 0.75                                               */
  float jbf_sigmai;  /* 0.025                                              */
  float jbf_coef;    /* blend coefficient (0.1)                            */
  int storage;       /* GFR_STORE_*: sinogram storage in HBM               */
} gfr_recon_params;

typedef struct gfr_ctx gfr_ctx;
typedef struct gfr_vlor gfr_vlor;
typedef struct gfr_recon gfr_recon;

/* ---------------- geometry (host only; no device needed) ---------------- */

/* Compiled constants of the reference (GFIRST.h:4-16, 31-34). */
int gfr_geometry_default(gfr_geometry* g);
/* Validates g and derives sizes (replaces the hand-set NSINOGS, GFIRST.h:9). */
int gfr_geometry_check(const gfr_geometry* g, gfr_geometry_info* info);
/* Segment table: zmin, sign and cumulative sinogram offsets (nseg+1) in the
 * reference's segment order (DRV:133-140, 163-168; GPU:54). Returns nsinogs. */
int gfr_seg_tables(const gfr_geometry* g, int* zmin, int* sgn, int* nz_acum);
/* Per-angle sin/cos of theta = -it/NANG*PI + PI (GPU:75-77, 139-141), from
 * libm in double rounded to float. */
int gfr_angle_tables(const gfr_geometry* g, float* sin_t, float* cos_t);
/* 9-tap PSF coefficients (GPU:322-330, calcGaussianCoefficients). */
int gfr_psf_taps(float psf, float* g9);
/* 5-tap z coefficients (CPU:169-173, gaussian_filterZ). */
int gfr_gz_taps(float* g5);

/* ---------------- context and memory ---------------- */

/* Replaces the implicit CUDA context + crearTexturas() (GPU:254-281) and the
 * cudaMalloc block of osem3d (DRV:62-83). hbm_budget = 0: no limit. */
int gfr_ctx_create(int device, size_t hbm_budget, gfr_ctx** out);
int gfr_ctx_destroy(gfr_ctx* ctx);
/* Replaces chequearError (GPU:305-320): returns the message, never exits. */
const char* gfr_last_error(gfr_ctx* ctx);
void* gfr_ctx_stream(gfr_ctx* ctx);
int gfr_sync(gfr_ctx* ctx, void* stream);
int gfr_alloc(gfr_ctx* ctx, size_t bytes, void** dptr);
int gfr_free(gfr_ctx* ctx, void* dptr);
/* Bytes currently allocated through gfr_alloc / reconstructions. */
size_t gfr_bytes_in_use(gfr_ctx* ctx);
/* Replaces UpdateImage / actualizarCorr / cudaMemcpy (GPU:221-252). */
int gfr_memcpy_h2d(gfr_ctx* ctx, void* dst, const void* src, size_t bytes, void* stream);
int gfr_memcpy_d2h(gfr_ctx* ctx, void* dst, const void* src, size_t bytes, void* stream);
/* Zero fill with a kernel (CleanMemory, GPU:207-219). */
int gfr_zero(gfr_ctx* ctx, void* dptr, size_t bytes, void* stream);

/* Launch shapes of the projectors for (geometry, nsub). Results do not
 * depend on them. A reconstruction session measures them once per context
 * (env GFR_AUTOTUNE=0 keeps the defaults); per-operation calls use the
 * context's current choice. No counterpart in the reference (its launch
 * shapes are fixed, GPU:105-107, 166-171). */
typedef struct gfr_tuning {
  int back_zb;    /* slices per backprojector thread: 1, 2 (back_lds 0), 4, 5, 6,
                     8 or 16                                                    */
  int back_lds;   /* 1: per-block LDS window of each view's rows; 0: cached reads */
  int back_waves; /* 6: 8-slice direct kernel built for 6 waves/SIMD; else 0     */
  int zsum_sg;    /* sinograms per forward z-sum wave: 4, 6 or 8                 */
  int rot_z;      /* planes per forward rotation block: 4, 16 or 32             */
  int back_bw;    /* waves per backprojector block: 4 (16x16 columns) or 1 (8x8);
                     1 only with back_lds 0 and back_waves 0                    */
  float ms_back;  /* measured time of the chosen backprojector (0 if not measured) */
  float ms_zsum;  /* measured time of the chosen forward projector              */
  int back_ahead; /* views whose rows are loaded before the first is summed: 1 (or
                     0), 2 or 4; above 1 only with back_zb 1, back_lds 0 and
                     back_waves 0                                               */
} gfr_tuning;
/* Returns 1 if the shapes were measured, 0 if defaults. */
int gfr_tuning_get(gfr_ctx* ctx, const gfr_geometry* g, int nsub, gfr_tuning* t);
int gfr_tuning_set(gfr_ctx* ctx, const gfr_geometry* g, int nsub, const gfr_tuning* t);

/* ---------------- per-operation entry points (stream ordered) ---------------- */

/* CreateInitialCylinder (CPU:35-56) into a device image. */
int gfr_fill_cylinder(gfr_ctx* ctx, const gfr_geometry* g, float val, float* d_img, void* stream);

/* convolutionXY with calcGaussianCoefficients(psf) (GPU:322-364): per plane a
 * 9-tap y filter (zero padded) then a 9-tap x filter for 4<=x<=RES-5; other
 * columns keep the raw input. d_in != d_out. */
int gfr_blur_xy(gfr_ctx* ctx, const gfr_geometry* g, float psf, const float* d_in, float* d_out,
                void* stream);

/* Forward projection of every view of subset `subset` (FORWPROJ + host
 * scatter, GPU:40-107, DRV:346-363): writes the subset's bins of d_sino
 * (NDATA layout); other bins are untouched. */
int gfr_forward_project(gfr_ctx* ctx, const gfr_geometry* g, const float* d_img, float* d_sino,
                        int subset, int nsub, void* stream);

/* Calibration mask (DRV:453-458): mask = 1 where PROJ_BLANK>0.01 and
 * norm>0.01, else 0. d_mask is NDATA bytes. */
int gfr_calib_mask(gfr_ctx* ctx, const gfr_geometry* g, const float* d_proj_blank,
                   const float* d_norm, uint8_t* d_mask, void* stream);

/* Ratio (DRV:413-430) over the subset's bins: corr = y/(p+max(bg,0)) if
 * p+bg >= 0.01 else -1000. d_bg may be NULL (no background). */
int gfr_ratio(gfr_ctx* ctx, const gfr_geometry* g, const float* d_proj, const float* d_y,
              const float* d_bg, float* d_corr, int subset, int nsub, void* stream);

/* Fused forward projection + ratio: writes d_corr (and the projection into
 * d_proj when non-NULL) for the subset's bins. Same results as
 * gfr_forward_project followed by gfr_ratio. */
int gfr_forward_ratio(gfr_ctx* ctx, const gfr_geometry* g, const float* d_img, const float* d_y,
                      const float* d_bg, float* d_corr, float* d_proj, int subset, int nsub,
                      void* stream);

/* Backprojection of the subset's views (per-view loop of DRV:436-469 with
 * BACKWPROJ, GPU:109-171): d_numer/d_denom receive the complete subset sums
 * (the reference's CleanMemory + per-view accumulation), voxel by voxel in
 * the reference's segment-major, angle-minor order. */
int gfr_back_project(gfr_ctx* ctx, const gfr_geometry* g, const float* d_corr,
                     const uint8_t* d_mask, float* d_numer, float* d_denom, int subset, int nsub,
                     void* stream);

/* gfr_back_project without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops): no
 * counterpart in the reference. The subset's packed rows go to a fixed 64 MiB
 * device array of the library (one per device, held until the process ends)
 * instead of context scratch; the kernels and results are gfr_back_project's.
 * Limits: the packed rows, 8*(NSINOGS + 2*NSEG)*(NRAD + 2) bytes per angle of
 * the subset, must fit 64 MiB (at the Inveon geometry: 16 or more subsets),
 * else GFR_EINVAL; the context must already hold the geometry's tables (any
 * earlier call with this geometry on this context creates them), else
 * GFR_ESTATE; the context's device must be current. Calls in flight at once
 * on one device share the array and race: order them on one stream. */
int gfr_back_project_noalloc(gfr_ctx* ctx, const gfr_geometry* g, const float* d_corr,
                             const uint8_t* d_mask, float* d_numer, float* d_denom, int subset,
                             int nsub, void* stream);

/* Sensitivity image of a subset: the calibration half of BACKWPROJ
 * (GPU:112-160 with CALIB_PART, DRV:447-467; ratio taken as 1, so only the
 * calibration and segment-edge gaps apply), accumulated over views in FP64.
 * d_sens: NVOX doubles, overwritten. */
int gfr_sensitivity(gfr_ctx* ctx, const gfr_geometry* g, const uint8_t* d_mask, double* d_sens,
                    int subset, int nsub, void* stream);
/* Image update (DRV:482-502; GPU:173-188; CPU:165-187): R = DENOM<0.001 ? 1 :
 * NUMER/DENOM; F = Gz(Gxy_psf_bw(R)); x *= F; slice 0 := slice 1 and slice
 * NZS-1 := slice NZS-2. d_work is NVOX floats of scratch. In place on d_x. */
int gfr_update(gfr_ctx* ctx, const gfr_geometry* g, float psf_bw, const float* d_numer,
               const float* d_denom, float* d_x, float* d_work, void* stream);

/* jointbilateralfilter3D (GPU:508-544). */
int gfr_jbf3d(gfr_ctx* ctx, const gfr_geometry* g, const float* d_in, const float* d_guide,
              float max_ref, int r, float sigma3d, float sigmai, float* d_out, void* stream);
/* Triples blend x = (1-coef)*x + coef*filtered (DRV:524). */
int gfr_blend(gfr_ctx* ctx, const gfr_geometry* g, float coef, const float* d_filt, float* d_x,
              void* stream);
/* max(0, max_i x[i]) (DRV:521); synchronous. */
int gfr_image_max(gfr_ctx* ctx, const gfr_geometry* g, const float* d_x, float* h_max,
                  void* stream);

/* VLOR soft assignment (DRV:229-247 pre-count and BLANKT; DRV:373-388 per
 * pass). `rec` is the record sequence the reference's read loop processes
 * (pairs of 1-based bin indices, last record already repeated; see
 * gfr_read_vlor_file). d_yT = triples prompts already scaled by RES. */
int gfr_vlor_create(gfr_ctx* ctx, const gfr_geometry* g, const int32_t* rec, long long nrec,
                    const float* d_yT, gfr_vlor** out);
/* d_dv[b] for the subset's bins b: sum over records touching b, in record
 * order, of proj[b]/(proj[i1]+proj[i2])*(BLANKT[i1]+BLANKT[i2]). */
int gfr_vlor_assign(gfr_ctx* ctx, gfr_vlor* v, const float* d_proj, float* d_dv, int subset,
                    int nsub, void* stream);
int gfr_vlor_destroy(gfr_vlor* v);

/* ---------------- raw files (host; CPU:104-142, DRV:229-240) ---------------- */

/* ReadSinogram_Type: out = 2*ndata floats (prompts, background); int32 file
 * unless the int view's max exceeds 1e6, then float32. A missing file gives
 * zeros and returns 1 (the reference prints "Sinogram Not Found!"); words
 * beyond 2*ndata are ignored and reported by *truncated. Returns 0/1 or <0. */
int gfr_read_sinogram_file(const char* path, long long ndata, float* out, int* is_float,
                           int* truncated);
/* VLOR file -> record sequence (last record repeated as by while(!feof)).
 * Call with rec=NULL to get the count; returns count or <0. */
long long gfr_read_vlor_file(const char* path, int32_t* rec, long long cap);

/* ---------------- list-mode binning (host; no counterpart in the reference,
 * which takes VLORs already binned, DRV:213-236) ---------------- */

/* Bin of the LOR through points a and b (mm, scanner frame, z axial, origin at
 * the centre) in the sinogram convention of the forward projector (GPU:75-96):
 * angle index by the LOR direction, radial bin by its distance to the axis in
 * units of pitch/2, segment by its ring difference (axial slope x DIAM/pitch)
 * and plane by its axial position at the point closest to the axis. Returns
 * the 0-based bin or -1 if the LOR lies outside the sinogram. */
long long gfr_lor_bin(const gfr_geometry* g, const double* a, const double* b);
/* Triple events as 3 hits (phi, z) on a cylinder of the given radius (6 floats
 * per event; hit 0 is the shared vertex): rec[2e], rec[2e+1] = 1-based bins of
 * the LORs (hit0, hit1) and (hit0, hit2), 0 where a LOR is outside the
 * sinogram (the VLORs.raw record format). Returns the number of events with
 * both LORs inside. */
long long gfr_vlor_bin_hits(const gfr_geometry* g, const float* hits, long long nev, float radius,
                            int32_t* rec);

/* ---------------- whole reconstruction (device-resident session) ---------------- */

/* osem3d (DRV:44-575) as a session: uploads the data once, keeps every
 * array resident in HBM, and advances one (type, subset) pass per step.
 * d2/t2: 2*ndata floats as read from file (unscaled); t2 NULL in mode 1;
 * norm: ndata floats; d2/t2/norm may be host or device memory (device data
 * complete before the call); rec/nrec: VLOR records in host memory (mode 3).
 * Admission checks the context's HBM budget. A NaN or infinite value in the
 * measured sinograms (prompts, background) fails with GFR_EINVAL, where the
 * reference would carry it into the image (DRV:423-426); in the
 * normalization it masks the bin, as in the reference (DRV:453-458). */
int gfr_recon_create(gfr_ctx* ctx, const gfr_geometry* g, const gfr_recon_params* p,
                     const float* d2, const float* t2, const float* norm, const int32_t* rec,
                     long long nrec, gfr_recon** out);
/* The whole of osem3d in one call (create, run, copy the images to the
 * host, destroy): xD (and xT in modes 2-3) receive NVOX floats. */
int gfr_osem_run(gfr_ctx* ctx, const gfr_geometry* g, const gfr_recon_params* p, const float* d2,
                 const float* t2, const float* norm, const int32_t* rec, long long nrec,
                 float* xD, float* xT);
/* Device bytes a reconstruction with these parameters holds. */
long long gfr_recon_footprint(const gfr_geometry* g, const gfr_recon_params* p);
/* One (type, subset) pass; returns 1 when the reconstruction is complete. */
int gfr_recon_step(gfr_recon* r);
/* Run the remaining passes. */
int gfr_recon_run(gfr_recon* r);
/* Position: iteration (1-based), type (0 D, 1 T), subset; total passes done. */
int gfr_recon_cursor(gfr_recon* r, int* iter, int* type, int* subset, long long* passes);
/* Copy state to the host: what = 0 doubles image, 1 triples image,
 * 2 PROJ_BLANK (NDATA), 3 VLOR pseudo-sinogram before scaling (NDATA). */
#define GFR_OUT_XD 0
#define GFR_OUT_XT 1
#define GFR_OUT_PROJ_BLANK 2
#define GFR_OUT_VLOR 3
int gfr_recon_get(gfr_recon* r, int what, float* host_out);
/* FP64 Poisson log-likelihood of the last pass, in count units: the sum over
 * the subset's bins whose ratio enters the update (valid calibration and
 * den = p + max(bg, 0) >= 0.01, DRV:423-429, 453-458) of
 * (y/RES) ln(den/RES) - den/RES, with y, bg RES-scaled as the ratio sees
 * them (the constant -ln y! omitted). The reference computes no likelihood
 * (DRV:413-422 only sums prompts, projection and background). Passes that
 * fit the VLOR pseudo-sinogram (triples in mode 3) have none: GFR_ESTATE. */
int gfr_recon_loglik(gfr_recon* r, double* out);
/* The session's stream. In the joint modes the triples passes run on a
 * second stream of the session (beside the next iteration's doubles
 * passes) that joins this one when the last pass is enqueued. The context
 * keeps the streams of destroyed sessions for later sessions: do not use it
 * after gfr_recon_destroy. */
void* gfr_recon_stream(gfr_recon* r);
/* Waits for every pass enqueued so far (both streams of a joint session). */
int gfr_recon_sync(gfr_recon* r);
int gfr_recon_destroy(gfr_recon* r);

/* Device-time tracing per phase of a pass (HIP events on the session's
 * stream; the reference times only with time() at 1 s resolution,
 * DRV:51-52, 572-573). Each phase is one kernel launch except UPDATE (two). */
#define GFR_PHASE_BLUR 0    /* PSF-FW blur                               */
#define GFR_PHASE_FORWARD 1 /* forward projection (+ ratio, FP64 LLH)    */
#define GFR_PHASE_VLOR 2    /* VLOR reassignment + ratio (mode 3)        */
#define GFR_PHASE_BACK 3    /* backprojection                            */
#define GFR_PHASE_UPDATE 4  /* ratio blur, z filter, multiply            */
#define GFR_PHASE_JBF 5     /* max + joint bilateral filter + blend       */
#define GFR_NPHASES 6
int gfr_recon_profile(gfr_recon* r, int enable);
/* Waits for the session's stream, then returns the accumulated milliseconds
 * and pass counts per phase since the last call (arrays of GFR_NPHASES). */
int gfr_recon_phase_times(gfr_recon* r, double* ms, long long* counts);

/* ---------------- batched frames (one graph launch per subset step) ---------------- */

/* Many doubles reconstructions (mode 1 of DRV:44-575) of one geometry, e.g.
 * the frames of a dynamic study, each with its own data, subset count and
 * iteration count. A step advances every active frame by one subset pass;
 * the whole step is one captured hipGraph (rebuilt only when frames join or
 * finish), with frames spread over `lanes` independent chains of scratch so
 * their kernels overlap on the device. Each frame's result is bit-identical
 * to a gfr_recon session with the same data and parameters. The reference
 * runs one study per process. */
typedef struct gfr_batch gfr_batch;
typedef struct gfr_batch_params {
  int storage;     /* GFR_STORE_* of every frame's sinograms */
  int max_frames;  /* frame slots */
  int min_nsub;    /* smallest subset count a frame may use (sizes the scratch) */
  int lanes;       /* independent scratch chains, >= 1 */
  float psf_fw, psf_bw;
  int group;       /* frames at the same subset of the same subset count that
                      share one backprojection launch: 1 (none), 2 or 3; 0 = 3.
                      Results do not depend on it. */
} gfr_batch_params;
/* Device bytes of the batch without frames, and of one frame. */
long long gfr_batch_footprint(const gfr_geometry* g, const gfr_batch_params* p);
long long gfr_batch_frame_footprint(const gfr_geometry* g, const gfr_batch_params* p, int has_bg);
/* norm: ndata floats (host); the calibration mask is shared by all frames. */
int gfr_batch_create(gfr_ctx* ctx, const gfr_geometry* g, const gfr_batch_params* p,
                     const float* norm, gfr_batch** out);
/* Adds a frame: prompts (and background, or NULL) are ndata floats, unscaled,
 * in host or device memory. Admission checks the context's HBM budget
 * (GFR_ENOMEM when it does not fit); a NaN or infinite count fails with
 * GFR_EINVAL. Returns the frame id in *frame. */
int gfr_batch_add(gfr_batch* b, int niter, int nsub, const float* prompts, const float* bg,
                  int* frame);
/* One subset pass of every active frame; returns the number of frames still
 * active afterwards (>= 0) or an error code. Enqueues only. */
int gfr_batch_step(gfr_batch* b);
/* Steps until every frame is complete, then waits. */
int gfr_batch_run(gfr_batch* b);
/* Frame position (as gfr_recon_cursor); returns 1 when the frame is complete. */
int gfr_batch_cursor(gfr_batch* b, int frame, int* iter, int* subset, long long* passes);
/* Copies a frame's image (nvox floats) to the host. */
int gfr_batch_get(gfr_batch* b, int frame, float* host_x);
/* FP64 log-likelihood of the frame's last pass (definition as
 * gfr_recon_loglik). */
int gfr_batch_loglik(gfr_batch* b, int frame, double* out);
/* Frees a frame's slot and memory (waits for the batch's work first). */
int gfr_batch_remove(gfr_batch* b, int frame);
/* 1: steps replay a captured graph (default); 0: the same kernels are
 * launched one by one. */
int gfr_batch_use_graph(gfr_batch* b, int enable);
/* Graph launches and graph (re)builds so far. */
int gfr_batch_stats(gfr_batch* b, long long* launches, long long* builds);
void* gfr_batch_stream(gfr_batch* b);
int gfr_batch_destroy(gfr_batch* b);

/* ---------------- serving loop ---------------- */

/* Reconstruction requests (any geometry, mode, storage and count level)
 * served from one context: each request is admitted as a device-resident
 * session (gfr_recon_create) when fewer than max_active are running and the
 * context's HBM budget has room (arrival order; a request that cannot fit on
 * an idle server fails with GFR_ENOMEM), and every admitted session advances
 * one pass at a time on its own stream, at most `inflight` passes queued
 * ahead of the device, so requests join and leave the running set between
 * passes and their kernels overlap. A request's result is bit-identical to
 * the same session run alone. Host threads inside the server drive the
 * device; the caller only submits and collects. The reference runs one study
 * per process (DRV:577-621). */
typedef struct gfr_server gfr_server;
#define GFR_REQ_PENDING 0
#define GFR_REQ_ACTIVE 1
#define GFR_REQ_DONE 2
#define GFR_REQ_FAILED 3
typedef struct gfr_request_info {
  int state;        /* GFR_REQ_*                                    */
  int error;        /* status code of a failed request              */
  double t_submit;  /* seconds on the server clock                  */
  double t_admit;   /* session created                              */
  double t_done;    /* result on the host (or failure)              */
  long long passes; /* passes enqueued so far                       */
  long long bytes;  /* device bytes of the session                  */
} gfr_request_info;
typedef struct gfr_server_totals {
  long long submitted, pending, active, done, failed, passes;
  double now; /* server clock, seconds */
} gfr_server_totals;
/* admitters: host threads creating sessions concurrently (uploads, masks). */
int gfr_server_create(gfr_ctx* ctx, int max_active, int inflight, int admitters,
                      gfr_server** out);
/* Queues a request; arguments as gfr_recon_create (d2/t2/norm in host or
 * device memory, device data complete before the call). The pointed-to
 * inputs must stay valid until the request has left GFR_REQ_PENDING. */
int gfr_server_submit(gfr_server* s, const gfr_geometry* g, const gfr_recon_params* p,
                      const float* d2, const float* t2, const float* norm, const int32_t* rec,
                      long long nrec, long long* id);
int gfr_server_info(gfr_server* s, long long id, gfr_request_info* info);
/* Blocks until the request is done or failed (timeout_s < 0: no limit);
 * returns its state. */
int gfr_server_wait(gfr_server* s, long long id, double timeout_s);
/* Copies the images of a done request (nvox floats each; xT in modes 2-3);
 * a failed request returns its error code. */
int gfr_server_result(gfr_server* s, long long id, float* xD, float* xT);
/* Forgets a done or failed request (frees its host result). */
int gfr_server_release(gfr_server* s, long long id);
/* Cancels one request: a pending one fails at once, an admitted one stops
 * before its next pass and its session is freed; either ends in
 * GFR_REQ_FAILED with error GFR_ECANCELED while the other requests go on
 * unchanged. GFR_ESTATE if the request has already completed. (The
 * reference has no request to cancel; its process runs one study.) */
int gfr_server_cancel(gfr_server* s, long long id);
int gfr_server_stats(gfr_server* s, gfr_server_totals* st);
/* Stops admitting, fails pending requests (GFR_ESTATE, waiters wake up),
 * waits for and frees running sessions. No other call on the server may
 * run concurrently with or after this one. */
int gfr_server_destroy(gfr_server* s);

/* ---------------- synthetic measurements ---------------- */

/* out[i] = Poisson(scale * mean[i] + scale2 * mean2[i]) as floats (mean2 may
 * be NULL), seeded counter-based RNG (Philox4x32-10): the draw of bin i
 * depends only on (seed, i). Device pointers; the reference reads measured
 * files only. */
int gfr_poisson(gfr_ctx* ctx, const float* d_mean, double scale, const float* d_mean2,
                double scale2, unsigned long long seed, float* d_out, long long n, void* stream);

#ifdef __cplusplus
}
#endif
#endif
