// gfr_internal.h — internal types shared by the HIP kernels and the host
// code of libgfirst.so (not part of the public C ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "gfirst.h"

namespace gfr {

// Per-segment constants, segment order 0,-1,+1,-2,+2,... (GFIRST.cu:163-168).
struct SegEntry {
  int zmin;   // minimum ring difference / first valid plane (GFIRST_GPU_Kernels.cu:54)
  int nza;    // first sinogram of the segment (nz_acum, GFIRST.cu:133-140)
  float dzf;  // forward-projector z step per sample (GFIRST_GPU_Kernels.cu:83)
  float fzb;  // backprojector z tilt (GFIRST_GPU_Kernels.cu:134, via double)
};

// Geometry as the kernels see it (by value). Pointers are device tables owned
// by the context, built once per distinct geometry.
struct DevGeom {
  int res, nzs, nrad, nang, nseg, nsinogs;
  int l0, l1;         // sample index range along a LOR
  float RADM, RESM;   // (NRAD-1)/2, (RES-1)/2
  float fxf;          // forward: detector pixel -> voxel (px/vx)
  float fxb;          // backward: voxel -> detector pixel (vx/px)
  float fzb_absmax;   // largest |fzb| over the segments
  long long ndata, nvox;
  const float* sin_t;     // [nang]
  const float* cos_t;     // [nang]
  const SegEntry* seg;    // [nseg]
  const int2* sino;       // [nsinogs]: (segment index, absolute plane)
  // [l1-l0+1][nsinogs]: z interpolation of each LOR sample, x = z0 | z1 << 16,
  // y = bits of alpha_z (GFIRST_GPU_Kernels.cu:88, 96 through the clamp filter)
  const uint2* ztab;
};

struct Taps9 { float g[9]; };

// Tile shapes of the projectors (chosen per geometry and subset count by the
// autotuner; none of them changes any summation order).
struct Tune {
  int back_zb = 8;   // slices per backprojector thread: 4, 5, 6, 8 or 16
  int zsum_sg = 8;   // consecutive sinograms per z-sum wave: 4, 6 or 8
  int back_lds = 0;  // 1: backprojector reads its rows through an LDS window
  int back_waves = 0;     // register budget for this many waves per SIMD (0: compiler's)
  int rot_z = 16;         // planes per rotation block: 4, 16 or 32
  int back_bw = 4;        // waves per direct backprojector block: 4 (16x16) or 1 (8x8)
  int back_ahead = 1;     // 1-slice shapes: views whose rows are loaded ahead (1, 2, 4)
  int multi_zb = 0;       // slices of the multi-frame backprojector: 4 or 6 (0: not measured)
  bool tuned = false;
  float ms_back = 0.f, ms_zsum = 0.f;
};
struct Taps5 { float g[5]; };

// Host-side geometry derivation (shared by the API and the kernels' launchers).
struct GeomHost {
  gfr_geometry g;
  gfr_geometry_info info;
  std::vector<int> zmin, sgn, nza;
  std::vector<float> sin_t, cos_t;
  std::vector<SegEntry> seg;
  std::vector<int2> sino;
  std::vector<uint2> ztab;
  float RADM, RESM, fxf, fxb;
};

// with_ztab: also build the per-sample z table (device tables only need it).
int derive_geometry(const gfr_geometry* g, GeomHost* out, std::string* err, bool with_ztab = false);
void psf_taps(float psf, float* g9);
void gz_taps(float* g5);

struct GeomDev {
  GeomHost host;
  DevGeom dev;
  void* block = nullptr;  // single device allocation holding the tables
  std::map<int, Tune> tune;  // per subset count
  // Sinograms that depend on the geometry alone, computed by the first
  // session that needs them and then only read: the projection of the
  // unblurred unit cylinder (PROJ_BLANK, the mask source) and, per forward
  // PSF (key: its bits), the projection of the blurred unit cylinder (the
  // joint mode's initial projection cache).
  std::mutex cache_mu;
  float* pb_cyl = nullptr;
  std::map<uint32_t, float*> ptc0;
};

}  // namespace gfr

namespace gfr {
// Device scratch shared by calls on different streams: a user waits for the
// previous user's event, so calls that share it are ordered.
struct Scratch {
  void* p = nullptr;
  size_t cap = 0;
  hipEvent_t ev = nullptr;
  bool used = false;
};
}  // namespace gfr

struct gfr_ctx {
  int device = 0;
  size_t budget = 0;
  size_t in_use = 0;
  hipStream_t stream = nullptr;
  std::string err;
  std::mutex err_mu;  // err is written by any thread using the context
  std::map<std::string, gfr::GeomDev*> geoms;
  std::map<void*, size_t> allocs;
  // device blocks of finished sessions kept for reuse, by exact size
  std::multimap<size_t, void*> kept;
  size_t kept_bytes = 0;
  size_t kept_max = 0;  // GFR_KEPT_MAX_MB (MiB), default kKeptMax
  std::vector<hipStream_t> kept_streams;  // idle non-blocking streams
  unsigned int* d_scratch_u32 = nullptr;  // 256 words for small reductions
  gfr::Scratch rot;                       // rotation stage of per-operation forward calls
  gfr::Scratch pack;                      // packed ratio of per-operation backprojections
  std::mutex mu;
};

namespace gfr {
// Session memory: sessions of one request class (geometry and parameters)
// allocate the same block sizes, so blocks of finished sessions are kept by
// exact size (up to kKeptMax bytes) and handed to the next session instead
// of a hipMalloc/hipFree pair per block (at most ctx->kept_max bytes kept:
// environment GFR_KEPT_MAX_MB, default 32 GiB, 0 keeps nothing). dev_free
// takes a block whose work is complete; dev_trim returns every kept block to
// the device (also before a retried allocation that ran out of memory).
constexpr size_t kKeptMax = (size_t)32 << 30;
hipError_t dev_alloc(gfr_ctx* ctx, void** p, size_t bytes);
void dev_free(gfr_ctx* ctx, void* p, size_t bytes);
void dev_trim(gfr_ctx* ctx);
// The same for session streams (non-blocking; at most kKeptStreams, none
// when kept_max is 0): stream_give takes an idle stream.
constexpr size_t kKeptStreams = 64;
hipError_t stream_take(gfr_ctx* ctx, hipStream_t* st);
void stream_give(gfr_ctx* ctx, hipStream_t st);
// Returns scratch of at least `bytes`, ordered after its previous user.
void* scratch_acquire(gfr_ctx* ctx, Scratch& s, size_t bytes, hipStream_t st, hipError_t* err);
void scratch_release(Scratch& s, hipStream_t st);
void scratch_free(Scratch& s);
// Angles per rotation chunk.
constexpr int kRotChunk = 16;
}  // namespace gfr

namespace gfr {

// Error helpers used by every entry point.
int set_err(gfr_ctx* ctx, int code, const std::string& msg);
int hip_err(gfr_ctx* ctx, hipError_t e, const char* where);
// True for device memory (host and NULL pointers: false).
bool is_device_ptr(const void* p);
#define GFR_HIP(ctx, call)                                   \
  do {                                                       \
    hipError_t _e = (call);                                  \
    if (_e != hipSuccess) return gfr::hip_err(ctx, _e, #call); \
  } while (0)

// Resolves (and caches) the device tables of a geometry.
GeomDev* geom_for(gfr_ctx* ctx, const gfr_geometry* g);
// The cached tables of a geometry, or NULL if the context has none yet
// (allocates nothing).
GeomDev* geom_lookup(gfr_ctx* ctx, const gfr_geometry* g);
// The tile shapes in force for (geometry, nsub), defaults if none (no
// measurement, no allocation).
Tune tune_lookup(gfr_ctx* ctx, GeomDev* gd, int nsub);
// Records the address of this device's fixed scratch array of
// gfr_back_project_noalloc; called with `device` current.
hipError_t noalloc_init(int device);
// Tile shapes for (geometry, nsub): measured once per context unless
// GFR_AUTOTUNE=0 (then the defaults); synchronous on first use.
Tune tune_for(gfr_ctx* ctx, GeomDev* gd, int nsub);
hipStream_t stream_of(gfr_ctx* ctx, void* s);

// Kernel launchers (gfr_kernels.hip). All return hipError_t of the launch.
hipError_t launch_fill_cylinder(const DevGeom& G, float val, float* img, hipStream_t st);
hipError_t launch_zero(void* p, size_t bytes, hipStream_t st);
hipError_t launch_blur_xy(const DevGeom& G, const Taps9& t, const float* in, float* out,
                          hipStream_t st);
// Forward projection outputs (any subset of them).
struct FwdOut {
  float* proj = nullptr;          // projection, full sinogram layout
  float* corr = nullptr;          // ratio, full layout
  float2* corr2 = nullptr;        // {ratio, calib} packed [angle of subset][s][r]
  float* corr4 = nullptr;         // or: channel `chan` of interleaved 3-frame rows
  int chan = 0;
  const uint8_t* mask = nullptr;  // calibration mask (needed with corr2)
  const void* y = nullptr;        // measured prompts (storage format, unscaled)
  const void* bg = nullptr;       // background or NULL
  int storage = GFR_STORE_F32;
  float yscale = 1.f;             // RES
  double* llh_part = nullptr;     // forward_llh_blocks() doubles, or NULL
};
// Scratch of the rotation stage: floats per angle; the launcher processes the
// subset's angles in chunks of r_cap_floats / rotate_floats_per_angle().
size_t rotate_floats_per_angle(const DevGeom& G);
int forward_llh_blocks(const DevGeom& G, int subset, int nsub, int sg);
// dsub: optional device pointer to the subset index (read by the kernels, so
// a captured graph stays valid while the index changes); grids then cover
// the largest subset.
hipError_t launch_forward(const DevGeom& G, const float* img, float* R, size_t r_cap_floats,
                          const FwdOut& out, int subset, int nsub, const int* dsub,
                          const Tune& tune, hipStream_t st);
hipError_t launch_back_packed(const DevGeom& G, const float2* corr2, float* numer, float* denom,
                              int subset, int nsub, const int* dsub, const Tune& tune,
                              hipStream_t st);
// Multi-frame backprojection: nf (2 or 3) frames' interleaved rows
// (packed4_bytes) -> numer[f], denom[f]; the sums of launch_back_packed on
// each frame's own rows, bit for bit. zb: 4 or 6 slices per thread.
constexpr int kMaxGroup = 3;
size_t packed4_bytes(const DevGeom& G, int subset, int nsub);
hipError_t launch_back_multi(const DevGeom& G, const float* corr4, float* const* numer,
                             float* const* denom, int nf, int subset, int nsub, const int* dsub,
                             int zb, hipStream_t st);
// Sensitivity of a subset: the backprojector's denominator half in FP64
// (packed rows carry ratio 1, so only calibration gaps apply).
hipError_t launch_sensitivity(const DevGeom& G, const float2* corr2, double* sens, int subset,
                              int nsub, hipStream_t st);
// FP64 log-likelihood of the last forward pass (sum of its block partials).
hipError_t launch_llh_sum(const DevGeom& G, const double* part, int subset, int nsub,
                          const int* dsub, int sg, double* out, hipStream_t st);
hipError_t launch_ratio(const DevGeom& G, const float* proj, const float* y, const void* bg,
                        int storage, float yscale, float* corr, float2* corr2,
                        const uint8_t* mask, int subset, int nsub, hipStream_t st);
hipError_t launch_calib_mask(const DevGeom& G, const float* pb, const float* norm,
                             uint8_t* mask, hipStream_t st);
// Full-layout ratio + mask -> packed {ratio, calib} of the subset (packed_bytes).
size_t packed_bytes(const DevGeom& G, int subset, int nsub);
hipError_t launch_pack(const DevGeom& G, const float* corr, const uint8_t* mask, float2* c2,
                       int subset, int nsub, hipStream_t st);
hipError_t launch_update(const DevGeom& G, const Taps9& t9, const Taps5& t5, const float* numer,
                         const float* denom, float* x, float* work, hipStream_t st);
hipError_t launch_jbf(const DevGeom& G, const float* in, const float* guide,
                      const float* d_max_ref, float max_ref, int r, float s3d, float sI,
                      float* out, hipStream_t st);
hipError_t launch_blend(const DevGeom& G, float coef, const float* filt, float* x,
                        hipStream_t st);
hipError_t launch_max(const DevGeom& G, const float* x, unsigned int* d_bits, hipStream_t st);
hipError_t launch_scale_copy(const float* in, float s, float* out, long long n, hipStream_t st);
hipError_t launch_convert_store(const float* in, void* out, int storage, long long n,
                                hipStream_t st);
hipError_t launch_llh_final(const double* part, int n, double* out, hipStream_t st);
// *flag = 1 if some v[i] is not <= 0 (NaN counts); *flag must start at 0.
hipError_t launch_any_positive(const float* v, long long n, unsigned int* flag, hipStream_t st);
// *flag = 1 if some v[i] is NaN or infinite; *flag must start at 0.
hipError_t launch_any_nonfinite(const float* v, long long n, unsigned int* flag,
                                hipStream_t st);
// sub[slots[i]] = (sub[slots[i]] + 1) mod nsub[slots[i]] for i < n.
hipError_t launch_advance(int* sub, const int* nsub, const int* slots, int n, hipStream_t st);
// Synthetic measurements (gfr_synth.hip).
hipError_t launch_poisson(const float* mean, double scale, const float* mean2, double scale2,
                          uint64_t seed, float* out, long long n, hipStream_t st);

// Largest number of FP64 partials a forward pass with `nsub` subsets writes
// (any tile shape).
int llh_capacity(const GeomHost& h, int nsub);
// Times each tile shape of the two projectors on the given buffers (subset 0)
// and returns the fastest; any shape gives identical results.
hipError_t tune_shapes(const DevGeom& G, int nsub, const FwdOut& fo, const float* img, float* R,
                       size_t r_floats, const float2* corr2, float* num, float* den,
                       hipStream_t st, Tune* out);

// VLOR.
hipError_t launch_vlor_count(const int* i1, const int* i2, long long nev, float* v0,
                             hipStream_t st);
hipError_t launch_vlor_blankt(const float* yT, const float* v0, float res, float* blankt,
                              long long nd, hipStream_t st);
hipError_t launch_vlor_event(const int* i1, const int* i2, const float* blankt,
                             const float* proj, float* sinv, float* valor, long long nev,
                             int want_valor, hipStream_t st);
hipError_t launch_vlor_zero_subset(const DevGeom& G, float* dv, int subset, int nsub,
                                   hipStream_t st);
// Valid events and bin -> event lists from raw records (device, `d_rec` holds
// nrec pairs of 1-based bins); sizes: i1/i2 nrec, bins/ev 2*nrec, start
// 2*nrec+1. Synchronises `st` and returns the counts.
hipError_t build_vlor_lists(const int* d_rec, long long nrec, long long nd, int* d_i1, int* d_i2,
                            long long* d_bins, int* d_start, int* d_ev, long long* nev_out,
                            long long* nbins_out, hipStream_t st);
hipError_t launch_vlor_bins(const DevGeom& G, const long long* bins, const int* start,
                            const int* ev, long long nbins, const float* proj,
                            const float* sinv, const float* valor, float* dv, int subset,
                            int nsub, hipStream_t st);

}  // namespace gfr

namespace gfr {
// gfr_vlor_create with its device work ordered on stream st.
int vlor_create_on(gfr_ctx* ctx, const gfr_geometry* g, const int32_t* rec, long long nrec,
                   const float* d_yT, hipStream_t st, gfr_vlor** out);
}  // namespace gfr

struct gfr_vlor {
  gfr_ctx* ctx = nullptr;
  gfr_geometry g;
  long long nev = 0;      // valid records
  long long nbins = 0;    // touched bins
  int* d_i1 = nullptr;    // [nev] 0-based
  int* d_i2 = nullptr;
  float* d_blankt = nullptr;  // [ndata]
  float* d_sinv = nullptr;    // [nev]
  float* d_valor = nullptr;   // [nev]
  long long* d_bins = nullptr;  // [nbins] bin ids, ascending
  int* d_start = nullptr;       // [nbins+1] offsets into d_ev
  int* d_ev = nullptr;          // [2*nev] event ids per bin, record order
  size_t bytes = 0;
};
