// gfr_recon.cpp — whole reconstruction as a device-resident session
// (osem3d, GFIRST.cu:44-575). Data are uploaded once; every array stays in
// HBM; a step is one (type, subset) pass and enqueues only kernels (no host
// round trip), so steps of different sessions overlap on their own streams.
//
// Loop nest: iter -> type (doubles, then triples) -> subset. Subset s holds
// the angles theta with theta % S == s; S = 1 is the reference's loop.
//
// Joint modes (2, 3) run two chains. The triples passes of iteration i read
// the doubles image only as the bilateral filter's guide (and its maximum),
// and the doubles passes of iteration i+1 do not read the triples image, so
// the triples passes go to a second stream with scratch of their own, guided
// by a snapshot of the doubles image taken when iteration i's doubles passes
// end; they run beside iteration i+1's doubles passes (the bilateral filter
// is bound by VALU issue, the projectors by the vector-memory path). Every
// pass sees exactly the inputs of the reference's sequential loop
// (GFIRST.cu:303-527), so the images are the same bits.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "gfr_internal.h"

using namespace gfr;

struct gfr_recon {
  gfr_ctx* ctx = nullptr;
  gfr_geometry g{};
  gfr_recon_params p{};
  GeomDev* gd = nullptr;
  hipStream_t st = nullptr;
  int ntypes = 1;
  // image-sized
  float *xD = nullptr, *xT = nullptr, *xb = nullptr, *num = nullptr, *den = nullptr,
        *work = nullptr, *jo = nullptr;
  // sinogram-sized
  float *pb = nullptr, *ptc = nullptr, *dv = nullptr;
  float2* corr2 = nullptr;  // {ratio, calib} of the current subset, [angle][s][r]
  float* R = nullptr;       // rotation stage of the forward projector
  size_t r_floats = 0;
  void *yD = nullptr, *bD = nullptr, *yT = nullptr, *bT = nullptr;
  uint8_t* mask = nullptr;
  gfr_vlor* vl = nullptr;
  double* llh_part = nullptr;
  double* d_llh = nullptr;
  unsigned int* d_maxbits = nullptr;
  int llh_valid = 0;
  // triples chain (modes 2-3): stream, scratch, the doubles snapshot
  hipStream_t stT = nullptr;
  float *xbT = nullptr, *numT = nullptr, *denT = nullptr, *workT = nullptr, *xDs = nullptr;
  float2* corr2T = nullptr;
  float* RT = nullptr;
  double *llh_partT = nullptr, *d_llhT = nullptr;
  hipEvent_t ev_snap = nullptr;   // snapshot of the doubles image taken
  hipEvent_t ev_tfree = nullptr;  // an iteration's triples passes are done with it
  hipEvent_t ev_join = nullptr;
  bool tfree_pending = false;
  int llh_chain = 0;  // chain of the last pass's log-likelihood (0 doubles, 1 triples)
  std::vector<std::pair<void*, size_t>> owned;  // blocks from dev_alloc
  size_t bytes = 0;
  Tune tune;
  int iter = 1, type = 0, subset = 0;
  long long passes = 0;
  bool done = false;
  // tracing: event pairs per phase, resolved lazily by gfr_recon_phase_times
  bool prof = false;
  std::vector<hipEvent_t> ev_free;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_pending;
  double phase_ms[GFR_NPHASES] = {0};
  long long phase_n[GFR_NPHASES] = {0};
};

namespace {

size_t store_bytes(int storage) { return storage == GFR_STORE_F32 ? 4 : 2; }

struct Plan {
  size_t img, sino_f32, sino_store, mask, llh, corr2, rot;
  int nimg, nsino_f32, nsino_store;
  bool two;  // second (triples) chain
};

bool any_positive(const float* v, long long n) {
  for (long long i = 0; i < n; i++)
    if (!(v[i] <= 0.f)) return true;  // NaN counts as present
  return false;
}

Plan plan_of(const GeomHost& h, const gfr_recon_params* p, bool bgD, bool bgT) {
  Plan pl{};
  pl.img = sizeof(float) * (size_t)h.info.nvox;
  pl.sino_f32 = sizeof(float) * (size_t)h.info.ndata;
  pl.sino_store = store_bytes(p->storage) * (size_t)h.info.ndata;
  pl.mask = (size_t)h.info.ndata;
  pl.llh = sizeof(double) * (size_t)llh_capacity(h, p->nsub) + 64;
  const size_t nav = (size_t)(h.g.nang + p->nsub - 1) / p->nsub;
  pl.corr2 = sizeof(float2) * nav * ((size_t)h.info.nsinogs + 2 * (size_t)h.g.nseg) *
             ((size_t)h.g.nrad + 2);  // packed rows: NRAD + 2 entries
  pl.rot = sizeof(float) * (size_t)(h.info.lambda1 - h.info.lambda0 + 1) * h.g.nzs * h.g.nrad *
           std::min(nav, (size_t)kRotChunk);
  const bool two = p->mode >= 2;
  pl.two = two;
  // xD num den work xb (+ xT jo, and the triples chain's xb num den work and
  // the doubles snapshot)
  pl.nimg = 6 + (two ? 7 : 0);
  pl.nsino_f32 = p->mode == 3 ? 2 : 0;              // ptc dv (PROJ_BLANK is per geometry)
  pl.nsino_store = 1 + (bgD ? 1 : 0) + (two ? 1 + (bgT ? 1 : 0) : 0);
  return pl;
}

size_t plan_bytes(const Plan& pl) {
  return pl.nimg * pl.img + pl.nsino_f32 * pl.sino_f32 + pl.nsino_store * pl.sino_store +
         pl.mask + (pl.two ? 2 : 1) * (pl.llh + pl.corr2 + pl.rot);
}

hipEvent_t take_event(gfr_recon* r) {
  if (!r->ev_free.empty()) {
    hipEvent_t e = r->ev_free.back();
    r->ev_free.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Brackets one phase of a pass (on the pass's stream) with events when
// tracing is on.
struct PhaseMark {
  gfr_recon* r;
  int phase;
  hipStream_t st;
  hipEvent_t e0 = nullptr;
  PhaseMark(gfr_recon* rr, int ph, hipStream_t s) : r(rr), phase(ph), st(s) {
    if (r->prof) {
      e0 = take_event(r);
      if (e0) (void)hipEventRecord(e0, st);
    }
  }
  void end() {
    if (!e0) return;
    hipEvent_t e1 = take_event(r);
    if (e1 && hipEventRecord(e1, st) == hipSuccess) {
      r->ev_pending.push_back({phase, {e0, e1}});
    } else {
      r->ev_free.push_back(e0);
      if (e1) r->ev_free.push_back(e1);
    }
    e0 = nullptr;
  }
  ~PhaseMark() { end(); }
};

// Tile shapes for this session's geometry and subset count: measured on its
// own buffers the first time (then shared with later sessions and batches of
// the same geometry and subset count), defaults when GFR_AUTOTUNE=0.
hipError_t autotune(gfr_recon* r) {
  GeomDev* gd = r->gd;
  const int S = r->p.nsub;
  {
    std::lock_guard<std::mutex> lk(r->ctx->mu);
    const Tune& c = gd->tune[S];
    if (c.tuned) {
      r->tune = c;
      return hipSuccess;
    }
  }
  const char* env = getenv("GFR_AUTOTUNE");
  if (env && env[0] == '0') {
    r->tune = Tune();
    return hipSuccess;
  }
  FwdOut fo;
  fo.corr2 = r->corr2;
  fo.mask = r->mask;
  fo.y = r->yD;
  fo.bg = r->bD;
  fo.storage = r->p.storage;
  fo.yscale = (float)r->g.res;
  fo.llh_part = r->llh_part;
  Tune t;
  hipError_t e = tune_shapes(gd->dev, S, fo, r->xb, r->R, r->r_floats, r->corr2, r->num, r->den,
                             r->st, &t);
  if (e != hipSuccess) return e;
  r->tune = t;
  std::lock_guard<std::mutex> lk(r->ctx->mu);
  gd->tune[S] = t;
  return hipSuccess;
}

// Whether a background half holds any value that is not <= 0 (NaN counts),
// for host or device memory.
int bg_present(gfr_ctx* ctx, const float* v, long long n, bool* out) {
  if (!is_device_ptr(v)) {
    *out = any_positive(v, n);
    return GFR_OK;
  }
  unsigned int* flag = nullptr;
  unsigned int h = 0;
  hipError_t e = hipMalloc((void**)&flag, sizeof(unsigned int));
  if (e == hipSuccess) e = hipMemcpy(flag, &h, sizeof(h), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = launch_any_positive(v, n, flag, nullptr);
  if (e == hipSuccess) e = hipMemcpy(&h, flag, sizeof(h), hipMemcpyDeviceToHost);
  if (flag) (void)hipFree(flag);
  if (e != hipSuccess) return hip_err(ctx, e, "gfr_recon_create background check");
  *out = h != 0;
  return GFR_OK;
}

int rfail(gfr_recon* r, int code) {
  gfr_recon_destroy(r);
  return code;
}

}  // namespace

extern "C" {

// One of the geometry's shared sinograms (PROJ_BLANK, or with `blurred` the
// projection of the unit cylinder blurred by `psf`), computed on the
// session's stream with its scratch buffers by the first session that needs
// it; later sessions of the geometry only read it.
hipError_t geom_sino(gfr_recon* r, bool blurred, float psf, const float** out) {
  GeomDev* gd = r->gd;
  gfr_ctx* ctx = r->ctx;
  std::lock_guard<std::mutex> lk(gd->cache_mu);
  uint32_t key = 0;
  memcpy(&key, &psf, sizeof(key));
  float* p = blurred ? (gd->ptc0.count(key) ? gd->ptc0[key] : nullptr) : gd->pb_cyl;
  if (!p) {
    const size_t bytes = sizeof(float) * (size_t)gd->host.info.ndata;
    {
      std::lock_guard<std::mutex> lc(ctx->mu);
      if (ctx->budget && ctx->in_use + bytes > ctx->budget) return hipErrorOutOfMemory;
    }
    hipError_t e = hipMalloc((void**)&p, bytes);
    if (e != hipSuccess) return e;
    const DevGeom& G = gd->dev;
    const float* src = r->work;
    e = launch_fill_cylinder(G, 1.f, r->work, r->st);
    if (e == hipSuccess && blurred) {
      Taps9 t9;
      psf_taps(psf, t9.g);
      e = launch_blur_xy(G, t9, r->work, r->xb, r->st);
      src = r->xb;
    }
    FwdOut fo;
    fo.proj = p;
    if (e == hipSuccess) e = launch_forward(G, src, r->R, r->r_floats, fo, 0, 1, nullptr, r->tune, r->st);
    if (e == hipSuccess) e = hipStreamSynchronize(r->st);
    if (e != hipSuccess) {
      (void)hipFree(p);
      return e;
    }
    if (blurred)
      gd->ptc0[key] = p;
    else
      gd->pb_cyl = p;
    std::lock_guard<std::mutex> lc(ctx->mu);
    ctx->in_use += bytes;
  }
  *out = p;
  return hipSuccess;
}

long long gfr_recon_footprint(const gfr_geometry* g, const gfr_recon_params* p) {
  GeomHost h;
  if (!p || derive_geometry(g, &h, nullptr) != GFR_OK) return GFR_EINVAL;
  return (long long)plan_bytes(plan_of(h, p, true, true));
}

int gfr_recon_create(gfr_ctx* ctx, const gfr_geometry* g, const gfr_recon_params* p,
                     const float* d2, const float* t2, const float* norm, const int32_t* rec,
                     long long nrec, gfr_recon** out) {
  if (!ctx || !out) return GFR_EINVAL;
  *out = nullptr;
  if (!p) return set_err(ctx, GFR_EINVAL, "params is NULL");
  GeomDev* gd = geom_for(ctx, g);
  if (!gd) return GFR_EINVAL;
  const GeomHost& h = gd->host;
  if (p->mode < 1 || p->mode > 3) return set_err(ctx, GFR_EINVAL, "mode must be 1, 2 or 3");
  if (p->niter < 0 || p->nsub < 1 || p->nsub > g->nang)
    return set_err(ctx, GFR_EINVAL, "niter >= 0 and 1 <= nsub <= nang required");
  if (p->storage < GFR_STORE_F32 || p->storage > GFR_STORE_BF16)
    return set_err(ctx, GFR_EINVAL, "unknown storage");
  if (g->nzs < 3) return set_err(ctx, GFR_EINVAL, "nzs must be >= 3");
  if (p->jbf_r < 0 || p->jbf_r > 8) return set_err(ctx, GFR_EINVAL, "jbf_r must be in [0, 8]");
  if (!d2 || !norm) return set_err(ctx, GFR_EINVAL, "doubles sinogram and norm are required");
  if (p->mode >= 2 && !t2) return set_err(ctx, GFR_EINVAL, "triples sinogram required");
  if (p->mode == 3 && nrec > 0 && !rec) return set_err(ctx, GFR_EINVAL, "VLOR records required");
  const long long nd = h.info.ndata;
  // Background halves that are all <= 0 are not stored (bg = 0 in the ratio).
  bool hasBgD = false, hasBgT = false;
  {
    int rc = bg_present(ctx, d2 + nd, nd, &hasBgD);
    if (rc == GFR_OK && p->mode >= 2) rc = bg_present(ctx, t2 + nd, nd, &hasBgT);
    if (rc != GFR_OK) return rc;
  }
  const Plan pl = plan_of(h, p, hasBgD, hasBgT);
  const size_t need = plan_bytes(pl);
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (ctx->budget && ctx->in_use + need > ctx->budget)
      return set_err(ctx, GFR_ENOMEM, "reconstruction exceeds the context's HBM budget");
    ctx->in_use += need;
  }
  gfr_recon* r = new gfr_recon();
  r->ctx = ctx;
  r->g = *g;
  r->p = *p;
  r->gd = gd;
  r->bytes = need;
  r->ntypes = p->mode >= 2 ? 2 : 1;
  if (stream_take(ctx, &r->st) != hipSuccess)
    return rfail(r, set_err(ctx, GFR_EHIP, "stream create"));
  auto A = [&](void** ptr, size_t b) -> bool {
    const size_t nb = b ? b : 4;
    if (dev_alloc(ctx, ptr, nb) != hipSuccess) return false;
    r->owned.push_back({*ptr, nb});
    return true;
  };
  bool ok = A((void**)&r->xD, pl.img) && A((void**)&r->num, pl.img) && A((void**)&r->den, pl.img) &&
            A((void**)&r->work, pl.img) && A((void**)&r->xb, pl.img) &&
            A((void**)&r->corr2, pl.corr2) && A((void**)&r->R, pl.rot) &&
            A((void**)&r->mask, pl.mask) && A((void**)&r->yD, pl.sino_store) &&
            A((void**)&r->llh_part, pl.llh) && A((void**)&r->d_maxbits, 64) &&
            A((void**)&r->d_llh, 64);
  if (ok && hasBgD) ok = A(&r->bD, pl.sino_store);
  if (ok && r->ntypes == 2)
    ok = A((void**)&r->xT, pl.img) && A((void**)&r->jo, pl.img) && A(&r->yT, pl.sino_store) &&
         (!hasBgT || A(&r->bT, pl.sino_store)) && A((void**)&r->xbT, pl.img) &&
         A((void**)&r->numT, pl.img) && A((void**)&r->denT, pl.img) &&
         A((void**)&r->workT, pl.img) && A((void**)&r->xDs, pl.img) &&
         A((void**)&r->corr2T, pl.corr2) && A((void**)&r->RT, pl.rot) && A((void**)&r->llh_partT, pl.llh) &&
         A((void**)&r->d_llhT, 64);
  if (ok && r->ntypes == 2)
    ok = stream_take(ctx, &r->stT) == hipSuccess &&
         hipEventCreateWithFlags(&r->ev_snap, hipEventDisableTiming) == hipSuccess &&
         hipEventCreateWithFlags(&r->ev_tfree, hipEventDisableTiming) == hipSuccess &&
         hipEventCreateWithFlags(&r->ev_join, hipEventDisableTiming) == hipSuccess;
  if (ok && p->mode == 3) ok = A((void**)&r->ptc, pl.sino_f32) && A((void**)&r->dv, pl.sino_f32);
  if (!ok) return rfail(r, set_err(ctx, GFR_ENOMEM, "reconstruction allocation failed"));
  hipStream_t st = r->st;
  const DevGeom& G = gd->dev;
  r->r_floats = pl.rot / sizeof(float);
  hipError_t e = hipSuccess;
  // Host arrays reach the device through a transient staging buffer.
  float* stage = nullptr;
  if (dev_alloc(ctx, (void**)&stage, pl.sino_f32) != hipSuccess)
    return rfail(r, set_err(ctx, GFR_ENOMEM, "reconstruction staging allocation failed"));
  // Measured sinograms with a NaN or infinite value fail the request: the
  // reference would carry it through the ratio into every voxel the LOR
  // touches (GFIRST.cu:423-426). The normalization needs no check: it only
  // selects the calibration mask through its > 0.01 test, where NaN masks
  // the bin and +inf keeps it, as there.
  unsigned int* bad = r->d_maxbits;
  auto upload = [&](const float* host, void* dst) -> hipError_t {
    hipError_t ee = hipMemcpyAsync(stage, host, pl.sino_f32, hipMemcpyDefault, st);
    if (ee == hipSuccess) ee = launch_any_nonfinite(stage, nd, bad, st);
    if (ee == hipSuccess) ee = launch_convert_store(stage, dst, p->storage, nd, st);
    if (ee == hipSuccess) ee = hipStreamSynchronize(st);
    return ee;
  };
  // Mask source: projection of the unblurred uniform cylinder (GFIRST.cu:155-181,
  // shared by the geometry's sessions) and the calibration mask
  // (GFIRST.cu:453-458) from the normalization.
  {
    const float* pb = nullptr;
    e = geom_sino(r, false, 0.f, &pb);
    r->pb = const_cast<float*>(pb);
  }
  if (e == hipSuccess) e = hipMemcpyAsync(stage, norm, pl.sino_f32, hipMemcpyDefault, st);
  if (e == hipSuccess) e = launch_calib_mask(G, r->pb, stage, r->mask, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  // Measured data, stored unscaled; kernels scale by RES on load (GFIRST.cu:224-260).
  if (e == hipSuccess) e = launch_zero(bad, sizeof(unsigned int), st);
  if (e == hipSuccess) e = upload(d2, r->yD);
  if (e == hipSuccess && hasBgD) e = upload(d2 + nd, r->bD);
  if (e == hipSuccess && r->ntypes == 2) e = upload(t2, r->yT);
  if (e == hipSuccess && hasBgT) e = upload(t2 + nd, r->bT);
  unsigned int hbad = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(&hbad, bad, sizeof(hbad), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e == hipSuccess && hbad) {
    dev_free(ctx, stage, pl.sino_f32);  // stream synchronised above
    return rfail(r, set_err(ctx, GFR_EINVAL, "non-finite value in the measured sinograms"));
  }
  // Initial estimates (GFIRST.cu:282-288).
  if (e == hipSuccess) e = launch_fill_cylinder(G, 1.f, r->xD, st);
  if (e == hipSuccess && r->ntypes == 2)
    e = launch_fill_cylinder(G, p->mode == 1 ? 0.f : 1.f, r->xT, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipStreamSynchronize(st);  // stage's work done before reuse
  dev_free(ctx, stage, pl.sino_f32);
  if (e != hipSuccess) return rfail(r, hip_err(ctx, e, "gfr_recon_create"));
  if (p->mode == 3) {
    // BLANKT needs T x RES (GFIRST.cu:225-226, 246): scale a float copy.
    e = hipMemcpyAsync(r->ptc, t2, pl.sino_f32, hipMemcpyDefault, st);
    if (e == hipSuccess) e = launch_scale_copy(r->ptc, (float)g->res, r->dv, nd, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return rfail(r, hip_err(ctx, e, "gfr_recon_create vlor"));
    int rc = vlor_create_on(ctx, g, rec, nrec, r->dv, st, &r->vl);
    if (rc != GFR_OK) return rfail(r, rc);
    // Projection cache of the triples estimate (bins outside the current
    // subset keep their last projection): it starts as the projection of the
    // blurred initial estimate, the unit cylinder (shared per geometry and PSF).
    e = launch_zero(r->dv, pl.sino_f32, st);
    const float* c0 = nullptr;
    if (e == hipSuccess) e = geom_sino(r, true, p->psf_fw, &c0);
    if (e == hipSuccess) e = hipMemcpyAsync(r->ptc, c0, pl.sino_f32, hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess) return rfail(r, hip_err(ctx, e, "gfr_recon_create vlor cache"));
  }
  e = hipStreamSynchronize(st);
  if (e == hipSuccess) e = autotune(r);
  if (e != hipSuccess) return rfail(r, hip_err(ctx, e, "gfr_recon_create sync"));
  r->done = (p->niter == 0);
  *out = r;
  return GFR_OK;
}

int gfr_recon_step(gfr_recon* r) {
  if (!r) return GFR_EINVAL;
  if (r->done) return 1;
  gfr_ctx* ctx = r->ctx;
  const DevGeom& G = r->gd->dev;
  const gfr_recon_params& p = r->p;
  const int ty = r->type, s = r->subset, S = p.nsub;
  // the pass's chain: doubles on the session's stream, triples on the second
  const bool tc = ty == 1;
  hipStream_t st = tc ? r->stT : r->st;
  float* x = tc ? r->xT : r->xD;
  float* xb = tc ? r->xbT : r->xb;
  float* R = tc ? r->RT : r->R;
  float2* corr2 = tc ? r->corr2T : r->corr2;
  float* num = tc ? r->numT : r->num;
  float* den = tc ? r->denT : r->den;
  float* work = tc ? r->workT : r->work;
  double* llh_part = tc ? r->llh_partT : r->llh_part;
  double* d_llh = tc ? r->d_llhT : r->d_llh;
  if (tc && s == 0) {
    // Fork: the doubles image at the end of this iteration's doubles passes
    // guides this iteration's triples passes; the previous iteration's
    // triples passes must be done with the snapshot first.
    if (r->tfree_pending) GFR_HIP(ctx, hipStreamWaitEvent(r->st, r->ev_tfree, 0));
    GFR_HIP(ctx, hipMemcpyAsync(r->xDs, r->xD, sizeof(float) * (size_t)r->gd->host.info.nvox,
                                hipMemcpyDeviceToDevice, r->st));
    GFR_HIP(ctx, hipEventRecord(r->ev_snap, r->st));
    GFR_HIP(ctx, hipStreamWaitEvent(r->stT, r->ev_snap, 0));
  }
  Taps9 tf, tb;
  Taps5 tz;
  psf_taps(p.psf_fw, tf.g);
  psf_taps(p.psf_bw, tb.g);
  gz_taps(tz.g);
  // PSF blur before projection (GFIRST.cu:324-340).
  {
    PhaseMark pm(r, GFR_PHASE_BLUR, st);
    GFR_HIP(ctx, launch_blur_xy(G, tf, x, xb, st));
  }
  if (tc && p.mode == 3) {
    // Forward projection into the cache, VLOR reassignment of the subset's bins
    // (GFIRST.cu:373-397), ratio against the pseudo-sinogram x RES.
    {
      PhaseMark pm(r, GFR_PHASE_FORWARD, st);
      FwdOut fc;
      fc.proj = r->ptc;
      GFR_HIP(ctx, launch_forward(G, xb, R, r->r_floats, fc, s, S, nullptr, r->tune, st));
    }
    PhaseMark pm(r, GFR_PHASE_VLOR, st);
    int rc = gfr_vlor_assign(ctx, r->vl, r->ptc, r->dv, s, S, st);
    if (rc != GFR_OK) return rc;
    GFR_HIP(ctx, launch_ratio(G, r->ptc, r->dv, r->bT, p.storage, (float)r->g.res, nullptr,
                              corr2, r->mask, s, S, st));
    r->llh_valid = 0;
  } else {
    const void* y = tc ? r->yT : r->yD;
    const void* bg = tc ? r->bT : r->bD;
    {
      PhaseMark pm(r, GFR_PHASE_FORWARD, st);
      FwdOut fo;
      fo.corr2 = corr2;
      fo.mask = r->mask;
      fo.y = y;
      fo.bg = bg;
      fo.storage = p.storage;
      fo.yscale = (float)r->g.res;
      fo.llh_part = llh_part;
      GFR_HIP(ctx, launch_forward(G, xb, R, r->r_floats, fo, s, S, nullptr, r->tune, st));
    }
    GFR_HIP(ctx, launch_llh_sum(G, llh_part, s, S, nullptr, r->tune.zsum_sg, d_llh, st));
    r->llh_valid = 1;
    r->llh_chain = tc ? 1 : 0;
  }
  // Backprojection of the subset (GFIRST.cu:436-469) and update (:482-502).
  {
    PhaseMark pm(r, GFR_PHASE_BACK, st);
    GFR_HIP(ctx, launch_back_packed(G, corr2, num, den, s, S, nullptr, r->tune, st));
  }
  {
    PhaseMark pm(r, GFR_PHASE_UPDATE, st);
    GFR_HIP(ctx, launch_update(G, tb, tz, num, den, x, work, st));
  }
  if (tc) {
    // Joint bilateral filter of triples guided by doubles + blend (:504-526).
    PhaseMark pm(r, GFR_PHASE_JBF, st);
    GFR_HIP(ctx, launch_max(G, r->xDs, r->d_maxbits, st));
    GFR_HIP(ctx, launch_jbf(G, r->xT, r->xDs, (const float*)r->d_maxbits, 0.f, p.jbf_r,
                            p.jbf_sigma3d, p.jbf_sigmai, r->jo, st));
    GFR_HIP(ctx, launch_blend(G, p.jbf_coef, r->jo, r->xT, st));
    if (s == S - 1) {
      GFR_HIP(ctx, hipEventRecord(r->ev_tfree, st));
      r->tfree_pending = true;
    }
  }
  r->passes++;
  if (++r->subset == S) {
    r->subset = 0;
    if (++r->type == r->ntypes) {
      r->type = 0;
      if (++r->iter > p.niter) r->done = true;
    }
  }
  if (r->done && r->stT) {
    // join: the session's stream completes after the triples chain
    GFR_HIP(ctx, hipEventRecord(r->ev_join, r->stT));
    GFR_HIP(ctx, hipStreamWaitEvent(r->st, r->ev_join, 0));
  }
  return r->done ? 1 : 0;
}

namespace {
// Waits for both chains of a session.
hipError_t sync_all(gfr_recon* r) {
  hipError_t e = r->stT ? hipStreamSynchronize(r->stT) : hipSuccess;
  if (e == hipSuccess) e = hipStreamSynchronize(r->st);
  return e;
}
}  // namespace

int gfr_recon_run(gfr_recon* r) {
  if (!r) return GFR_EINVAL;
  int rc = 0;
  while ((rc = gfr_recon_step(r)) == 0) {
  }
  if (rc < 0) return rc;
  GFR_HIP(r->ctx, sync_all(r));
  return GFR_OK;
}

int gfr_recon_cursor(gfr_recon* r, int* iter, int* type, int* subset, long long* passes) {
  if (!r) return GFR_EINVAL;
  if (iter) *iter = r->iter;
  if (type) *type = r->type;
  if (subset) *subset = r->subset;
  if (passes) *passes = r->passes;
  return r->done ? 1 : 0;
}

int gfr_recon_get(gfr_recon* r, int what, float* host_out) {
  if (!r || !host_out) return GFR_EINVAL;
  const GeomHost& h = r->gd->host;
  const void* src = nullptr;
  size_t bytes = 0;
  switch (what) {
    case GFR_OUT_XD: src = r->xD; bytes = sizeof(float) * h.info.nvox; break;
    case GFR_OUT_XT: src = r->xT; bytes = sizeof(float) * h.info.nvox; break;
    case GFR_OUT_PROJ_BLANK: src = r->pb; bytes = sizeof(float) * h.info.ndata; break;
    case GFR_OUT_VLOR: src = r->dv; bytes = sizeof(float) * h.info.ndata; break;
    default: return set_err(r->ctx, GFR_EINVAL, "gfr_recon_get: unknown output");
  }
  if (!src) return set_err(r->ctx, GFR_ESTATE, "gfr_recon_get: output not held in this mode");
  GFR_HIP(r->ctx, sync_all(r));
  GFR_HIP(r->ctx, hipMemcpyAsync(host_out, src, bytes, hipMemcpyDeviceToHost, r->st));
  GFR_HIP(r->ctx, hipStreamSynchronize(r->st));
  return GFR_OK;
}

int gfr_recon_loglik(gfr_recon* r, double* out) {
  if (!r || !out) return GFR_EINVAL;
  if (!r->llh_valid) return set_err(r->ctx, GFR_ESTATE, "no log-likelihood for the last pass");
  GFR_HIP(r->ctx, sync_all(r));
  GFR_HIP(r->ctx, hipMemcpyAsync(out, r->llh_chain ? r->d_llhT : r->d_llh, sizeof(double),
                                 hipMemcpyDeviceToHost, r->st));
  GFR_HIP(r->ctx, hipStreamSynchronize(r->st));
  return GFR_OK;
}

void* gfr_recon_stream(gfr_recon* r) { return r ? (void*)r->st : nullptr; }

int gfr_recon_sync(gfr_recon* r) {
  if (!r) return GFR_EINVAL;
  GFR_HIP(r->ctx, sync_all(r));
  return GFR_OK;
}

int gfr_recon_profile(gfr_recon* r, int enable) {
  if (!r) return GFR_EINVAL;
  r->prof = enable != 0;
  return GFR_OK;
}

int gfr_recon_phase_times(gfr_recon* r, double* ms, long long* counts) {
  if (!r) return GFR_EINVAL;
  GFR_HIP(r->ctx, sync_all(r));
  for (auto& pe : r->ev_pending) {
    float t = 0.f;
    if (hipEventElapsedTime(&t, pe.second.first, pe.second.second) == hipSuccess) {
      r->phase_ms[pe.first] += t;
      r->phase_n[pe.first] += 1;
    }
    r->ev_free.push_back(pe.second.first);
    r->ev_free.push_back(pe.second.second);
  }
  r->ev_pending.clear();
  for (int i = 0; i < GFR_NPHASES; i++) {
    if (ms) ms[i] = r->phase_ms[i];
    if (counts) counts[i] = r->phase_n[i];
    r->phase_ms[i] = 0.0;
    r->phase_n[i] = 0;
  }
  return GFR_OK;
}

int gfr_osem_run(gfr_ctx* ctx, const gfr_geometry* g, const gfr_recon_params* p, const float* d2,
                 const float* t2, const float* norm, const int32_t* rec, long long nrec,
                 float* xD, float* xT) {
  if (!ctx) return GFR_EINVAL;
  if (!xD || (p && p->mode >= 2 && !xT))
    return set_err(ctx, GFR_EINVAL, "gfr_osem_run: output images required");
  gfr_recon* r = nullptr;
  int rc = gfr_recon_create(ctx, g, p, d2, t2, norm, rec, nrec, &r);
  if (rc != GFR_OK) return rc;
  rc = gfr_recon_run(r);
  if (rc == GFR_OK) rc = gfr_recon_get(r, GFR_OUT_XD, xD);
  if (rc == GFR_OK && p->mode >= 2) rc = gfr_recon_get(r, GFR_OUT_XT, xT);
  gfr_recon_destroy(r);
  return rc;
}

int gfr_recon_destroy(gfr_recon* r) {
  if (!r) return GFR_EINVAL;
  if (r->stT) (void)hipStreamSynchronize(r->stT);
  if (r->st) (void)hipStreamSynchronize(r->st);
  if (r->vl) gfr_vlor_destroy(r->vl);
  for (auto& pe : r->ev_pending) {
    (void)hipEventDestroy(pe.second.first);
    (void)hipEventDestroy(pe.second.second);
  }
  for (hipEvent_t e : r->ev_free) (void)hipEventDestroy(e);
  for (auto& q : r->owned) dev_free(r->ctx, q.first, q.second);  // streams synchronised above
  for (hipEvent_t e : {r->ev_snap, r->ev_tfree, r->ev_join})
    if (e) (void)hipEventDestroy(e);
  stream_give(r->ctx, r->stT);  // both synchronised above
  stream_give(r->ctx, r->st);
  {
    std::lock_guard<std::mutex> lk(r->ctx->mu);
    r->ctx->in_use -= r->bytes;
  }
  delete r;
  return GFR_OK;
}

}  // extern "C"
