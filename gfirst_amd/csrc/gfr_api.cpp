// gfr_api.cpp — context, memory and per-operation entry points of the C ABI.
// Each entry validates its arguments, resolves the geometry's device tables
// (cached per context) and enqueues kernels on the given stream; nothing here
// synchronises except the explicitly synchronous calls.
#include <cstdlib>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "gfr_internal.h"

namespace gfr {

int set_err(gfr_ctx* ctx, int code, const std::string& msg) {
  if (ctx) {
    std::lock_guard<std::mutex> lk(ctx->err_mu);
    ctx->err = msg;
  }
  return code;
}

bool is_device_ptr(const void* p) {
  if (!p) return false;
  hipPointerAttribute_t at{};
  const bool dev = hipPointerGetAttributes(&at, p) == hipSuccess && at.type == hipMemoryTypeDevice;
  (void)hipGetLastError();  // a pageable host pointer reports an error: not sticky
  return dev;
}

int hip_err(gfr_ctx* ctx, hipError_t e, const char* where) {
  char buf[512];
  snprintf(buf, sizeof(buf), "%s: %s", where, hipGetErrorString(e));
  return set_err(ctx, e == hipErrorOutOfMemory ? GFR_ENOMEM : GFR_EHIP, buf);
}

hipStream_t stream_of(gfr_ctx* ctx, void* s) { return s ? (hipStream_t)s : ctx->stream; }

Tune tune_for(gfr_ctx* ctx, GeomDev* gd, int nsub) {
  std::lock_guard<std::mutex> lk(ctx->mu);
  return gd->tune[nsub];
}

void* scratch_acquire(gfr_ctx* ctx, Scratch& s, size_t bytes, hipStream_t st, hipError_t* err) {
  (void)ctx;
  *err = hipSuccess;
  if (!s.ev) {
    *err = hipEventCreateWithFlags(&s.ev, hipEventDisableTiming);
    if (*err != hipSuccess) return nullptr;
  }
  if (s.cap < bytes) {
    if (s.p) {
      (void)hipEventSynchronize(s.ev);
      (void)hipFree(s.p);
      s.p = nullptr;
      s.cap = 0;
      s.used = false;
    }
    *err = hipMalloc(&s.p, bytes);
    if (*err != hipSuccess) {
      s.p = nullptr;
      return nullptr;
    }
    s.cap = bytes;
  }
  if (s.used) {
    *err = hipStreamWaitEvent(st, s.ev, 0);
    if (*err != hipSuccess) return nullptr;
  }
  return s.p;
}

void scratch_release(Scratch& s, hipStream_t st) {
  if (hipEventRecord(s.ev, st) == hipSuccess) s.used = true;
}

void scratch_free(Scratch& s) {
  if (s.ev) {
    (void)hipEventSynchronize(s.ev);
    (void)hipEventDestroy(s.ev);
  }
  if (s.p) (void)hipFree(s.p);
  s = Scratch();
}

GeomDev* geom_lookup(gfr_ctx* ctx, const gfr_geometry* g) {
  if (!g) return nullptr;
  std::string key((const char*)g, sizeof(gfr_geometry));
  std::lock_guard<std::mutex> lk(ctx->mu);
  auto it = ctx->geoms.find(key);
  return it == ctx->geoms.end() ? nullptr : it->second;
}

Tune tune_lookup(gfr_ctx* ctx, GeomDev* gd, int nsub) {
  std::lock_guard<std::mutex> lk(ctx->mu);
  auto it = gd->tune.find(nsub);
  return it == gd->tune.end() ? Tune() : it->second;
}

GeomDev* geom_for(gfr_ctx* ctx, const gfr_geometry* g) {
  if (!g) {
    set_err(ctx, GFR_EINVAL, "geometry is NULL");
    return nullptr;
  }
  std::string key((const char*)g, sizeof(gfr_geometry));
  std::lock_guard<std::mutex> lk(ctx->mu);
  auto it = ctx->geoms.find(key);
  if (it != ctx->geoms.end()) return it->second;
  GeomDev* gd = new GeomDev();
  std::string err;
  if (derive_geometry(g, &gd->host, &err, true) != GFR_OK) {
    set_err(ctx, GFR_EINVAL, err);
    delete gd;
    return nullptr;
  }
  const GeomHost& h = gd->host;
  const size_t n_ang = (size_t)g->nang, n_seg = (size_t)g->nseg, n_sino = h.sino.size();
  const size_t off_cos = n_ang * sizeof(float);
  const size_t off_seg = ((2 * n_ang * sizeof(float)) + 15) & ~size_t(15);
  const size_t off_sino = (off_seg + n_seg * sizeof(SegEntry) + 15) & ~size_t(15);
  const size_t off_zt = (off_sino + n_sino * sizeof(int2) + 15) & ~size_t(15);
  const size_t total = off_zt + h.ztab.size() * sizeof(uint2);
  std::vector<char> blob(total, 0);
  memcpy(blob.data(), h.sin_t.data(), n_ang * sizeof(float));
  memcpy(blob.data() + off_cos, h.cos_t.data(), n_ang * sizeof(float));
  memcpy(blob.data() + off_seg, h.seg.data(), n_seg * sizeof(SegEntry));
  memcpy(blob.data() + off_sino, h.sino.data(), n_sino * sizeof(int2));
  memcpy(blob.data() + off_zt, h.ztab.data(), h.ztab.size() * sizeof(uint2));
  hipError_t e = hipMalloc(&gd->block, total);
  if (e == hipSuccess) e = hipMemcpy(gd->block, blob.data(), total, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    hip_err(ctx, e, "geometry tables");
    if (gd->block) (void)hipFree(gd->block);
    delete gd;
    return nullptr;
  }
  DevGeom& d = gd->dev;
  d.res = g->res;
  d.nzs = g->nzs;
  d.nrad = g->nrad;
  d.nang = g->nang;
  d.nseg = g->nseg;
  d.nsinogs = h.info.nsinogs;
  d.l0 = h.info.lambda0;
  d.l1 = h.info.lambda1;
  d.RADM = h.RADM;
  d.RESM = h.RESM;
  d.fxf = h.fxf;
  d.fxb = h.fxb;
  d.fzb_absmax = 0.f;
  for (const SegEntry& e : h.seg) d.fzb_absmax = fmaxf(d.fzb_absmax, fabsf(e.fzb));
  d.ndata = h.info.ndata;
  d.nvox = h.info.nvox;
  char* base = (char*)gd->block;
  d.sin_t = (const float*)base;
  d.cos_t = (const float*)(base + off_cos);
  d.seg = (const SegEntry*)(base + off_seg);
  d.sino = (const int2*)(base + off_sino);
  d.ztab = (const uint2*)(base + off_zt);
  ctx->geoms[key] = gd;
  return gd;
}

hipError_t dev_alloc(gfr_ctx* ctx, void** p, size_t bytes) {
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    auto it = ctx->kept.find(bytes);
    if (it != ctx->kept.end()) {
      *p = it->second;
      ctx->kept.erase(it);
      ctx->kept_bytes -= bytes;
      return hipSuccess;
    }
  }
  hipError_t e = hipMalloc(p, bytes);
  if (e == hipErrorOutOfMemory) {
    (void)hipGetLastError();
    dev_trim(ctx);
    e = hipMalloc(p, bytes);
  }
  return e;
}

void dev_free(gfr_ctx* ctx, void* p, size_t bytes) {
  if (!p) return;
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (ctx->kept_bytes + bytes <= ctx->kept_max) {
      ctx->kept.emplace(bytes, p);
      ctx->kept_bytes += bytes;
      return;
    }
  }
  (void)hipFree(p);
}

void dev_trim(gfr_ctx* ctx) {
  std::multimap<size_t, void*> k;
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    k.swap(ctx->kept);
    ctx->kept_bytes = 0;
  }
  for (auto& kv : k) (void)hipFree(kv.second);
}

hipError_t stream_take(gfr_ctx* ctx, hipStream_t* st) {
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (!ctx->kept_streams.empty()) {
      *st = ctx->kept_streams.back();
      ctx->kept_streams.pop_back();
      return hipSuccess;
    }
  }
  return hipStreamCreateWithFlags(st, hipStreamNonBlocking);
}

void stream_give(gfr_ctx* ctx, hipStream_t st) {
  if (!st) return;
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (ctx->kept_max > 0 && ctx->kept_streams.size() < kKeptStreams) {
      ctx->kept_streams.push_back(st);
      return;
    }
  }
  (void)hipStreamDestroy(st);
}

}  // namespace gfr

using namespace gfr;

namespace {

bool bad_subset(gfr_ctx* ctx, const GeomDev* gd, int subset, int nsub) {
  if (nsub < 1 || nsub > gd->host.g.nang || subset < 0 || subset >= nsub) {
    set_err(ctx, GFR_EINVAL, "subset out of range (need 0 <= subset < nsub <= nang)");
    return true;
  }
  return false;
}

// Per-operation forward projection through the context's rotation scratch.
int run_forward(gfr_ctx* ctx, GeomDev* gd, const float* d_img, const FwdOut& o, int subset,
                int nsub, hipStream_t st) {
  const int nav = (gd->dev.nang - subset + nsub - 1) / nsub;
  const size_t per = rotate_floats_per_angle(gd->dev);
  const size_t floats = per * (size_t)std::min(nav, kRotChunk);
  hipError_t e = hipSuccess;
  float* R;
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    R = (float*)scratch_acquire(ctx, ctx->rot, floats * sizeof(float), st, &e);
    if (R) e = launch_forward(gd->dev, d_img, R, floats, o, subset, nsub, nullptr, gd->tune[nsub], st);
    if (R) scratch_release(ctx->rot, st);
  }
  if (e != hipSuccess) return hip_err(ctx, e, "forward projection");
  return GFR_OK;
}

#define GFR_CHECK_CTX(ctx) \
  if (!(ctx)) return GFR_EINVAL
#define GFR_GEOM(ctx, g, gd)             \
  GeomDev* gd = geom_for((ctx), (g));    \
  if (!gd) return GFR_EINVAL
#define GFR_NONNULL(ctx, p, name) \
  if (!(p)) return set_err((ctx), GFR_EINVAL, name " is NULL")

}  // namespace

extern "C" {

int gfr_ctx_create(int device, size_t hbm_budget, gfr_ctx** out) {
  if (!out) return GFR_EINVAL;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return GFR_EHIP;
  if (device < 0 || device >= n) return GFR_EINVAL;
  if (hipSetDevice(device) != hipSuccess) return GFR_EHIP;
  if (noalloc_init(device) != hipSuccess) return GFR_EHIP;
  gfr_ctx* c = new gfr_ctx();
  c->device = device;
  c->budget = hbm_budget;
  c->kept_max = kKeptMax;
  if (const char* env = getenv("GFR_KEPT_MAX_MB")) {
    const long long mb = atoll(env);
    c->kept_max = mb > 0 ? (size_t)mb << 20 : 0;
  }
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return GFR_EHIP;
  }
  if (hipMalloc(&c->d_scratch_u32, 256 * sizeof(unsigned int)) != hipSuccess) {
    (void)hipStreamDestroy(c->stream);
    delete c;
    return GFR_ENOMEM;
  }
  *out = c;
  return GFR_OK;
}

int gfr_ctx_destroy(gfr_ctx* ctx) {
  if (!ctx) return GFR_EINVAL;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  dev_trim(ctx);
  for (hipStream_t st : ctx->kept_streams) (void)hipStreamDestroy(st);
  for (auto& kv : ctx->geoms) {
    if (kv.second->block) (void)hipFree(kv.second->block);
    if (kv.second->pb_cyl) (void)hipFree(kv.second->pb_cyl);
    for (auto& pc : kv.second->ptc0) (void)hipFree(pc.second);
    delete kv.second;
  }
  for (auto& kv : ctx->allocs) (void)hipFree(kv.first);
  (void)hipFree(ctx->d_scratch_u32);
  scratch_free(ctx->rot);
  scratch_free(ctx->pack);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return GFR_OK;
}

const char* gfr_last_error(gfr_ctx* ctx) {
  if (!ctx) return "no context";
  // copy per calling thread: valid until that thread's next call
  static thread_local std::string copy;
  std::lock_guard<std::mutex> lk(ctx->err_mu);
  copy = ctx->err;
  return copy.c_str();
}

void* gfr_ctx_stream(gfr_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int gfr_sync(gfr_ctx* ctx, void* stream) {
  GFR_CHECK_CTX(ctx);
  GFR_HIP(ctx, hipStreamSynchronize(stream_of(ctx, stream)));
  return GFR_OK;
}

int gfr_alloc(gfr_ctx* ctx, size_t bytes, void** dptr) {
  GFR_CHECK_CTX(ctx);
  GFR_NONNULL(ctx, dptr, "dptr");
  *dptr = nullptr;
  std::lock_guard<std::mutex> lk(ctx->mu);
  if (ctx->budget && ctx->in_use + bytes > ctx->budget)
    return set_err(ctx, GFR_ENOMEM, "allocation exceeds the context's HBM budget");
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, bytes ? bytes : 4);
  if (e != hipSuccess) return hip_err(ctx, e, "hipMalloc");
  ctx->allocs[p] = bytes;
  ctx->in_use += bytes;
  *dptr = p;
  return GFR_OK;
}

int gfr_free(gfr_ctx* ctx, void* dptr) {
  GFR_CHECK_CTX(ctx);
  if (!dptr) return GFR_OK;
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    auto it = ctx->allocs.find(dptr);
    if (it == ctx->allocs.end()) return set_err(ctx, GFR_EINVAL, "gfr_free: unknown pointer");
    ctx->in_use -= it->second;
    ctx->allocs.erase(it);
  }
  GFR_HIP(ctx, hipFree(dptr));
  return GFR_OK;
}

size_t gfr_bytes_in_use(gfr_ctx* ctx) { return ctx ? ctx->in_use : 0; }

int gfr_memcpy_h2d(gfr_ctx* ctx, void* dst, const void* src, size_t bytes, void* stream) {
  GFR_CHECK_CTX(ctx);
  if (!bytes) return GFR_OK;
  GFR_NONNULL(ctx, dst, "dst");
  GFR_NONNULL(ctx, src, "src");
  hipStream_t st = stream_of(ctx, stream);
  GFR_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st));
  GFR_HIP(ctx, hipStreamSynchronize(st));
  return GFR_OK;
}

int gfr_memcpy_d2h(gfr_ctx* ctx, void* dst, const void* src, size_t bytes, void* stream) {
  GFR_CHECK_CTX(ctx);
  if (!bytes) return GFR_OK;
  GFR_NONNULL(ctx, dst, "dst");
  GFR_NONNULL(ctx, src, "src");
  hipStream_t st = stream_of(ctx, stream);
  GFR_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st));
  GFR_HIP(ctx, hipStreamSynchronize(st));
  return GFR_OK;
}

int gfr_zero(gfr_ctx* ctx, void* dptr, size_t bytes, void* stream) {
  GFR_CHECK_CTX(ctx);
  GFR_NONNULL(ctx, dptr, "dptr");
  GFR_HIP(ctx, launch_zero(dptr, bytes, stream_of(ctx, stream)));
  return GFR_OK;
}

int gfr_fill_cylinder(gfr_ctx* ctx, const gfr_geometry* g, float val, float* d_img,
                      void* stream) {
  GFR_CHECK_CTX(ctx);
  GFR_GEOM(ctx, g, gd);
  GFR_NONNULL(ctx, d_img, "d_img");
  GFR_HIP(ctx, launch_fill_cylinder(gd->dev, val, d_img, stream_of(ctx, stream)));
  return GFR_OK;
}

int gfr_blur_xy(gfr_ctx* ctx, const gfr_geometry* g, float psf, const float* d_in, float* d_out,
                void* stream) {
  GFR_CHECK_CTX(ctx);
  GFR_GEOM(ctx, g, gd);
  GFR_NONNULL(ctx, d_in, "d_in");
  GFR_NONNULL(ctx, d_out, "d_out");
  if (d_in == d_out) return set_err(ctx, GFR_EINVAL, "gfr_blur_xy: d_in == d_out");
  Taps9 t;
  psf_taps(psf, t.g);
  GFR_HIP(ctx, launch_blur_xy(gd->dev, t, d_in, d_out, stream_of(ctx, stream)));
  return GFR_OK;
}

int gfr_forward_project(gfr_ctx* ctx, const gfr_geometry* g, const float* d_img, float* d_sino,
                        int subset, int nsub, void* stream) {
  GFR_CHECK_CTX(ctx);
  GFR_GEOM(ctx, g, gd);
  GFR_NONNULL(ctx, d_img, "d_img");
  GFR_NONNULL(ctx, d_sino, "d_sino");
  if (bad_subset(ctx, gd, subset, nsub)) return GFR_EINVAL;
  FwdOut o;
  o.proj = d_sino;
  return run_forward(ctx, gd, d_img, o, subset, nsub, stream_of(ctx, stream));
}

int gfr_calib_mask(gfr_ctx* ctx, const gfr_geometry* g, const float* d_proj_blank,
                   const float* d_norm, uint8_t* d_mask, void* stream) {
  GFR_CHECK_CTX(ctx);
  GFR_GEOM(ctx, g, gd);
  GFR_NONNULL(ctx, d_proj_blank, "d_proj_blank");
  GFR_NONNULL(ctx, d_norm, "d_norm");
  GFR_NONNULL(ctx, d_mask, "d_mask");
  GFR_HIP(ctx, launch_calib_mask(gd->dev, d_proj_blank, d_norm, d_mask, stream_of(ctx, stream)));
  return GFR_OK;
}

int gfr_ratio(gfr_ctx* ctx, const gfr_geometry* g, const float* d_proj, const float* d_y,
              const float* d_bg, float* d_corr, int subset, int nsub, void* stream) {
  GFR_CHECK_CTX(ctx);
  GFR_GEOM(ctx, g, gd);
  GFR_NONNULL(ctx, d_proj, "d_proj");
  GFR_NONNULL(ctx, d_y, "d_y");
  GFR_NONNULL(ctx, d_corr, "d_corr");
  if (bad_subset(ctx, gd, subset, nsub)) return GFR_EINVAL;
  GFR_HIP(ctx, launch_ratio(gd->dev, d_proj, d_y, d_bg, GFR_STORE_F32, 1.f, d_corr, nullptr, nullptr, subset, nsub,
                            stream_of(ctx, stream)));
  return GFR_OK;
}

int gfr_forward_ratio(gfr_ctx* ctx, const gfr_geometry* g, const float* d_img, const float* d_y,
                      const float* d_bg, float* d_corr, float* d_proj, int subset, int nsub,
                      void* stream) {
  GFR_CHECK_CTX(ctx);
  GFR_GEOM(ctx, g, gd);
  GFR_NONNULL(ctx, d_img, "d_img");
  GFR_NONNULL(ctx, d_y, "d_y");
  GFR_NONNULL(ctx, d_corr, "d_corr");
  if (bad_subset(ctx, gd, subset, nsub)) return GFR_EINVAL;
  FwdOut o;
  o.proj = d_proj;
  o.corr = d_corr;
  o.y = d_y;
  o.bg = d_bg;
  return run_forward(ctx, gd, d_img, o, subset, nsub, stream_of(ctx, stream));
}

int gfr_back_project(gfr_ctx* ctx, const gfr_geometry* g, const float* d_corr,
                     const uint8_t* d_mask, float* d_numer, float* d_denom, int subset, int nsub,
                     void* stream) {
  GFR_CHECK_CTX(ctx);
  GFR_GEOM(ctx, g, gd);
  GFR_NONNULL(ctx, d_corr, "d_corr");
  GFR_NONNULL(ctx, d_mask, "d_mask");
  GFR_NONNULL(ctx, d_numer, "d_numer");
  GFR_NONNULL(ctx, d_denom, "d_denom");
  if (bad_subset(ctx, gd, subset, nsub)) return GFR_EINVAL;
  // Pack the subset's {ratio, calib} rows, then the session's backprojector.
  hipStream_t st = stream_of(ctx, stream);
  const size_t bytes = packed_bytes(gd->dev, subset, nsub);
  hipError_t e = hipSuccess;
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    float2* c2 = (float2*)scratch_acquire(ctx, ctx->pack, bytes, st, &e);
    if (c2) e = launch_pack(gd->dev, d_corr, d_mask, c2, subset, nsub, st);
    if (c2 && e == hipSuccess)
      e = launch_back_packed(gd->dev, c2, d_numer, d_denom, subset, nsub, nullptr, gd->tune[nsub], st);
    if (c2) scratch_release(ctx->pack, st);
  }
  if (e != hipSuccess) return hip_err(ctx, e, "back projection");
  return GFR_OK;
}

int gfr_sensitivity(gfr_ctx* ctx, const gfr_geometry* g, const uint8_t* d_mask, double* d_sens,
                    int subset, int nsub, void* stream) {
  GFR_CHECK_CTX(ctx);
  GFR_GEOM(ctx, g, gd);
  GFR_NONNULL(ctx, d_mask, "d_mask");
  GFR_NONNULL(ctx, d_sens, "d_sens");
  if (bad_subset(ctx, gd, subset, nsub)) return GFR_EINVAL;
  hipStream_t st = stream_of(ctx, stream);
  const size_t bytes = packed_bytes(gd->dev, subset, nsub);
  hipError_t e = hipSuccess;
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    float2* c2 = (float2*)scratch_acquire(ctx, ctx->pack, bytes, st, &e);
    if (c2) e = launch_pack(gd->dev, nullptr, d_mask, c2, subset, nsub, st);
    if (c2 && e == hipSuccess) e = launch_sensitivity(gd->dev, c2, d_sens, subset, nsub, st);
    if (c2) scratch_release(ctx->pack, st);
  }
  if (e != hipSuccess) return hip_err(ctx, e, "sensitivity");
  return GFR_OK;
}

int gfr_update(gfr_ctx* ctx, const gfr_geometry* g, float psf_bw, const float* d_numer,
               const float* d_denom, float* d_x, float* d_work, void* stream) {
  GFR_CHECK_CTX(ctx);
  GFR_GEOM(ctx, g, gd);
  GFR_NONNULL(ctx, d_numer, "d_numer");
  GFR_NONNULL(ctx, d_denom, "d_denom");
  GFR_NONNULL(ctx, d_x, "d_x");
  GFR_NONNULL(ctx, d_work, "d_work");
  if (g->nzs < 3) return set_err(ctx, GFR_EINVAL, "gfr_update: nzs must be >= 3");
  Taps9 t9;
  Taps5 t5;
  psf_taps(psf_bw, t9.g);
  gz_taps(t5.g);
  GFR_HIP(ctx, launch_update(gd->dev, t9, t5, d_numer, d_denom, d_x, d_work,
                             stream_of(ctx, stream)));
  return GFR_OK;
}

int gfr_jbf3d(gfr_ctx* ctx, const gfr_geometry* g, const float* d_in, const float* d_guide,
              float max_ref, int r, float sigma3d, float sigmai, float* d_out, void* stream) {
  GFR_CHECK_CTX(ctx);
  GFR_GEOM(ctx, g, gd);
  GFR_NONNULL(ctx, d_in, "d_in");
  GFR_NONNULL(ctx, d_guide, "d_guide");
  GFR_NONNULL(ctx, d_out, "d_out");
  if (r < 0 || r > 8) return set_err(ctx, GFR_EINVAL, "gfr_jbf3d: radius must be in [0, 8]");
  GFR_HIP(ctx, launch_jbf(gd->dev, d_in, d_guide, nullptr, max_ref, r, sigma3d, sigmai, d_out,
                          stream_of(ctx, stream)));
  return GFR_OK;
}

int gfr_blend(gfr_ctx* ctx, const gfr_geometry* g, float coef, const float* d_filt, float* d_x,
              void* stream) {
  GFR_CHECK_CTX(ctx);
  GFR_GEOM(ctx, g, gd);
  GFR_NONNULL(ctx, d_filt, "d_filt");
  GFR_NONNULL(ctx, d_x, "d_x");
  GFR_HIP(ctx, launch_blend(gd->dev, coef, d_filt, d_x, stream_of(ctx, stream)));
  return GFR_OK;
}

int gfr_image_max(gfr_ctx* ctx, const gfr_geometry* g, const float* d_x, float* h_max,
                  void* stream) {
  GFR_CHECK_CTX(ctx);
  GFR_GEOM(ctx, g, gd);
  GFR_NONNULL(ctx, d_x, "d_x");
  GFR_NONNULL(ctx, h_max, "h_max");
  hipStream_t st = stream_of(ctx, stream);
  GFR_HIP(ctx, launch_max(gd->dev, d_x, ctx->d_scratch_u32, st));
  unsigned int bits = 0;
  GFR_HIP(ctx, hipMemcpyAsync(&bits, ctx->d_scratch_u32, 4, hipMemcpyDeviceToHost, st));
  GFR_HIP(ctx, hipStreamSynchronize(st));
  memcpy(h_max, &bits, 4);
  return GFR_OK;
}

// ---------------------------------------------------------------------------
// VLOR: the static part of the soft assignment (valid records, BLANKT and the
// bin -> record lists) is built once; each pass recomputes 1/(s1+s2) per
// record and sums each touched bin's contributions in record order.
// ---------------------------------------------------------------------------
int gfr_vlor_create(gfr_ctx* ctx, const gfr_geometry* g, const int32_t* rec, long long nrec,
                    const float* d_yT, gfr_vlor** out) {
  GFR_CHECK_CTX(ctx);
  return vlor_create_on(ctx, g, rec, nrec, d_yT, ctx->stream, out);
}

int gfr_vlor_assign(gfr_ctx* ctx, gfr_vlor* v, const float* d_proj, float* d_dv, int subset,
                    int nsub, void* stream) {
  GFR_CHECK_CTX(ctx);
  GFR_NONNULL(ctx, v, "vlor");
  GFR_NONNULL(ctx, d_proj, "d_proj");
  GFR_NONNULL(ctx, d_dv, "d_dv");
  GFR_GEOM(ctx, &v->g, gd);
  if (bad_subset(ctx, gd, subset, nsub)) return GFR_EINVAL;
  hipStream_t st = stream_of(ctx, stream);
  GFR_HIP(ctx, launch_vlor_event(v->d_i1, v->d_i2, v->d_blankt, d_proj, v->d_sinv, v->d_valor,
                                 v->nev, 0, st));
  GFR_HIP(ctx, launch_vlor_zero_subset(gd->dev, d_dv, subset, nsub, st));
  GFR_HIP(ctx, launch_vlor_bins(gd->dev, v->d_bins, v->d_start, v->d_ev, v->nbins, d_proj,
                                v->d_sinv, v->d_valor, d_dv, subset, nsub, st));
  return GFR_OK;
}

int gfr_vlor_destroy(gfr_vlor* v) {
  if (!v) return GFR_EINVAL;
  if (v->ctx) {
    std::lock_guard<std::mutex> lk(v->ctx->mu);
    if (v->ctx->in_use >= v->bytes) v->ctx->in_use -= v->bytes;
  }
  void* ps[] = {v->d_i1, v->d_i2, v->d_blankt, v->d_sinv, v->d_valor, v->d_bins, v->d_start, v->d_ev};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  delete v;
  return GFR_OK;
}

}  // extern "C"

extern "C" {

int gfr_tuning_get(gfr_ctx* ctx, const gfr_geometry* g, int nsub, gfr_tuning* out) {
  GFR_CHECK_CTX(ctx);
  GFR_GEOM(ctx, g, gd);
  GFR_NONNULL(ctx, out, "tuning");
  if (nsub < 1 || nsub > g->nang) return set_err(ctx, GFR_EINVAL, "nsub out of range");
  const Tune t = tune_for(ctx, gd, nsub);
  out->back_zb = t.back_zb;
  out->back_lds = t.back_lds;
  out->back_waves = t.back_waves;
  out->zsum_sg = t.zsum_sg;
  out->rot_z = t.rot_z;
  out->back_bw = t.back_bw;
  out->ms_back = t.ms_back;
  out->ms_zsum = t.ms_zsum;
  out->back_ahead = t.back_ahead;
  return t.tuned ? 1 : 0;
}

int gfr_tuning_set(gfr_ctx* ctx, const gfr_geometry* g, int nsub, const gfr_tuning* in) {
  GFR_CHECK_CTX(ctx);
  GFR_GEOM(ctx, g, gd);
  GFR_NONNULL(ctx, in, "tuning");
  if (nsub < 1 || nsub > g->nang) return set_err(ctx, GFR_EINVAL, "nsub out of range");
  const int zb = in->back_zb, sg = in->zsum_sg;
  if (!(zb == 1 || zb == 2 || zb == 4 || zb == 5 || zb == 6 || zb == 8 || zb == 16) ||
      !(sg == 4 || sg == 6 || sg == 8))
    return set_err(ctx, GFR_EINVAL,
                   "back_zb must be 1, 2, 4, 5, 6, 8 or 16 and zsum_sg 4, 6 or 8");
  if (in->back_lds != 0 && in->back_lds != 1)
    return set_err(ctx, GFR_EINVAL, "back_lds must be 0 or 1");
  if (zb < 4 && in->back_lds)
    return set_err(ctx, GFR_EINVAL, "back_zb 1 or 2 only with cached reads (back_lds 0)");
  if (in->back_waves != 0 && !(in->back_waves == 6 && in->back_zb == 8 && !in->back_lds))
    return set_err(ctx, GFR_EINVAL, "back_waves: 0, or 6 with 8 slices and direct reads");
  if (in->rot_z != 4 && in->rot_z != 16 && in->rot_z != 32)
    return set_err(ctx, GFR_EINVAL, "rot_z must be 4, 16 or 32");
  if (in->back_bw != 4 && !(in->back_bw == 1 && !in->back_lds && in->back_waves == 0))
    return set_err(ctx, GFR_EINVAL, "back_bw: 4, or 1 with cached reads and back_waves 0");
  const int ahead = in->back_ahead == 0 ? 1 : in->back_ahead;
  if (!(ahead == 1 || ahead == 2 || ahead == 4) ||
      (ahead > 1 && !(zb == 1 && !in->back_lds && in->back_waves == 0)))
    return set_err(ctx, GFR_EINVAL,
                   "back_ahead: 1 (or 0), or 2 or 4 with back_zb 1 and cached reads");
  std::lock_guard<std::mutex> lk(ctx->mu);
  Tune& t = gd->tune[nsub];
  t.back_zb = in->back_zb;
  t.back_lds = in->back_lds;
  t.back_waves = in->back_waves;
  t.zsum_sg = in->zsum_sg;
  t.rot_z = in->rot_z;
  t.back_bw = in->back_bw;
  t.back_ahead = ahead;
  t.tuned = true;
  t.ms_back = t.ms_zsum = 0.f;
  return GFR_OK;
}

}  // extern "C"

namespace gfr {

int vlor_create_on(gfr_ctx* ctx, const gfr_geometry* g, const int32_t* rec, long long nrec,
                   const float* d_yT, hipStream_t st, gfr_vlor** out) {
  GFR_GEOM(ctx, g, gd);
  GFR_NONNULL(ctx, out, "out");
  GFR_NONNULL(ctx, d_yT, "d_yT");
  if (nrec < 0 || (nrec > 0 && !rec)) return set_err(ctx, GFR_EINVAL, "gfr_vlor_create: records");
  // list offsets are int: 2 * nrec must stay below 2^31
  if (nrec >= (1LL << 30)) return set_err(ctx, GFR_EINVAL, "gfr_vlor_create: too many records");
  *out = nullptr;
  const long long nd = gd->host.info.ndata;
  if (nd >= (1LL << 31)) return set_err(ctx, GFR_EINVAL, "gfr_vlor_create: sinogram too large");
  // Valid records (GFIRST.cu:235-237) and the bin -> record lists are built on
  // the device from the raw records (build_vlor_lists); buffers are sized for
  // every record being valid.
  const size_t ne = 2 * (size_t)nrec;
  gfr_vlor* v = new gfr_vlor();
  v->ctx = ctx;
  v->g = *g;
  auto dalloc = [&](void** p, size_t b) -> hipError_t {
    v->bytes += b;
    return hipMalloc(p, b ? b : 4);
  };
  float* d_v0 = nullptr;
  int* d_rec = nullptr;
  hipError_t e = hipSuccess;
  if (e == hipSuccess) e = dalloc((void**)&v->d_i1, sizeof(int) * nrec);
  if (e == hipSuccess) e = dalloc((void**)&v->d_i2, sizeof(int) * nrec);
  if (e == hipSuccess) e = dalloc((void**)&v->d_blankt, sizeof(float) * nd);
  if (e == hipSuccess) e = dalloc((void**)&v->d_sinv, sizeof(float) * nrec);
  if (e == hipSuccess) e = dalloc((void**)&v->d_valor, sizeof(float) * nrec);
  if (e == hipSuccess) e = dalloc((void**)&v->d_bins, sizeof(long long) * ne);
  if (e == hipSuccess) e = dalloc((void**)&v->d_start, sizeof(int) * (ne + 1));
  if (e == hipSuccess) e = dalloc((void**)&v->d_ev, sizeof(int) * ne);
  if (e == hipSuccess) e = hipMalloc((void**)&d_v0, sizeof(float) * nd);
  if (e == hipSuccess && nrec) e = hipMalloc((void**)&d_rec, 2 * sizeof(int) * nrec);
  if (e == hipSuccess && nrec)
    e = hipMemcpyAsync(d_rec, rec, 2 * sizeof(int) * nrec, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) {
    long long nev = 0, nbins = 0;
    e = build_vlor_lists(d_rec, nrec, nd, v->d_i1, v->d_i2, v->d_bins, v->d_start, v->d_ev, &nev,
                         &nbins, st);
    v->nev = nev;
    v->nbins = nbins;
  }
  if (d_rec) (void)hipFree(d_rec);
  const long long nev = v->nev;
  if (e == hipSuccess && nrec == 0) {
    const int zero = 0;
    e = hipMemcpyAsync(v->d_start, &zero, sizeof(int), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
  }
  // datos_v0 counts on the device (GFIRST.cu:238-239)
  if (e == hipSuccess) e = launch_zero(d_v0, sizeof(float) * nd, st);
  if (e == hipSuccess) e = launch_vlor_count(v->d_i1, v->d_i2, nev, d_v0, st);
  // BLANKT = T/(RES*v0) where v0 > 0.01 (GFIRST.cu:244-247); T already x RES.
  if (e == hipSuccess) e = launch_vlor_blankt(d_yT, d_v0, (float)g->res, v->d_blankt, nd, st);
  // valor per record is static: BLANKT[i1] + BLANKT[i2].
  if (e == hipSuccess)
    e = launch_vlor_event(v->d_i1, v->d_i2, v->d_blankt, v->d_blankt, v->d_sinv, v->d_valor, nev,
                          1, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (d_v0) (void)hipFree(d_v0);
  if (e != hipSuccess) {
    v->ctx = nullptr;  // its bytes were never counted in the context
    gfr_vlor_destroy(v);
    return hip_err(ctx, e, "gfr_vlor_create");
  }
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->in_use += v->bytes;
  }
  *out = v;
  return GFR_OK;
}

}  // namespace gfr
