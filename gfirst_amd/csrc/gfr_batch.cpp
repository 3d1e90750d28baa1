// gfr_batch.cpp — many doubles reconstructions of one geometry advanced
// together (the frames of a dynamic study, or studies of one scanner).
//
// Per frame, only what is the frame's own stays resident: the image estimate,
// the prompts and background in the storage format, and a device-side subset
// index. Everything a pass needs transiently (blurred image, rotation stage,
// packed ratio, numerator/denominator, FP64 partials) lives in per-lane
// scratch, and frames are spread over `lanes` independent chains so the
// kernels of different frames overlap on the device.
//
// One step = one subset pass of every active frame, recorded once into a
// hipGraph: the kernels take the subset from device memory (the subset index
// of frame f is sub[f]) and a final kernel advances every index, so the same
// graph serves every step until a frame joins or finishes. The pass itself is
// the reference's (GFIRST.cu:324-340 blur, 341-469 forward/ratio/backward,
// 482-502 update), identical kernel for kernel to gfr_recon_step in mode 1,
// except that frames at the same subset of the same subset count share their
// backprojection: up to `group` of them write their ratios into one set of
// interleaved rows and one launch backprojects them all (the per-view
// geometry, calibration interpolation and row loads are shared; each frame's
// sums keep their bits).
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "gfr_internal.h"

using namespace gfr;

namespace {

struct Lane {
  hipStream_t st = nullptr;
  hipEvent_t join = nullptr;
  float *xb = nullptr, *num = nullptr, *den = nullptr, *work = nullptr, *R = nullptr;
  float2* corr2 = nullptr;
  float* corr4 = nullptr;                // interleaved rows of a frame group
  float* num_g[kMaxGroup] = {};          // numerators / denominators of a group
  float* den_g[kMaxGroup] = {};          // ([0] = num / den)
  double* llh_part = nullptr;
  size_t r_floats = 0;
};

struct Frame {
  bool used = false;
  int niter = 0, nsub = 1;
  long long passes = 0;
  float* x = nullptr;
  void* y = nullptr;
  void* bg = nullptr;
  size_t bytes = 0;
  Tune tune;
  bool done() const { return passes >= (long long)niter * nsub; }
};

size_t store_bytes(int storage) { return storage == GFR_STORE_F32 ? 4 : 2; }

// Frames per shared backprojection (0 asks for the largest).
int group_of(const gfr_batch_params* p) { return p->group == 0 ? kMaxGroup : p->group; }

struct Sizes {
  size_t img, corr2, corr4, rot, llh, frame_nobg, frame_bg, shared;
};

Sizes sizes_of(const GeomHost& h, const gfr_batch_params* p) {
  Sizes z{};
  const size_t nav = (size_t)(h.g.nang + p->min_nsub - 1) / p->min_nsub;
  z.img = sizeof(float) * (size_t)h.info.nvox;
  z.corr2 = sizeof(float2) * nav * ((size_t)h.info.nsinogs + 2 * (size_t)h.g.nseg) *
            ((size_t)h.g.nrad + 2);  // packed rows: NRAD + 2 entries
  const int grp = group_of(p);
  z.corr4 = grp > 1 ? 2 * z.corr2 : 0;
  z.rot = sizeof(float) * (size_t)(h.info.lambda1 - h.info.lambda0 + 1) * h.g.nzs * h.g.nrad *
          std::min(nav, (size_t)kRotChunk);
  z.llh = sizeof(double) * (size_t)llh_capacity(h, p->min_nsub) + 64;
  const size_t sino = store_bytes(p->storage) * (size_t)h.info.ndata;
  z.frame_nobg = z.img + sino;
  z.frame_bg = z.img + 2 * sino;
  const size_t per_lane =
      (4 + 2 * (size_t)(grp - 1)) * z.img + z.corr2 + z.corr4 + z.rot + z.llh;
  const size_t slots = (size_t)p->max_frames * (3 * sizeof(int) + sizeof(double)) + 256;
  z.shared = (size_t)p->lanes * per_lane + (size_t)h.info.ndata /* mask */ + slots;
  return z;
}

}  // namespace

struct gfr_batch {
  gfr_ctx* ctx = nullptr;
  gfr_geometry g{};
  gfr_batch_params p{};
  GeomDev* gd = nullptr;
  hipStream_t st = nullptr;  // origin stream: steps are ordered on it
  hipEvent_t fork = nullptr;
  std::vector<Lane> lanes;
  std::vector<Frame> frames;
  uint8_t* mask = nullptr;
  int* d_sub = nullptr;     // [max_frames] subset index per slot
  int* d_nsub = nullptr;    // [max_frames]
  int* d_slots = nullptr;   // [max_frames] active slots of the current graph
  double* d_llh = nullptr;  // [max_frames] FP64 log-likelihood of the last pass
  std::vector<void*> owned;
  size_t bytes = 0;
  bool use_graph = true;
  bool dirty = true;  // membership changed since the graph was built
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<int> active;  // slots in the current step program
  long long launches = 0, builds = 0;
};

namespace {

int bfail(gfr_batch* b, int code) {
  gfr_batch_destroy(b);
  return code;
}

void drop_graph(gfr_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

// Enqueues the blur and forward projection of frame f on lane L (kernels
// read the subset index from d_sub[f]); the ratio goes to the lane's packed
// rows (chan < 0) or to channel chan of its interleaved group rows.
hipError_t enqueue_forward(gfr_batch* b, int f, const Lane& L, int chan) {
  const DevGeom& G = b->gd->dev;
  const Frame& fr = b->frames[f];
  const int* dsub = b->d_sub + f;
  Taps9 tf;
  psf_taps(b->p.psf_fw, tf.g);
  hipError_t e = launch_blur_xy(G, tf, fr.x, L.xb, L.st);
  FwdOut fo;
  if (chan < 0) {
    fo.corr2 = L.corr2;
  } else {
    fo.corr4 = L.corr4;
    fo.chan = chan;
  }
  fo.mask = b->mask;
  fo.y = fr.y;
  fo.bg = fr.bg;
  fo.storage = b->p.storage;
  fo.yscale = (float)b->g.res;
  fo.llh_part = L.llh_part;
  if (e == hipSuccess)
    e = launch_forward(G, L.xb, L.R, L.r_floats, fo, 0, fr.nsub, dsub, fr.tune, L.st);
  if (e == hipSuccess)
    e = launch_llh_sum(G, L.llh_part, 0, fr.nsub, dsub, fr.tune.zsum_sg, b->d_llh + f, L.st);
  return e;
}

hipError_t enqueue_update(gfr_batch* b, int f, const float* num, const float* den,
                          const Lane& L) {
  Taps9 tb;
  Taps5 tz;
  psf_taps(b->p.psf_bw, tb.g);
  gz_taps(tz.g);
  return launch_update(b->gd->dev, tb, tz, num, den, b->frames[f].x, L.work, L.st);
}

// Enqueues one subset pass of the frames of a group on lane L: every
// member's forward projection, one backprojection, every member's update.
hipError_t enqueue_group(gfr_batch* b, const std::vector<int>& grp, const Lane& L) {
  const DevGeom& G = b->gd->dev;
  const Frame& f0 = b->frames[grp[0]];
  hipError_t e = hipSuccess;
  if (grp.size() == 1) {
    e = enqueue_forward(b, grp[0], L, -1);
    if (e == hipSuccess)
      e = launch_back_packed(G, L.corr2, L.num, L.den, 0, f0.nsub, b->d_sub + grp[0], f0.tune,
                             L.st);
    if (e == hipSuccess) e = enqueue_update(b, grp[0], L.num, L.den, L);
    return e;
  }
  for (size_t i = 0; i < grp.size() && e == hipSuccess; i++)
    e = enqueue_forward(b, grp[i], L, (int)i);
  if (e == hipSuccess)
    e = launch_back_multi(G, L.corr4, L.num_g, L.den_g, (int)grp.size(), 0, f0.nsub,
                          b->d_sub + grp[0], f0.tune.multi_zb == 6 ? 6 : 4, L.st);
  for (size_t i = 0; i < grp.size() && e == hipSuccess; i++)
    e = enqueue_update(b, grp[i], L.num_g[i], L.den_g[i], L);
  return e;
}

// Work of one pass of a group, for spreading groups over lanes (a shared
// backprojection costs about 0.7 of the separate ones per frame).
double group_cost(const gfr_batch* b, const std::vector<int>& grp) {
  const int S = b->frames[grp[0]].nsub;
  const double one = (double)((b->g.nang + S - 1) / S) + 2.0;
  return grp.size() == 1 ? one : 0.7 * one * (double)grp.size();
}

// Frames at the same subset of the same subset count (their subset indices
// stay equal while the membership is unchanged: every active frame advances
// once per step) form groups of at most `group`.
std::vector<std::vector<int>> make_groups(const gfr_batch* b) {
  std::vector<int> order = b->active;
  auto key = [&](int f) {
    const Frame& fr = b->frames[f];
    return std::make_pair(fr.nsub, (int)(fr.passes % fr.nsub));
  };
  std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return key(a) < key(c); });
  const int gmax = group_of(&b->p);
  std::vector<std::vector<int>> groups;
  for (int f : order) {
    if (!groups.empty() && (int)groups.back().size() < gmax && key(groups.back()[0]) == key(f))
      groups.back().push_back(f);
    else
      groups.push_back({f});
  }
  return groups;
}

// Enqueues one step (every active frame, then the index advance) on the
// origin stream; under capture this records the step program.
hipError_t enqueue_step(gfr_batch* b) {
  const int nl = (int)b->lanes.size();
  std::vector<std::vector<int>> groups = make_groups(b);
  std::stable_sort(groups.begin(), groups.end(), [&](const std::vector<int>& a,
                                                     const std::vector<int>& c) {
    return group_cost(b, a) > group_cost(b, c);
  });
  std::vector<double> load(nl, 0.0);
  std::vector<std::vector<const std::vector<int>*>> per(nl);
  for (const auto& grp : groups) {
    const int l = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    load[l] += group_cost(b, grp);
    per[l].push_back(&grp);
  }
  hipError_t e = hipEventRecord(b->fork, b->st);
  for (int l = 0; l < nl && e == hipSuccess; l++) {
    if (per[l].empty()) continue;
    Lane& L = b->lanes[l];
    e = hipStreamWaitEvent(L.st, b->fork, 0);
    for (size_t i = 0; i < per[l].size() && e == hipSuccess; i++)
      e = enqueue_group(b, *per[l][i], L);
    if (e == hipSuccess) e = hipEventRecord(L.join, L.st);
    if (e == hipSuccess) e = hipStreamWaitEvent(b->st, L.join, 0);
  }
  if (e == hipSuccess)
    e = launch_advance(b->d_sub, b->d_nsub, b->d_slots, (int)b->active.size(), b->st);
  return e;
}

// Rebuilds the active list (and the device copy the advance kernel reads)
// and, in graph mode, records and instantiates the step program.
int rebuild(gfr_batch* b) {
  gfr_ctx* ctx = b->ctx;
  drop_graph(b);
  b->active.clear();
  for (int f = 0; f < (int)b->frames.size(); f++)
    if (b->frames[f].used && !b->frames[f].done()) b->active.push_back(f);
  if (!b->active.empty())
    GFR_HIP(ctx, hipMemcpyAsync(b->d_slots, b->active.data(), sizeof(int) * b->active.size(),
                                hipMemcpyHostToDevice, b->st));
  GFR_HIP(ctx, hipStreamSynchronize(b->st));
  b->dirty = false;
  if (!b->use_graph || b->active.empty()) return GFR_OK;
  GFR_HIP(ctx, hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
  hipError_t e = enqueue_step(b);
  hipGraph_t gr = nullptr;
  hipError_t e2 = hipStreamEndCapture(b->st, &gr);
  if (e == hipSuccess) e = e2;
  if (e != hipSuccess) {
    if (gr) (void)hipGraphDestroy(gr);
    return hip_err(ctx, e, "gfr_batch graph capture");
  }
  b->graph = gr;
  GFR_HIP(ctx, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  b->builds++;
  return GFR_OK;
}

// Best of two timed group backprojections (after a warm-up) of subset 0.
float time_group_back(gfr_batch* b, Lane& L, int nsub, int zb, hipError_t* err) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  float best = 1e30f;
  for (int rep = 0; rep < 3 && e == hipSuccess; rep++) {
    e = hipEventRecord(e0, b->st);
    if (e == hipSuccess)
      e = launch_back_multi(b->gd->dev, L.corr4, L.num_g, L.den_g, group_of(&b->p), 0, nsub,
                            nullptr, zb, b->st);
    if (e == hipSuccess) e = hipEventRecord(e1, b->st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e == hipSuccess && rep > 0 && ms < best) best = ms;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  *err = e;
  return best;
}

bool any_positive(const float* v, long long n) {
  for (long long i = 0; i < n; i++)
    if (!(v[i] <= 0.f)) return true;
  return false;
}

}  // namespace

extern "C" {

long long gfr_batch_footprint(const gfr_geometry* g, const gfr_batch_params* p) {
  GeomHost h;
  if (!p || p->lanes < 1 || p->min_nsub < 1 || p->max_frames < 1 || p->group < 0 ||
      p->group > kMaxGroup || derive_geometry(g, &h, nullptr) != GFR_OK)
    return GFR_EINVAL;
  return (long long)sizes_of(h, p).shared;
}

long long gfr_batch_frame_footprint(const gfr_geometry* g, const gfr_batch_params* p, int has_bg) {
  GeomHost h;
  if (!p || p->min_nsub < 1 || derive_geometry(g, &h, nullptr) != GFR_OK) return GFR_EINVAL;
  const Sizes z = sizes_of(h, p);
  return (long long)(has_bg ? z.frame_bg : z.frame_nobg);
}

int gfr_batch_create(gfr_ctx* ctx, const gfr_geometry* g, const gfr_batch_params* p,
                     const float* norm, gfr_batch** out) {
  if (!ctx || !out) return GFR_EINVAL;
  *out = nullptr;
  if (!p || !norm) return set_err(ctx, GFR_EINVAL, "batch params and norm are required");
  GeomDev* gd = geom_for(ctx, g);
  if (!gd) return GFR_EINVAL;
  if (p->storage < GFR_STORE_F32 || p->storage > GFR_STORE_BF16)
    return set_err(ctx, GFR_EINVAL, "unknown storage");
  if (p->max_frames < 1 || p->max_frames > 65536)
    return set_err(ctx, GFR_EINVAL, "max_frames must be in [1, 65536]");
  if (p->lanes < 1 || p->lanes > 16) return set_err(ctx, GFR_EINVAL, "lanes must be in [1, 16]");
  if (p->min_nsub < 1 || p->min_nsub > g->nang)
    return set_err(ctx, GFR_EINVAL, "min_nsub must be in [1, nang]");
  if (p->group < 0 || p->group > kMaxGroup)
    return set_err(ctx, GFR_EINVAL, "group must be in [0, 3]");
  if (g->nzs < 3) return set_err(ctx, GFR_EINVAL, "nzs must be >= 3");
  const GeomHost& h = gd->host;
  const Sizes z = sizes_of(h, p);
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (ctx->budget && ctx->in_use + z.shared > ctx->budget)
      return set_err(ctx, GFR_ENOMEM, "batch exceeds the context's HBM budget");
    ctx->in_use += z.shared;
  }
  gfr_batch* b = new gfr_batch();
  b->ctx = ctx;
  b->g = *g;
  b->p = *p;
  b->gd = gd;
  b->bytes = z.shared;
  b->frames.resize(p->max_frames);
  if (hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess)
    return bfail(b, set_err(ctx, GFR_EHIP, "batch stream/event create"));
  auto A = [&](void** ptr, size_t n) -> bool {
    if (hipMalloc(ptr, n ? n : 4) != hipSuccess) return false;
    b->owned.push_back(*ptr);
    return true;
  };
  bool ok = A((void**)&b->mask, (size_t)h.info.ndata) &&
            A((void**)&b->d_sub, sizeof(int) * p->max_frames) &&
            A((void**)&b->d_nsub, sizeof(int) * p->max_frames) &&
            A((void**)&b->d_slots, sizeof(int) * p->max_frames) &&
            A((void**)&b->d_llh, sizeof(double) * p->max_frames);
  b->lanes.resize(p->lanes);
  for (Lane& L : b->lanes) {
    if (!ok) break;
    ok = hipStreamCreateWithFlags(&L.st, hipStreamNonBlocking) == hipSuccess &&
         hipEventCreateWithFlags(&L.join, hipEventDisableTiming) == hipSuccess &&
         A((void**)&L.xb, z.img) && A((void**)&L.num, z.img) && A((void**)&L.den, z.img) &&
         A((void**)&L.work, z.img) && A((void**)&L.corr2, z.corr2) && A((void**)&L.R, z.rot) &&
         A((void**)&L.llh_part, z.llh);
    L.num_g[0] = L.num;
    L.den_g[0] = L.den;
    if (ok && z.corr4) {
      ok = A((void**)&L.corr4, z.corr4);
      for (int i = 1; i < group_of(p) && ok; i++)
        ok = A((void**)&L.num_g[i], z.img) && A((void**)&L.den_g[i], z.img);
      // defined contents for the first timing of the group backprojector
      if (ok) ok = launch_zero(L.corr4, z.corr4, b->st) == hipSuccess;
    }
    L.r_floats = z.rot / sizeof(float);
  }
  if (!ok) return bfail(b, set_err(ctx, GFR_ENOMEM, "batch allocation failed"));
  // Shared calibration mask from the projection of the unblurred uniform
  // cylinder and the normalization (GFIRST.cu:155-181, 453-458).
  const DevGeom& G = gd->dev;
  float *pb = nullptr, *dn = nullptr;
  const size_t sino = sizeof(float) * (size_t)h.info.ndata;
  if (hipMalloc((void**)&pb, sino) != hipSuccess) return bfail(b, set_err(ctx, GFR_ENOMEM, "batch staging"));
  if (hipMalloc((void**)&dn, sino) != hipSuccess) {
    (void)hipFree(pb);
    return bfail(b, set_err(ctx, GFR_ENOMEM, "batch staging"));
  }
  Lane& L0 = b->lanes[0];
  FwdOut fo;
  fo.proj = pb;
  hipError_t e = launch_fill_cylinder(G, 1.f, L0.xb, b->st);
  if (e == hipSuccess)
    e = launch_forward(G, L0.xb, L0.R, L0.r_floats, fo, 0, 1, nullptr, Tune(), b->st);
  if (e == hipSuccess) e = hipMemcpyAsync(dn, norm, sino, hipMemcpyHostToDevice, b->st);
  if (e == hipSuccess) e = launch_calib_mask(G, pb, dn, b->mask, b->st);
  if (e == hipSuccess) e = launch_zero(b->d_sub, sizeof(int) * p->max_frames, b->st);
  if (e == hipSuccess) e = hipStreamSynchronize(b->st);
  (void)hipFree(pb);
  (void)hipFree(dn);
  if (e != hipSuccess) return bfail(b, hip_err(ctx, e, "gfr_batch_create"));
  *out = b;
  return GFR_OK;
}

int gfr_batch_add(gfr_batch* b, int niter, int nsub, const float* prompts, const float* bg,
                  int* frame) {
  if (!b || !frame) return GFR_EINVAL;
  gfr_ctx* ctx = b->ctx;
  *frame = -1;
  if (!prompts) return set_err(ctx, GFR_EINVAL, "prompts are required");
  if (niter < 0 || nsub < b->p.min_nsub || nsub > b->g.nang)
    return set_err(ctx, GFR_EINVAL, "niter >= 0 and min_nsub <= nsub <= nang required");
  int f = -1;
  for (int i = 0; i < (int)b->frames.size(); i++)
    if (!b->frames[i].used) {
      f = i;
      break;
    }
  if (f < 0) return set_err(ctx, GFR_ESTATE, "no free frame slot");
  const GeomHost& h = b->gd->host;
  const long long nd = h.info.ndata;
  // A background given in host memory that is all zero is dropped (as
  // gfr_recon_create does); device-memory backgrounds are kept as given.
  const bool bg_dev = is_device_ptr(bg);
  const bool has_bg = bg && (bg_dev || any_positive(bg, nd));
  const Sizes z = sizes_of(h, &b->p);
  const size_t need = has_bg ? z.frame_bg : z.frame_nobg;
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (ctx->budget && ctx->in_use + need > ctx->budget)
      return set_err(ctx, GFR_ENOMEM, "frame exceeds the context's HBM budget");
    ctx->in_use += need;
  }
  Frame fr;
  fr.niter = niter;
  fr.nsub = nsub;
  fr.bytes = need;
  const size_t sb = store_bytes(b->p.storage) * (size_t)nd;
  bool ok = hipMalloc((void**)&fr.x, z.img) == hipSuccess;
  ok = ok && hipMalloc(&fr.y, sb) == hipSuccess;
  if (ok && has_bg) ok = hipMalloc(&fr.bg, sb) == hipSuccess;
  float* stage = nullptr;  // nd floats, then the non-finite flag
  const size_t nd_al = ((size_t)nd + 63) & ~(size_t)63;
  if (ok) ok = hipMalloc((void**)&stage, sizeof(float) * nd_al + 256) == hipSuccess;
  auto undo = [&](int code) {
    if (stage) (void)hipFree(stage);
    if (fr.x) (void)hipFree(fr.x);
    if (fr.y) (void)hipFree(fr.y);
    if (fr.bg) (void)hipFree(fr.bg);
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->in_use -= need;
    return code;
  };
  if (!ok) return undo(set_err(ctx, GFR_ENOMEM, "frame allocation failed"));
  const DevGeom& G = b->gd->dev;
  hipStream_t st = b->st;
  // Stored unscaled in the storage format; kernels scale by RES on load
  // (GFIRST.cu:224-260). Initial estimate: uniform cylinder (GFIRST.cu:282-288).
  // A NaN or infinite count fails the frame (as gfr_recon_create).
  unsigned int* bad = (unsigned int*)(stage + nd_al);
  hipError_t e = launch_zero(bad, sizeof(unsigned int), st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(stage, prompts, sizeof(float) * (size_t)nd, hipMemcpyDefault, st);
  if (e == hipSuccess) e = launch_any_nonfinite(stage, nd, bad, st);
  if (e == hipSuccess) e = launch_convert_store(stage, fr.y, b->p.storage, nd, st);
  if (e == hipSuccess && has_bg)
    e = hipMemcpyAsync(stage, bg, sizeof(float) * (size_t)nd, hipMemcpyDefault, st);
  if (e == hipSuccess && has_bg) e = launch_any_nonfinite(stage, nd, bad, st);
  if (e == hipSuccess && has_bg) e = launch_convert_store(stage, fr.bg, b->p.storage, nd, st);
  if (e == hipSuccess) e = launch_fill_cylinder(G, 1.f, fr.x, st);
  const int cur[2] = {0, nsub};  // host source of the two copies, alive until the sync
  if (e == hipSuccess) e = hipMemcpyAsync(b->d_sub + f, &cur[0], sizeof(int), hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(b->d_nsub + f, &cur[1], sizeof(int), hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return undo(hip_err(ctx, e, "gfr_batch_add"));
  unsigned int hbad = 0;
  e = hipMemcpyAsync(&hbad, bad, sizeof(hbad), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return undo(hip_err(ctx, e, "gfr_batch_add"));
  if (hbad) return undo(set_err(ctx, GFR_EINVAL, "non-finite value in the frame's sinograms"));
  // Tile shapes for this subset count, measured on the first lane's scratch.
  Tune t;
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    t = b->gd->tune[nsub];
  }
  const char* env = getenv("GFR_AUTOTUNE");
  if (!t.tuned && !(env && env[0] == '0')) {
    Lane& L = b->lanes[0];
    FwdOut fo;
    fo.corr2 = L.corr2;
    fo.mask = b->mask;
    fo.y = fr.y;
    fo.bg = fr.bg;
    fo.storage = b->p.storage;
    fo.yscale = (float)b->g.res;
    fo.llh_part = L.llh_part;
    e = tune_shapes(G, nsub, fo, fr.x, L.R, L.r_floats, L.corr2, L.num, L.den, st, &t);
    if (e != hipSuccess) return undo(hip_err(ctx, e, "gfr_batch_add tuning"));
    std::lock_guard<std::mutex> lk(ctx->mu);
    b->gd->tune[nsub] = t;
  }
  // Slices of the group backprojector, measured once per subset count.
  if (t.multi_zb == 0 && group_of(&b->p) > 1 && !(env && env[0] == '0')) {
    Lane& L = b->lanes[0];
    float best = 1e30f;
    for (int zb : {4, 6}) {
      const float ms = time_group_back(b, L, nsub, zb, &e);
      if (e != hipSuccess) return undo(hip_err(ctx, e, "gfr_batch_add tuning"));
      if (ms < best) {
        best = ms;
        t.multi_zb = zb;
      }
    }
    std::lock_guard<std::mutex> lk(ctx->mu);
    b->gd->tune[nsub].multi_zb = t.multi_zb;
  }
  (void)hipFree(stage);
  stage = nullptr;
  fr.tune = t;
  fr.used = true;
  b->frames[f] = fr;
  b->dirty = true;
  *frame = f;
  return GFR_OK;
}

int gfr_batch_step(gfr_batch* b) {
  if (!b) return GFR_EINVAL;
  if (b->dirty) {
    int rc = rebuild(b);
    if (rc != GFR_OK) return rc;
  }
  if (b->active.empty()) return 0;
  if (b->use_graph) {
    GFR_HIP(b->ctx, hipGraphLaunch(b->exec, b->st));
    b->launches++;
  } else {
    GFR_HIP(b->ctx, enqueue_step(b));
  }
  int left = 0;
  for (int f : b->active) {
    Frame& fr = b->frames[f];
    fr.passes++;
    if (fr.done()) b->dirty = true;
    else left++;
  }
  return left;
}

int gfr_batch_run(gfr_batch* b) {
  if (!b) return GFR_EINVAL;
  int rc;
  do {
    rc = gfr_batch_step(b);
  } while (rc > 0);
  if (rc < 0) return rc;
  GFR_HIP(b->ctx, hipStreamSynchronize(b->st));
  return GFR_OK;
}

int gfr_batch_cursor(gfr_batch* b, int frame, int* iter, int* subset, long long* passes) {
  if (!b || frame < 0 || frame >= (int)b->frames.size() || !b->frames[frame].used)
    return b ? set_err(b->ctx, GFR_EINVAL, "unknown frame") : GFR_EINVAL;
  const Frame& fr = b->frames[frame];
  if (iter) *iter = fr.done() ? fr.niter + 1 : (int)(fr.passes / fr.nsub) + 1;
  if (subset) *subset = fr.done() ? 0 : (int)(fr.passes % fr.nsub);
  if (passes) *passes = fr.passes;
  return fr.done() ? 1 : 0;
}

int gfr_batch_get(gfr_batch* b, int frame, float* host_x) {
  if (!b || !host_x) return GFR_EINVAL;
  if (frame < 0 || frame >= (int)b->frames.size() || !b->frames[frame].used)
    return set_err(b->ctx, GFR_EINVAL, "unknown frame");
  const size_t n = sizeof(float) * (size_t)b->gd->host.info.nvox;
  GFR_HIP(b->ctx, hipMemcpyAsync(host_x, b->frames[frame].x, n, hipMemcpyDeviceToHost, b->st));
  GFR_HIP(b->ctx, hipStreamSynchronize(b->st));
  return GFR_OK;
}

int gfr_batch_loglik(gfr_batch* b, int frame, double* out) {
  if (!b || !out) return GFR_EINVAL;
  if (frame < 0 || frame >= (int)b->frames.size() || !b->frames[frame].used)
    return set_err(b->ctx, GFR_EINVAL, "unknown frame");
  if (b->frames[frame].passes == 0)
    return set_err(b->ctx, GFR_ESTATE, "no log-likelihood before the first pass");
  GFR_HIP(b->ctx, hipMemcpyAsync(out, b->d_llh + frame, sizeof(double), hipMemcpyDeviceToHost,
                                 b->st));
  GFR_HIP(b->ctx, hipStreamSynchronize(b->st));
  return GFR_OK;
}

int gfr_batch_remove(gfr_batch* b, int frame) {
  if (!b) return GFR_EINVAL;
  if (frame < 0 || frame >= (int)b->frames.size() || !b->frames[frame].used)
    return set_err(b->ctx, GFR_EINVAL, "unknown frame");
  GFR_HIP(b->ctx, hipStreamSynchronize(b->st));
  Frame& fr = b->frames[frame];
  (void)hipFree(fr.x);
  (void)hipFree(fr.y);
  if (fr.bg) (void)hipFree(fr.bg);
  {
    std::lock_guard<std::mutex> lk(b->ctx->mu);
    b->ctx->in_use -= fr.bytes;
  }
  const bool was_active = !fr.done();
  fr = Frame();
  if (was_active) b->dirty = true;
  return GFR_OK;
}

int gfr_batch_use_graph(gfr_batch* b, int enable) {
  if (!b) return GFR_EINVAL;
  b->use_graph = enable != 0;
  b->dirty = true;
  return GFR_OK;
}

int gfr_batch_stats(gfr_batch* b, long long* launches, long long* builds) {
  if (!b) return GFR_EINVAL;
  if (launches) *launches = b->launches;
  if (builds) *builds = b->builds;
  return GFR_OK;
}

void* gfr_batch_stream(gfr_batch* b) { return b ? (void*)b->st : nullptr; }

int gfr_batch_destroy(gfr_batch* b) {
  if (!b) return GFR_EINVAL;
  if (b->st) (void)hipStreamSynchronize(b->st);
  drop_graph(b);
  size_t frame_bytes = 0;
  for (Frame& fr : b->frames) {
    if (!fr.used) continue;
    (void)hipFree(fr.x);
    (void)hipFree(fr.y);
    if (fr.bg) (void)hipFree(fr.bg);
    frame_bytes += fr.bytes;
  }
  for (Lane& L : b->lanes) {
    if (L.st) {
      (void)hipStreamSynchronize(L.st);
      (void)hipStreamDestroy(L.st);
    }
    if (L.join) (void)hipEventDestroy(L.join);
  }
  for (void* q : b->owned) (void)hipFree(q);
  if (b->fork) (void)hipEventDestroy(b->fork);
  if (b->st) (void)hipStreamDestroy(b->st);
  {
    std::lock_guard<std::mutex> lk(b->ctx->mu);
    b->ctx->in_use -= b->bytes + frame_bytes;
  }
  delete b;
  return GFR_OK;
}

int gfr_poisson(gfr_ctx* ctx, const float* d_mean, double scale, const float* d_mean2,
                double scale2, unsigned long long seed, float* d_out, long long n, void* stream) {
  if (!ctx) return GFR_EINVAL;
  if ((!d_mean || !d_out) && n > 0) return set_err(ctx, GFR_EINVAL, "gfr_poisson: NULL pointer");
  if (n < 0 || !(scale >= 0.0) || !(scale2 >= 0.0))
    return set_err(ctx, GFR_EINVAL, "gfr_poisson: bad size or scale");
  GFR_HIP(ctx, launch_poisson(d_mean, scale, d_mean2, scale2, (uint64_t)seed, d_out, n,
                              stream_of(ctx, stream)));
  return GFR_OK;
}

}  // extern "C"
