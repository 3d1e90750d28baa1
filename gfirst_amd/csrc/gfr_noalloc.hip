// gfr_noalloc.hip — backprojection for callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc). gfr_back_project packs the subset's {ratio, calib} rows
// into scratch the context allocates on demand; gfr_back_project_noalloc
// packs them into a fixed file-scope device array instead and runs the same
// kernels, so both give the same bits.
//
// Each device that loads this library holds its own copy of the array until
// the process ends. Its address on a device is taken once, at context
// creation, with that device current (noalloc_init); host code uses that
// address, never the array's name, which in host code is only a stand-in.
// Calls share the device's array: two calls in flight at once on one device
// (different streams) race. Every entry the backprojector reads is rewritten
// by the pack of the same call, so nothing needs clearing between calls.
#include <mutex>

#include "gfr_device.h"

namespace gfr {

namespace {

constexpr size_t kScratchBytes = size_t(64) << 20;  // 64 MiB
static __device__ __align__(256) float2 scratch[kScratchBytes / sizeof(float2)];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};

}  // namespace

// With `device` current: records the address of this device's scratch array.
hipError_t noalloc_init(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_scratch[device]) return hipSuccess;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_scratch[device] = p;
  return hipSuccess;
}

}  // namespace gfr

using namespace gfr;

extern "C" int gfr_back_project_noalloc(gfr_ctx* ctx, const gfr_geometry* g, const float* d_corr,
                                        const uint8_t* d_mask, float* d_numer, float* d_denom,
                                        int subset, int nsub, void* stream) {
  if (!ctx) return GFR_EINVAL;
  if (!d_corr || !d_mask || !d_numer || !d_denom)
    return set_err(ctx, GFR_EINVAL, "gfr_back_project_noalloc: a device pointer is NULL");
  int cur = -1;
  if (hipGetDevice(&cur) != hipSuccess || cur != ctx->device)
    return set_err(ctx, GFR_ESTATE, "gfr_back_project_noalloc: the context's device is not current");
  void* p = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    p = (cur >= 0 && cur < kMaxDevices) ? g_scratch[cur] : nullptr;
  }
  if (!p) return set_err(ctx, GFR_ESTATE, "gfr_back_project_noalloc: scratch not set up");
  GeomDev* gd = geom_lookup(ctx, g);
  if (!gd)
    return set_err(ctx, GFR_ESTATE,
                   "gfr_back_project_noalloc: the context holds no tables for this geometry "
                   "(any earlier call with the geometry on this context creates them)");
  if (nsub < 1 || nsub > gd->host.g.nang || subset < 0 || subset >= nsub)
    return set_err(ctx, GFR_EINVAL, "subset out of range (need 0 <= subset < nsub <= nang)");
  const size_t bytes = packed_bytes(gd->dev, subset, nsub);
  if (bytes > kScratchBytes)
    return set_err(ctx, GFR_EINVAL,
                   "gfr_back_project_noalloc: the subset's packed rows need " +
                       std::to_string(bytes) + " bytes, more than the 64 MiB scratch "
                       "(use more subsets, or gfr_back_project)");
  const hipStream_t st = stream_of(ctx, stream);
  const Tune t = tune_lookup(ctx, gd, nsub);
  hipError_t e = launch_pack(gd->dev, d_corr, d_mask, (float2*)p, subset, nsub, st);
  if (e == hipSuccess)
    e = launch_back_packed(gd->dev, (const float2*)p, d_numer, d_denom, subset, nsub, nullptr, t,
                           st);
  if (e != hipSuccess) return hip_err(ctx, e, "gfr_back_project_noalloc");
  return GFR_OK;
}
