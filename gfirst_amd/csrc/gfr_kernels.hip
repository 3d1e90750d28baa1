// gfr_kernels.hip — gfx950 kernels of the OSEM hot path.
//
// Rounding model (shared with the CPU oracle so results compare bit for bit):
// the file is compiled with -ffp-contract=off; the only fused multiply-adds
// are explicit fmaf() calls. Interpolation weights are exact FP32:
// lerp(a,b,t) = fma(t, b-a, a), x first, then y, then z.
#include <math.h>

#include "gfr_device.h"

namespace gfr {

namespace {

// Ratio alone (GFIRST.cu:415-430) over the subset's bins; the background is
// in storage format S.
template <int S>
__global__ void k_ratio(DevGeom G, const float* __restrict__ proj, const float* __restrict__ y,
                        const void* __restrict__ bg, float yscale, float* __restrict__ corr,
                        float2* __restrict__ corr2, const uint8_t* __restrict__ mask, int subset,
                        int nsub) {
  const int r = blockIdx.x * kWave + threadIdx.x;
  const int s = blockIdx.y * 4 + threadIdx.y;
  const int it = subset + (int)blockIdx.z * nsub;
  if (r >= G.nrad || s >= G.nsinogs) return;
  const size_t bin = ((size_t)s * G.nang + it) * G.nrad + r;
  const float p = proj[bin];
  const float yv = yscale * y[bin];
  float bv = bg ? yscale * load_store<S>(bg, bin) : 0.f;
  if (bv < 0.f) bv = 0.f;
  const float den = p + bv;
  const float c = (den >= 0.01f) ? yv / den : -1000.f;
  if (corr) corr[bin] = c;
  if (corr2) store_packed(G, corr2, blockIdx.z, s, r, make_float2(c, mask[bin] ? 1.f : -1000.f));
}

// ---------------------------------------------------------------------------
// XY blur (convolutionXY_GPU, GFIRST_GPU_Kernels.cu:332-358). One block per
// (plane, 4 rows); the plane's rows y-4..y+4 go through LDS once. SRC reads
// the input value (plain load, or the update ratio NUMER/DENOM).
// ---------------------------------------------------------------------------
struct SrcPlain {
  const float* p;
  __device__ __forceinline__ float operator()(size_t i) const { return p[i]; }
};
struct SrcRatio {  // divVxVGPU (GFIRST_GPU_Kernels.cu:173-182)
  const float* num;
  const float* den;
  __device__ __forceinline__ float operator()(size_t i) const {
    const float b = den[i];
    return (b < 0.001f) ? 1.0f : num[i] / b;
  }
};

// One block per (plane, kBlurRows output rows): the input rows y-4..y+kBlurRows+3
// are loaded once into LDS (zero outside the plane), y-filtered in LDS, then
// x-filtered; the SRC functor is evaluated once per loaded element.
constexpr int kBlurRows = 16;

template <class Src>
__global__ __launch_bounds__(256) void k_blur_xy(int res, int nzs, Taps9 t, Src src,
                                                 float* __restrict__ out) {
  extern __shared__ float lds[];
  float* in = lds;                           // (kBlurRows + 8) x res
  float* yc = lds + (kBlurRows + 8) * res;   // kBlurRows x res
  const int z = blockIdx.y;
  const int y0 = blockIdx.x * kBlurRows;
  const int nrows = min(kBlurRows, res - y0);
  const size_t pb = (size_t)z * res * res;
  const int tid = threadIdx.x;
  const int nin = (nrows + 8) * res;
  for (int i = tid; i < nin; i += 256) {
    const int row = i / res, col = i - row * res;
    const int y = y0 + row - 4;
    in[i] = (y >= 0 && y < res) ? src(pb + (size_t)y * res + col) : 0.f;
  }
  __syncthreads();
  const int nout = nrows * res;
  for (int i = tid; i < nout; i += 256) {
    const int row = i / res, col = i - row * res;
    float acc = in[row * res + col] * t.g[0];
#pragma unroll
    for (int j = 1; j < 9; j++) acc = fmaf(in[(row + j) * res + col], t.g[j], acc);
    yc[i] = acc;
  }
  __syncthreads();
  for (int i = tid; i < nout; i += 256) {
    const int row = i / res, x = i - row * res;
    float o;
    if (x > 3 && x < res - 4) {
      const float* yr = yc + row * res + x - 4;
      o = yr[0] * t.g[0];
#pragma unroll
      for (int k = 1; k < 9; k++) o = fmaf(yr[k], t.g[k], o);
    } else {
      o = in[(row + 4) * res + x];  // raw input at the x borders
    }
    out[pb + (size_t)(y0 + row) * res + x] = o;
  }
}

// Gz with the reference's index shift, multiply, slice copies
// (GFIRST_CPU_Kernels.cpp:174-186; GFIRST.cu:491-502). One thread per voxel
// of slices 1..NZS-2; slices 0 and NZS-1 receive copies of slices 1 and
// NZS-2, so each voxel is read and written by one thread only (in place).
__global__ __launch_bounds__(256) void k_gz_mul(int res, int nzs, Taps5 t,
                                                const float* __restrict__ f1,
                                                float* __restrict__ x) {
  const int ix = blockIdx.x * kWave + threadIdx.x;
  const int iy = blockIdx.y * 4 + threadIdx.y;
  const int z = blockIdx.z + 1;
  if (ix >= res || iy >= res || z > nzs - 2) return;
  const long long r2 = (long long)res * res;
  const long long nv = r2 * nzs;
  const long long n = z * r2 + (long long)iy * res + ix;
  const int K = (int)((n - 1) / r2);
  auto rd = [&](long long j) -> float { return (j >= 0 && j < nv) ? f1[j] : 0.f; };
  float v;
  if (K == 0) {
    v = 0.5f * rd(n) + 0.5f * rd(n + r2);
  } else if (K == nzs - 1) {
    v = 0.5f * rd(n - r2) + 0.5f * rd(n);
  } else if (K == 1 || K == nzs - 2) {
    v = 0.333f * rd(n - r2) + 0.333f * rd(n) + 0.333f * rd(n + r2);
  } else {
    v = t.g[0] * rd(n - 2 * r2) + t.g[1] * rd(n - r2) + t.g[2] * rd(n) + t.g[3] * rd(n + r2) +
        t.g[4] * rd(n + 2 * r2);
  }
  const float nx = x[n] * v;
  x[n] = nx;
  if (z == 1) x[n - r2] = nx;
  if (z == nzs - 2) x[n + r2] = nx;
}

// ---------------------------------------------------------------------------
// Joint bilateral filter (jointbilateralfilter3D_K, GFIRST_GPU_Kernels.cu:
// 508-539). Spatial weights come from a host table of exp(-d2/(2 sigma)^2)
// (d2 is an integer); intensity weights exp() in double, stored as float.
// ---------------------------------------------------------------------------
struct GTab { float g[193]; };  // d2 = 0..3*8*8

__global__ __launch_bounds__(256) void k_jbf(int res, int nzs, const float* __restrict__ in,
                                             const float* __restrict__ guide,
                                             const float* __restrict__ d_max, float max_ref,
                                             int r, double kI, GTab gt,
                                             float* __restrict__ out) {
  const int ix = blockIdx.x * kWave + threadIdx.x;
  const int iy = blockIdx.y * 4 + threadIdx.y;
  const int iz = blockIdx.z;
  if (ix >= res || iy >= res) return;
  const size_t iV = ((size_t)iz * res + iy) * res + ix;
  if (ix < r || iy < r || iz < r) {
    out[iV] = 0.f;
    return;
  }
  const float mref = d_max ? __uint_as_float(*(const unsigned int*)d_max) : max_ref;
  const float rv = guide[iV];
  float norm = 0.f, sum = 0.f;
  const int z1 = min(iz + r, nzs - 1), y1 = min(iy + r, res - 1), x1 = min(ix + r, res - 1);
  for (int zz = iz - r; zz <= z1; zz++) {
    for (int yy = iy - r; yy <= y1; yy++) {
      const size_t rowb = ((size_t)zz * res + yy) * res;
      for (int xx = ix - r; xx <= x1; xx++) {
        const size_t iW = rowb + xx;
        const float tq = (guide[iW] - rv) / mref;
        const float di2 = tq * tq;
        const int d2 = (xx - ix) * (xx - ix) + (yy - iy) * (yy - iy) + (zz - iz) * (zz - iz);
        const float H = (float)exp(-(double)di2 / kI);
        const float hg = H * gt.g[d2];
        norm += hg;
        sum = fmaf(hg, in[iW], sum);
      }
    }
  }
  out[iV] = ((double)norm > 1e-8) ? sum / norm : 0.f;
}

// The reference's radius (4, GFIRST.cu:512) with the neighbour offsets as
// loop counters: offsets and the spatial-weight index are the same for every
// lane (scalar), a row's nine texels are read by buffer loads from one 32-bit
// row offset (texels past the volume's end read 0 and are not summed), and
// each voxel still sums its neighbours in the reference's order (z, then y,
// then x), skipping those past the volume's upper faces.
//
// Each intensity weight H = (float)exp(-(double)di2 / kI) is first computed
// on a fast path whose result is provably the same float:
//  - tq = (g - rv) / mref by a reciprocal and two FMA corrections with
//    y = RN(1/mref) (Markstein's theorem: the second correction of a
//    faithful quotient is the correctly rounded one). It holds while nothing
//    underflows or overflows: for mref in [2^-20, 2^20) and a first quotient
//    q0 of magnitude in [2^-60, 2^90). Outside that range the weight is the
//    same on both paths when kI is in [2^-80, 2^40]: |q0| < 2^-60 gives
//    |x| < 2^-40 and H = 1; |q0| >= 2^90 gives di2 = inf (H = 0) or a NaN,
//    which the range test below sends to the reference expression. With
//    mref or kI outside those bounds every weight takes that expression.
//  - x' = di2 * (-1/kI) instead of the division (|x' - x| <= 3 ulp(x)), and
//    exp(x') = 2^(n/1024) * p(r): n rounded by the 1.5 * 2^52 shift (its low
//    word is n), r = x' - n ln2/1024 in one FMA, a 1024-entry table of
//    2^(j/1024) and a cubic on |r| <= ln2/2048. For x' >= -87 (a normal
//    float result) the relative error is below 2^-44.
//  - the float rounding of E drops its 29 low mantissa bits; if they are
//    more than 2^13 units (2^-40 relative) away from the halfway point, the
//    library's exp of the exact x rounds to the same float. Below x' = -104.5
//    the float is 0 either way.
// A weight near a rounding midpoint (about 3e-5 of them) or with x' in
// [-104.5, -87) (a denormal float) is recomputed by the reference's
// expression; the row's weights are then summed in order.
namespace jbfx {
constexpr int kBits = 10, kTab = 1 << kBits;
constexpr double kCN = 1477.3197218702985;                  // 1024 / ln2
constexpr double kShift = 0x1.8p52;                          // ulp 1
constexpr double kLn2N = 0x1.62e42fefa39efp-1 / kTab;
constexpr unsigned kHalf = 1u << 28, kBand = 1u << 13;
}  // namespace jbfx

__global__ __launch_bounds__(256) void k_jbf_r4(int res, int nzs, const float* __restrict__ in,
                                                const float* __restrict__ guide,
                                                const float* __restrict__ d_max, float max_ref,
                                                double kI, double nki, GTab gt,
                                                float* __restrict__ out) {
  constexpr int R = 4;
  __shared__ double p2[jbfx::kTab];  // 2^(j/1024)
  for (int j = threadIdx.y * kWave + threadIdx.x; j < jbfx::kTab; j += 4 * kWave)
    p2[j] = exp2((double)j * (1.0 / jbfx::kTab));
  __syncthreads();
  const int ix = blockIdx.x * kWave + threadIdx.x;
  const int iy = blockIdx.y * 4 + threadIdx.y;
  const int iz = blockIdx.z;
  if (ix >= res || iy >= res) return;
  const int r2 = res * res;
  const int iV = (iz * res + iy) * res + ix;
  if (ix < R || iy < R || iz < R) {
    out[iV] = 0.f;
    return;
  }
  const unsigned vbytes = (unsigned)(r2 * nzs) * 4u;
  const __amdgpu_buffer_rsrc_t rg = rsrc(guide, vbytes), ri = rsrc(in, vbytes);
  const float mref = d_max ? __uint_as_float(*(const unsigned int*)d_max) : max_ref;
  const float ym = 1.0f / mref;
  const unsigned allslow =
      (mref >= 0x1p-20f && mref < 0x1p20f && kI >= 0x1p-80 && kI <= 0x1p40) ? 0u : 0x1ffu;
  const float rv = guide[iV];
  float norm = 0.f, sum = 0.f;
  for (int dz = -R; dz <= R; dz++) {
    for (int dy = -R; dy <= R; dy++) {
      const bool vzy = iz + dz <= nzs - 1 && iy + dy <= res - 1;
      // bit k: neighbour ix + k - R lies inside the volume
      const unsigned vmask = vzy ? (0x1ffu >> max(0, ix + R - (res - 1))) : 0u;
      unsigned voff = (unsigned)(iV + dz * r2 + dy * res - R) * 4u;
      asm volatile("" : "+v"(voff));  // row start as the base: texel k at immediate 4k
      float gv[2 * R + 1], iv[2 * R + 1], H[2 * R + 1];
#pragma unroll
      for (int k = 0; k <= 2 * R; k++) {
        gv[k] = ld_f1(rg, voff + 4u * k, 0);
        iv[k] = ld_f1(ri, voff + 4u * k, 0);
      }
      unsigned slow = 0;
#pragma unroll
      for (int k = 0; k <= 2 * R; k++) {
        const float a = gv[k] - rv;
        const float q0 = a * ym;
        const float q1 = fmaf(fmaf(-q0, mref, a), ym, q0);
        const float tq = fmaf(fmaf(-q1, mref, a), ym, q1);
        const float di2 = tq * tq;
        const double x = (double)di2 * nki;
        const double sh = fma(x, jbfx::kCN, jbfx::kShift);
        const int n = __double2loint(sh);
        const double r = fma(-(sh - jbfx::kShift), jbfx::kLn2N, x);
        double p = fma(r, 1.0 / 6, 0.5);
        p = fma(p, r, 1.0);
        p = fma(p, r, 1.0);
        const double E = ldexp(p2[n & (jbfx::kTab - 1)] * p, n >> jbfx::kBits);
        const unsigned low = (unsigned)__double2loint(E) & 0x1fffffffu;
        const bool amb = low - (jbfx::kHalf - jbfx::kBand) < 2u * jbfx::kBand;
        const bool under = x < -104.5;
        H[k] = under ? 0.f : (float)E;
        const bool bad = !under & (!(x >= -87.0) | amb);
        slow |= bad ? 1u << k : 0u;
      }
      slow = (slow | allslow) & vmask;
      if (slow) {
        for (int k = 0; k <= 2 * R; k++) {
          if ((slow >> k) & 1u) {
            const float tq = (gv[k] - rv) / mref;
            const float di2 = tq * tq;
            H[k] = (float)exp(-(double)di2 / kI);
          }
        }
      }
      const int d2yz = dz * dz + dy * dy;
#pragma unroll
      for (int k = 0; k <= 2 * R; k++) {
        if ((vmask >> k) & 1u) {
          const float hg = H[k] * gt.g[d2yz + (k - R) * (k - R)];
          norm += hg;
          sum = fmaf(hg, iv[k], sum);
        }
      }
    }
  }
  out[iV] = ((double)norm > 1e-8) ? sum / norm : 0.f;
}

__global__ void k_blend(long long nv, float coef, const float* __restrict__ filt,
                        float* __restrict__ x) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  x[i] = (float)((1.0 - (double)coef) * (double)x[i] + (double)(coef * filt[i]));
}

// max(0, max x): non-negative floats order like their bit patterns.
__global__ void k_max(long long nv, const float* __restrict__ x, unsigned int* bits) {
  __shared__ unsigned int red[256];
  unsigned int m = 0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nv;
       i += (long long)gridDim.x * blockDim.x) {
    const float v = x[i];
    if (v > 0.f) m = max(m, __float_as_uint(v));
  }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = max(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(bits, red[0]);
}

__global__ void k_fill_cylinder(int res, int nzs, float val, float* __restrict__ img) {
  const int x = blockIdx.x * kWave + threadIdx.x;
  const int y = blockIdx.y * 4 + threadIdx.y;
  const int z = blockIdx.z;
  if (x >= res || y >= res) return;
  const int h = res / 2;
  const long long d = (long long)(x - h) * (x - h) + (long long)(y - h) * (y - h);
  img[((size_t)z * res + y) * res + x] = (d <= (long long)h * h) ? val : 0.f;
}

__global__ void k_zero(uint32_t* __restrict__ p, size_t nwords) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nwords;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0u;
}

__global__ void k_zero_bytes(uint8_t* __restrict__ p, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// Calibration mask (GFIRST.cu:453-458); the reference compares in double.
__global__ void k_calib_mask(long long nd, const float* __restrict__ pb,
                             const float* __restrict__ norm, uint8_t* __restrict__ mask) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nd) return;
  mask[i] = ((double)pb[i] > 0.01 && (double)norm[i] > 0.01) ? 1 : 0;
}

__global__ void k_scale_copy(const float* __restrict__ in, float s, float* __restrict__ out,
                             long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[i] * s;
}

__global__ void k_convert_store(const float* __restrict__ in, void* __restrict__ out,
                                int storage, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float v = in[i];
  if (storage == GFR_STORE_F32) {
    ((float*)out)[i] = v;
  } else if (storage == GFR_STORE_F16) {
    ((_Float16*)out)[i] = (_Float16)v;
  } else {
    // round to nearest even; NaN kept quiet
    uint32_t b = __float_as_uint(v);
    uint16_t h;
    if ((b & 0x7fffffffu) > 0x7f800000u) {
      h = (uint16_t)((b >> 16) | 0x40u);
    } else {
      b += 0x7fffu + ((b >> 16) & 1u);
      h = (uint16_t)(b >> 16);
    }
    ((uint16_t*)out)[i] = h;
  }
}

// Deterministic final sum of per-block FP64 partials (fixed order).
__global__ void k_llh_final(const double* __restrict__ part, int n, double* __restrict__ out) {
  __shared__ double red[256];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) acc += part[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

// ---------------------------------------------------------------------------
// VLOR (GFIRST.cu:229-247, 373-388).
// ---------------------------------------------------------------------------
__global__ void k_vlor_blankt(const float* __restrict__ yT, const float* __restrict__ v0,
                              float res, float* __restrict__ blankt, long long nd) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nd) return;
  const float val = v0[i];
  blankt[i] = ((double)val > 0.01) ? yT[i] / (res * val) : 0.f;
}

// datos_v0[i]++ for both LORs of every valid record (GFIRST.cu:238-239):
// counts stay exact integers in float (below 2^24), so the order of the
// atomic additions does not matter.
__global__ void k_vlor_count(const int* __restrict__ i1, const int* __restrict__ i2,
                             long long nev, float* __restrict__ v0) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nev) return;
  atomicAdd(&v0[i1[e]], 1.f);
  atomicAdd(&v0[i2[e]], 1.f);
}

// Per event: s_inv = 1/(s1+s2) in double, rounded to float (GFIRST.cu:384-385)
// and, once, valor = BLANKT[i1] + BLANKT[i2] (GFIRST.cu:381).
__global__ void k_vlor_event(const int* __restrict__ i1, const int* __restrict__ i2,
                             const float* __restrict__ blankt, const float* __restrict__ proj,
                             float* __restrict__ sinv, float* __restrict__ valor, long long nev,
                             int want_valor) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nev) return;
  const int a = i1[e], b = i2[e];
  if (want_valor) valor[e] = blankt[a] + blankt[b];
  const float s = proj[a] + proj[b];
  sinv[e] = (s > 0.01f) ? (float)(1.0 / (double)s) : 0.f;
}

__global__ void k_vlor_zero_subset(DevGeom G, float* __restrict__ dv, int subset, int nsub) {
  const int r = blockIdx.x * kWave + threadIdx.x;
  const int s = blockIdx.y * 4 + threadIdx.y;
  const int it = subset + (int)blockIdx.z * nsub;
  if (r >= G.nrad || s >= G.nsinogs) return;
  dv[((size_t)s * G.nang + it) * G.nrad + r] = 0.f;
}

// One thread per touched bin: its contributions in record order
// (datos_v[i] += s_i * s_inv * valor, GFIRST.cu:386-387).
__global__ void k_vlor_bins(DevGeom G, const long long* __restrict__ bins,
                            const int* __restrict__ start, const int* __restrict__ ev,
                            long long nbins, const float* __restrict__ proj,
                            const float* __restrict__ sinv, const float* __restrict__ valor,
                            float* __restrict__ dv, int subset, int nsub) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nbins) return;
  const long long b = bins[k];
  const int it = (int)((b / G.nrad) % G.nang);
  if (it % nsub != subset) return;
  const float sb = proj[b];
  float acc = 0.f;
  for (int j = start[k]; j < start[k + 1]; j++) {
    const int e = ev[j];
    acc += sb * sinv[e] * valor[e];
  }
  dv[b] = acc;
}

inline dim3 grid_img(const DevGeom& G, int nz) {
  return dim3((G.res + kWave - 1) / kWave, (G.res + 3) / 4, nz);
}

inline int nav_of(const DevGeom& G, int subset, int nsub) {
  return (G.nang - subset + nsub - 1) / nsub;
}

}  // namespace

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
hipError_t launch_fill_cylinder(const DevGeom& G, float val, float* img, hipStream_t st) {
  hipLaunchKernelGGL(k_fill_cylinder, grid_img(G, G.nzs), dim3(kWave, 4), 0, st, G.res, G.nzs,
                     val, img);
  return hipGetLastError();
}

namespace {
// Advances the subset index of each listed slot: sub = (sub + 1) mod nsub.
__global__ void k_advance(int* __restrict__ sub, const int* __restrict__ nsub,
                          const int* __restrict__ slots, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int s = slots[i];
  const int v = sub[s] + 1;
  sub[s] = v >= nsub[s] ? 0 : v;
}
}  // namespace

namespace {
__global__ void k_any_positive(const float* __restrict__ v, long long n,
                               unsigned int* __restrict__ flag) {
  bool any = false;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    any |= !(v[i] <= 0.f);
  if (any) *flag = 1u;
}

// *flag = 1 if some v[i] is NaN or infinite.
__global__ void k_any_nonfinite(const float* __restrict__ v, long long n,
                                unsigned int* __restrict__ flag) {
  bool bad = false;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    bad |= (__float_as_uint(v[i]) & 0x7f800000u) == 0x7f800000u;
  if (bad) *flag = 1u;
}
}  // namespace

hipError_t launch_any_nonfinite(const float* v, long long n, unsigned int* flag,
                                hipStream_t st) {
  if (n <= 0) return hipSuccess;
  long long blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_any_nonfinite, dim3((unsigned)blocks), dim3(256), 0, st, v, n, flag);
  return hipGetLastError();
}

hipError_t launch_any_positive(const float* v, long long n, unsigned int* flag, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  long long blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_any_positive, dim3((unsigned)blocks), dim3(256), 0, st, v, n, flag);
  return hipGetLastError();
}

hipError_t launch_advance(int* sub, const int* nsub, const int* slots, int n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_advance, dim3((n + 255) / 256), dim3(256), 0, st, sub, nsub, slots, n);
  return hipGetLastError();
}

hipError_t launch_zero(void* p, size_t bytes, hipStream_t st) {
  if (bytes == 0) return hipSuccess;
  size_t nw = bytes / 4;
  if (nw) {
    size_t blocks = (nw + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_zero, dim3((unsigned)blocks), dim3(256), 0, st, (uint32_t*)p, nw);
  }
  size_t rem = bytes - nw * 4;
  if (rem) {
    hipLaunchKernelGGL(k_zero_bytes, dim3(1), dim3(64), 0, st, (uint8_t*)p + nw * 4, rem);
  }
  return hipGetLastError();
}

template <class Src>
hipError_t blur_launch(const DevGeom& G, const Taps9& t, Src src, float* out, hipStream_t st) {
  const size_t lds = sizeof(float) * (size_t)(2 * kBlurRows + 8) * G.res;
  if (lds > 160 * 1024) return hipErrorInvalidValue;
  dim3 grid((G.res + kBlurRows - 1) / kBlurRows, G.nzs);
  hipLaunchKernelGGL(k_blur_xy<Src>, grid, dim3(256), lds, st, G.res, G.nzs, t, src, out);
  return hipGetLastError();
}

hipError_t launch_blur_xy(const DevGeom& G, const Taps9& t, const float* in, float* out,
                          hipStream_t st) {
  return blur_launch(G, t, SrcPlain{in}, out, st);
}

hipError_t launch_ratio(const DevGeom& G, const float* proj, const float* y, const void* bg,
                        int storage, float yscale, float* corr, float2* corr2,
                        const uint8_t* mask, int subset, int nsub, hipStream_t st) {
  const int nav = nav_of(G, subset, nsub);
  dim3 grid((G.nrad + kWave - 1) / kWave, (G.nsinogs + 3) / 4, nav);
  if (storage == GFR_STORE_F16)
    hipLaunchKernelGGL(k_ratio<GFR_STORE_F16>, grid, dim3(kWave, 4), 0, st, G, proj, y, bg,
                       yscale, corr, corr2, mask, subset, nsub);
  else if (storage == GFR_STORE_BF16)
    hipLaunchKernelGGL(k_ratio<GFR_STORE_BF16>, grid, dim3(kWave, 4), 0, st, G, proj, y, bg,
                       yscale, corr, corr2, mask, subset, nsub);
  else
    hipLaunchKernelGGL(k_ratio<GFR_STORE_F32>, grid, dim3(kWave, 4), 0, st, G, proj, y, bg,
                       yscale, corr, corr2, mask, subset, nsub);
  return hipGetLastError();
}

hipError_t launch_calib_mask(const DevGeom& G, const float* pb, const float* norm,
                             uint8_t* mask, hipStream_t st) {
  const long long nd = G.ndata;
  hipLaunchKernelGGL(k_calib_mask, dim3((unsigned)((nd + 255) / 256)), dim3(256), 0, st, nd, pb,
                     norm, mask);
  return hipGetLastError();
}

hipError_t launch_update(const DevGeom& G, const Taps9& t9, const Taps5& t5, const float* numer,
                         const float* denom, float* x, float* work, hipStream_t st) {
  hipError_t e = blur_launch(G, t9, SrcRatio{numer, denom}, work, st);
  if (e != hipSuccess) return e;
  if (G.nzs >= 3)
    hipLaunchKernelGGL(k_gz_mul, grid_img(G, G.nzs - 2), dim3(kWave, 4), 0, st, G.res, G.nzs, t5,
                       work, x);
  return hipGetLastError();
}

hipError_t launch_jbf(const DevGeom& G, const float* in, const float* guide,
                      const float* d_max_ref, float max_ref, int r, float s3d, float sI,
                      float* out, hipStream_t st) {
  GTab gt;
  const double kS = (2.0 * (double)s3d) * (2.0 * (double)s3d);
  for (int d2 = 0; d2 < 193; d2++) gt.g[d2] = (float)exp(-(double)(float)d2 / kS);
  const double kI = (2.0 * (double)sI) * (2.0 * (double)sI);
  // the radius-4 kernel addresses the volume with 32-bit byte offsets
  if (r == 4 && G.nvox < (1ll << 29))
    hipLaunchKernelGGL(k_jbf_r4, grid_img(G, G.nzs), dim3(kWave, 4), 0, st, G.res, G.nzs, in,
                       guide, d_max_ref, max_ref, kI, -1.0 / kI, gt, out);
  else
    hipLaunchKernelGGL(k_jbf, grid_img(G, G.nzs), dim3(kWave, 4), 0, st, G.res, G.nzs, in,
                       guide, d_max_ref, max_ref, r, kI, gt, out);
  return hipGetLastError();
}

hipError_t launch_blend(const DevGeom& G, float coef, const float* filt, float* x,
                        hipStream_t st) {
  const long long nv = G.nvox;
  hipLaunchKernelGGL(k_blend, dim3((unsigned)((nv + 255) / 256)), dim3(256), 0, st, nv, coef,
                     filt, x);
  return hipGetLastError();
}

hipError_t launch_max(const DevGeom& G, const float* x, unsigned int* d_bits, hipStream_t st) {
  hipError_t e = launch_zero(d_bits, 4, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_max, dim3(1024), dim3(256), 0, st, G.nvox, x, d_bits);
  return hipGetLastError();
}

hipError_t launch_scale_copy(const float* in, float s, float* out, long long n, hipStream_t st) {
  hipLaunchKernelGGL(k_scale_copy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, in, s, out,
                     n);
  return hipGetLastError();
}

hipError_t launch_convert_store(const float* in, void* out, int storage, long long n,
                                hipStream_t st) {
  hipLaunchKernelGGL(k_convert_store, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, in,
                     out, storage, n);
  return hipGetLastError();
}

hipError_t launch_llh_final(const double* part, int n, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_llh_final, dim3(1), dim3(256), 0, st, part, n, out);
  return hipGetLastError();
}

hipError_t launch_vlor_blankt(const float* yT, const float* v0, float res, float* blankt,
                              long long nd, hipStream_t st) {
  hipLaunchKernelGGL(k_vlor_blankt, dim3((unsigned)((nd + 255) / 256)), dim3(256), 0, st, yT, v0,
                     res, blankt, nd);
  return hipGetLastError();
}

hipError_t launch_vlor_count(const int* i1, const int* i2, long long nev, float* v0,
                             hipStream_t st) {
  if (nev == 0) return hipSuccess;
  hipLaunchKernelGGL(k_vlor_count, dim3((unsigned)((nev + 255) / 256)), dim3(256), 0, st, i1, i2,
                     nev, v0);
  return hipGetLastError();
}

hipError_t launch_vlor_event(const int* i1, const int* i2, const float* blankt,
                             const float* proj, float* sinv, float* valor, long long nev,
                             int want_valor, hipStream_t st) {
  if (nev == 0) return hipSuccess;
  hipLaunchKernelGGL(k_vlor_event, dim3((unsigned)((nev + 255) / 256)), dim3(256), 0, st, i1, i2,
                     blankt, proj, sinv, valor, nev, want_valor);
  return hipGetLastError();
}

hipError_t launch_vlor_zero_subset(const DevGeom& G, float* dv, int subset, int nsub,
                                   hipStream_t st) {
  const int nav = nav_of(G, subset, nsub);
  dim3 grid((G.nrad + kWave - 1) / kWave, (G.nsinogs + 3) / 4, nav);
  hipLaunchKernelGGL(k_vlor_zero_subset, grid, dim3(kWave, 4), 0, st, G, dv, subset, nsub);
  return hipGetLastError();
}

hipError_t launch_vlor_bins(const DevGeom& G, const long long* bins, const int* start,
                            const int* ev, long long nbins, const float* proj,
                            const float* sinv, const float* valor, float* dv, int subset,
                            int nsub, hipStream_t st) {
  if (nbins == 0) return hipSuccess;
  hipLaunchKernelGGL(k_vlor_bins, dim3((unsigned)((nbins + 255) / 256)), dim3(256), 0, st, G,
                     bins, start, ev, nbins, proj, sinv, valor, dv, subset, nsub);
  return hipGetLastError();
}

}  // namespace gfr
