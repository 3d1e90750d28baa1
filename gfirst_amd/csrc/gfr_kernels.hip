// gfr_kernels.hip — gfx950 kernels of the OSEM hot path.
//
// Rounding model (shared with the CPU oracle so results compare bit for bit):
// the file is compiled with -ffp-contract=off; the only fused multiply-adds
// are explicit fmaf() calls. Interpolation weights are exact FP32:
// lerp(a,b,t) = fma(t, b-a, a), x first, then y, then z.
#include <hip/hip_runtime.h>
#include <math.h>

#include "gfr_internal.h"

namespace gfr {

namespace {

constexpr int kWave = 64;

__device__ __forceinline__ float lerp1(float a, float b, float t) { return fmaf(t, b - a, a); }

// Unnormalized coordinate -> (i0, i1, alpha) with clamp-to-edge addressing:
// texel i covers [i, i+1), so the filter blends floor(c-0.5) and the next one
// (GFIRST_GPU_Kernels.cu:260-263, 270-279 set linear filter + clamp).
__device__ __forceinline__ void axis(float c, int n, int& i0, int& i1, float& a) {
  float t = c - 0.5f;
  float f = floorf(t);
  int i = (int)f;
  a = t - f;
  i0 = min(max(i, 0), n - 1);
  i1 = min(max(i + 1, 0), n - 1);
}

// Trilinear sample of the image at an unnormalized (x, y, z): replaces
// tex3D(imgEstTex, x, y, z) (GFIRST_GPU_Kernels.cu:97).
__device__ __forceinline__ float tex3(const float* __restrict__ v, int W, int H, int D, float x,
                                      float y, float z) {
  int x0, x1, y0, y1, z0, z1;
  float ax, ay, az;
  axis(x, W, x0, x1, ax);
  axis(y, H, y0, y1, ay);
  axis(z, D, z0, z1, az);
  const float* p0 = v + (size_t)z0 * W * H;
  const float* p1 = v + (size_t)z1 * W * H;
  float b0 = lerp1(lerp1(p0[y0 * W + x0], p0[y0 * W + x1], ax),
                   lerp1(p0[y1 * W + x0], p0[y1 * W + x1], ax), ay);
  float b1 = lerp1(lerp1(p1[y0 * W + x0], p1[y0 * W + x1], ax),
                   lerp1(p1[y1 * W + x0], p1[y1 * W + x1], ax), ay);
  return lerp1(b0, b1, az);
}

template <int S>
__device__ __forceinline__ float load_store(const void* p, size_t i) {
  if constexpr (S == GFR_STORE_F32) {
    return ((const float*)p)[i];
  } else if constexpr (S == GFR_STORE_F16) {
    return (float)((const _Float16*)p)[i];
  } else {
    uint32_t b = (uint32_t)((const uint16_t*)p)[i] << 16;
    return __uint_as_float(b);
  }
}

// ---------------------------------------------------------------------------
// Forward projection (+ optional ratio and FP64 log-likelihood epilogue).
// One thread per LOR (view = (segment, angle), plane, radial bin): the LOR
// sum of RES+1 trilinear samples of FORWPROJ_K (GFIRST_GPU_Kernels.cu:73-98),
// then the ratio of GFIRST.cu:415-430 for the same bin.
// Block = 64 radial bins x 4 sinograms; grid.z = angles of the subset.
// ---------------------------------------------------------------------------
template <int S, bool RATIO, bool LLH>
__global__ __launch_bounds__(256) void k_forward(DevGeom G, const float* __restrict__ img,
                                                 float* __restrict__ proj,
                                                 const void* __restrict__ y,
                                                 const void* __restrict__ bg, float yscale,
                                                 float* __restrict__ corr,
                                                 double* __restrict__ llh_part, int subset,
                                                 int nsub) {
  const int r = blockIdx.x * kWave + threadIdx.x;
  const int s = blockIdx.y * 4 + threadIdx.y;
  const int it = subset + (int)blockIdx.z * nsub;
  double ll = 0.0;
  if (r < G.nrad && s < G.nsinogs) {
    const int2 sp = G.sino[s];
    const SegEntry se = G.seg[sp.x];
    const float st = G.sin_t[it], ct = G.cos_t[it];
    const float dx = st, dy = -ct, dz = se.dzf;
    const float zc = (float)sp.y + 0.5f;
    const float x0 = G.fxf * ((float)r - G.RADM);
    const float xc = fmaf(x0, ct, G.RESM) + 0.5f;
    const float yc = fmaf(x0, st, G.RESM) + 0.5f;
    float acc = 0.f;
    for (int l = G.l0; l <= G.l1; l++) {
      const float fl = (float)l;
      acc += tex3(img, G.res, G.res, G.nzs, fmaf(fl, dx, xc), fmaf(fl, dy, yc), fmaf(fl, dz, zc));
    }
    const size_t bin = ((size_t)s * G.nang + it) * G.nrad + r;
    if (proj) proj[bin] = acc;
    if constexpr (RATIO) {
      const float yv = yscale * load_store<S>(y, bin);
      float bv = bg ? yscale * load_store<S>(bg, bin) : 0.f;
      if (bv < 0.f) bv = 0.f;
      const float den = acc + bv;
      corr[bin] = (den >= 0.01f) ? yv / den : -1000.f;
      if constexpr (LLH) {
        if (den > 0.f) ll = (double)yv * log((double)den) - (double)den;
      }
    }
  }
  if constexpr (LLH) {
    __shared__ double red[256];
    const int t = threadIdx.y * kWave + threadIdx.x;
    red[t] = ll;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if (t < w) red[t] += red[t + w];
      __syncthreads();
    }
    if (t == 0) {
      const size_t b = ((size_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
      llh_part[b] = red[0];
    }
  }
}

// Ratio alone (GFIRST.cu:415-430) over the subset's bins; the background is
// in storage format S.
template <int S>
__global__ void k_ratio(DevGeom G, const float* __restrict__ proj, const float* __restrict__ y,
                        const void* __restrict__ bg, float yscale, float* __restrict__ corr,
                        int subset, int nsub) {
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
  corr[bin] = (den >= 0.01f) ? yv / den : -1000.f;
}

// ---------------------------------------------------------------------------
// Backprojection: one thread per voxel, views in the reference's order
// (segment-major, then angle; GFIRST.cu:439-465). Per view: the rotation and
// bilinear fetch of BACKWPROJ_K (GFIRST_GPU_Kernels.cu:136-160) from the
// sinogram rows of the view; planes outside [zmin, NZS-zmin) read -1000 as
// the host's CORR_PART/CALIB_PART fill does (GFIRST.cu:447).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_back(DevGeom G, const float* __restrict__ corr,
                                              const uint8_t* __restrict__ mask,
                                              float* __restrict__ numer,
                                              float* __restrict__ denom, int subset, int nsub,
                                              int nav) {
  const int ix = blockIdx.x * kWave + threadIdx.x;
  const int iy = blockIdx.y * 4 + threadIdx.y;
  const int iz = blockIdx.z;
  if (ix >= G.res || iy >= G.res) return;
  const float xv = (float)ix - G.RESM, yv = (float)iy - G.RESM;
  const size_t rowstride = (size_t)G.nang * G.nrad;
  float n = 0.f, d = 0.f;
  for (int si = 0; si < G.nseg; si++) {
    const SegEntry se = G.seg[si];
    const int zm = se.zmin;
    const int zhi = G.nzs - zm;
    for (int a = 0; a < nav; a++) {
      const int it = subset + a * nsub;
      const float st = G.sin_t[it], ct = G.cos_t[it];
      const float xr = fmaf(yv, st, xv * ct);
      const float yr = fmaf(yv, ct, -(xv * st));
      const float u = fmaf(xr, G.fxb, G.RADM) + 0.5f;
      const float w = fmaf(yr, se.fzb, (float)iz) + 0.5f;
      float av = 0.f, bv = 0.f;
      if (w >= (float)zm && w < (float)zhi) {
        int r0, r1, p0, p1;
        float au, aw;
        axis(u, G.nrad, r0, r1, au);
        axis(w, G.nzs, p0, p1, aw);
        const size_t rb = ((size_t)se.nza * G.nang + it) * G.nrad;
        const bool v0 = (p0 >= zm && p0 < zhi), v1 = (p1 >= zm && p1 < zhi);
        const size_t o0 = rb + (size_t)(p0 - zm) * rowstride;
        const size_t o1 = rb + (size_t)(p1 - zm) * rowstride;
        const float c00 = v0 ? corr[o0 + r0] : -1000.f, c01 = v0 ? corr[o0 + r1] : -1000.f;
        const float c10 = v1 ? corr[o1 + r0] : -1000.f, c11 = v1 ? corr[o1 + r1] : -1000.f;
        const float k00 = (v0 && mask[o0 + r0]) ? 1.f : -1000.f;
        const float k01 = (v0 && mask[o0 + r1]) ? 1.f : -1000.f;
        const float k10 = (v1 && mask[o1 + r0]) ? 1.f : -1000.f;
        const float k11 = (v1 && mask[o1 + r1]) ? 1.f : -1000.f;
        av = lerp1(lerp1(c00, c01, au), lerp1(c10, c11, au), aw);
        bv = lerp1(lerp1(k00, k01, au), lerp1(k10, k11, au), aw);
      }
      if (bv < 0.f || av < 0.f) {
        av = 0.f;
        bv = 0.f;
      }
      n += av;
      d += bv;
    }
  }
  const size_t iV = ((size_t)iz * G.res + iy) * G.res + ix;
  numer[iV] = n;
  denom[iV] = d;
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

template <class Src>
__global__ __launch_bounds__(256) void k_blur_xy(int res, int nzs, Taps9 t, Src src,
                                                 float* __restrict__ out) {
  const int x = blockIdx.x * kWave + threadIdx.x;
  const int y = blockIdx.y * 4 + threadIdx.y;
  const int z = blockIdx.z;
  if (x >= res || y >= res) return;
  const size_t pb = (size_t)z * res * res;
  float o;
  if (x > 3 && x < res - 4) {
    float acc = 0.f;
#pragma unroll
    for (int k = 0; k < 9; k++) {
      const int xx = x + k - 4;
      float yc = 0.f;
#pragma unroll
      for (int j = 0; j < 9; j++) {
        const int yy = y + j - 4;
        const float v = (yy >= 0 && yy < res) ? src(pb + (size_t)yy * res + xx) : 0.f;
        yc = (j == 0) ? v * t.g[0] : fmaf(v, t.g[j], yc);
      }
      acc = (k == 0) ? yc * t.g[0] : fmaf(yc, t.g[k], acc);
    }
    o = acc;
  } else {
    o = src(pb + (size_t)y * res + x);
  }
  out[pb + (size_t)y * res + x] = o;
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

hipError_t launch_blur_xy(const DevGeom& G, const Taps9& t, const float* in, float* out,
                          hipStream_t st) {
  hipLaunchKernelGGL(k_blur_xy<SrcPlain>, grid_img(G, G.nzs), dim3(kWave, 4), 0, st, G.res,
                     G.nzs, t, SrcPlain{in}, out);
  return hipGetLastError();
}

hipError_t launch_forward(const DevGeom& G, const float* img, float* proj, const void* y,
                          const void* bg, int storage, float yscale, float* corr,
                          double* llh_part, int subset, int nsub, hipStream_t st) {
  const int nav = nav_of(G, subset, nsub);
  dim3 grid((G.nrad + kWave - 1) / kWave, (G.nsinogs + 3) / 4, nav);
  dim3 blk(kWave, 4);
  if (!corr) {
    hipLaunchKernelGGL((k_forward<GFR_STORE_F32, false, false>), grid, blk, 0, st, G, img, proj,
                       nullptr, nullptr, 1.f, nullptr, nullptr, subset, nsub);
    return hipGetLastError();
  }
#define GFR_FWD(S)                                                                              \
  if (llh_part)                                                                                 \
    hipLaunchKernelGGL((k_forward<S, true, true>), grid, blk, 0, st, G, img, proj, y, bg,      \
                       yscale, corr, llh_part, subset, nsub);                                   \
  else                                                                                          \
    hipLaunchKernelGGL((k_forward<S, true, false>), grid, blk, 0, st, G, img, proj, y, bg,     \
                       yscale, corr, nullptr, subset, nsub);
  if (storage == GFR_STORE_F16) {
    GFR_FWD(GFR_STORE_F16)
  } else if (storage == GFR_STORE_BF16) {
    GFR_FWD(GFR_STORE_BF16)
  } else {
    GFR_FWD(GFR_STORE_F32)
  }
#undef GFR_FWD
  return hipGetLastError();
}

hipError_t launch_ratio(const DevGeom& G, const float* proj, const float* y, const void* bg,
                        int storage, float yscale, float* corr, int subset, int nsub,
                        hipStream_t st) {
  const int nav = nav_of(G, subset, nsub);
  dim3 grid((G.nrad + kWave - 1) / kWave, (G.nsinogs + 3) / 4, nav);
  if (storage == GFR_STORE_F16)
    hipLaunchKernelGGL(k_ratio<GFR_STORE_F16>, grid, dim3(kWave, 4), 0, st, G, proj, y, bg,
                       yscale, corr, subset, nsub);
  else if (storage == GFR_STORE_BF16)
    hipLaunchKernelGGL(k_ratio<GFR_STORE_BF16>, grid, dim3(kWave, 4), 0, st, G, proj, y, bg,
                       yscale, corr, subset, nsub);
  else
    hipLaunchKernelGGL(k_ratio<GFR_STORE_F32>, grid, dim3(kWave, 4), 0, st, G, proj, y, bg,
                       yscale, corr, subset, nsub);
  return hipGetLastError();
}

hipError_t launch_calib_mask(const DevGeom& G, const float* pb, const float* norm,
                             uint8_t* mask, hipStream_t st) {
  const long long nd = G.ndata;
  hipLaunchKernelGGL(k_calib_mask, dim3((unsigned)((nd + 255) / 256)), dim3(256), 0, st, nd, pb,
                     norm, mask);
  return hipGetLastError();
}

hipError_t launch_back(const DevGeom& G, const float* corr, const uint8_t* mask, float* numer,
                       float* denom, int subset, int nsub, hipStream_t st) {
  const int nav = nav_of(G, subset, nsub);
  hipLaunchKernelGGL(k_back, grid_img(G, G.nzs), dim3(kWave, 4), 0, st, G, corr, mask, numer,
                     denom, subset, nsub, nav);
  return hipGetLastError();
}

hipError_t launch_update(const DevGeom& G, const Taps9& t9, const Taps5& t5, const float* numer,
                         const float* denom, float* x, float* work, hipStream_t st) {
  hipLaunchKernelGGL(k_blur_xy<SrcRatio>, grid_img(G, G.nzs), dim3(kWave, 4), 0, st, G.res,
                     G.nzs, t9, SrcRatio{numer, denom}, work);
  hipError_t e = hipGetLastError();
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
  hipLaunchKernelGGL(k_jbf, grid_img(G, G.nzs), dim3(kWave, 4), 0, st, G.res, G.nzs, in, guide,
                     d_max_ref, max_ref, r, kI, gt, out);
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
