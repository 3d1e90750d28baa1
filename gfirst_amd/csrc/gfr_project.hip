// gfr_project.hip — the two projectors of the OSEM path on gfx950.
//
// Forward projection (FORWPROJ_K, GFIRST_GPU_Kernels.cu:40-102): a LOR of view
// (segment, angle theta), plane p, radial bin r sums RES+1 trilinear samples of
// the blurred image at (xc + l dx, yc + l dy, p + 0.5 + l dz). The x/y sample
// positions depend only on (theta, r, l); only z depends on the segment and
// plane. Trilinear = lerp_z(bilinear_xy(plane z0), bilinear_xy(plane z1)), so
// the kernel pair below computes every bilinear value once per angle:
//   k_rotate: R[a][l][z][r] = bilinear_xy(image plane z, x(theta,r,l), y(theta,r,l))
//   k_zsum  : P(s, r) = sum_l lerp(R[a][l][z0][r], R[a][l][z1][r], alpha_z)
// which is the same arithmetic, operation for operation, as sampling the
// volume directly (bit-identical results). The z interpolation of sinogram s
// at sample l, (z0, z1, alpha_z), is a per-geometry table, so a sample costs
// two coalesced buffer loads with scalar offsets and one lerp.
//
// Backprojection (BACKWPROJ_K, GFIRST_GPU_Kernels.cu:109-164): one thread
// per voxel column and ZB consecutive slices, views in the reference's order
// (segment-major, angle-minor, GFIRST.cu:439-465). Per view the radial
// interpolation (r0, r1, alpha_u) is shared by the thread's slices; the
// ZB+1 planes its slices need are fetched together (one memory round trip
// per view, row addresses without tests thanks to the padded packed rows)
// and radially interpolated once each. Each slice keeps its own exactly
// computed plane index and weight, as the reference's per-voxel sampling.
#include "gfr_device.h"

namespace gfr {

namespace {

// Raw buffer resource over [p, p + bytes): 32-bit offsets, loads past the end
// return 0 without touching memory.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc(const void* p, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes,
                                           0x00020000);
}
__device__ __forceinline__ float ld_f32(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
__device__ __forceinline__ float2 ld_f2(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
  return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}

// ---------------------------------------------------------------------------
// k_rotate: one thread per (angle a of the chunk, sample l, plane z, bin r).
// Logical block order (r-block, l-block, z, a) with XCD-contiguous ranges, so
// the work-groups sharing an image plane share an L2.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_rotate(DevGeom G, const float* __restrict__ img,
                                                float* __restrict__ R, int subset, int nsub,
                                                const int* __restrict__ dsub, int a0, int nl) {
  if (dsub) subset = *dsub;
  const unsigned b = xcd_remap(blockIdx.x, gridDim.x);
  const unsigned nrb = (G.nrad + kWave - 1) / kWave, nlb = (nl + 3) / 4;
  const unsigned rb = b % nrb;
  unsigned q = b / nrb;
  const unsigned lb = q % nlb;
  q /= nlb;
  const int z = (int)(q % G.nzs);
  const int a = (int)(q / G.nzs);
  const int r = rb * kWave + threadIdx.x;
  const int li = lb * 4 + threadIdx.y;
  if (r >= G.nrad || li >= nl || a0 + a >= (G.nang - subset + nsub - 1) / nsub) return;
  const int it = subset + (a0 + a) * nsub;
  const float st = G.sin_t[it], ct = G.cos_t[it];
  const float x0 = G.fxf * ((float)r - G.RADM);
  const float xc = fmaf(x0, ct, G.RESM) + 0.5f;
  const float yc = fmaf(x0, st, G.RESM) + 0.5f;
  const float fl = (float)(G.l0 + li);
  const float x = fmaf(fl, st, xc);   // dx = sin(theta)   (GFIRST_GPU_Kernels.cu:81)
  const float y = fmaf(fl, -ct, yc);  // dy = -cos(theta)  (:82)
  int x0i, x1i, y0i, y1i;
  float ax, ay;
  axis(x, G.res, x0i, x1i, ax);
  axis(y, G.res, y0i, y1i, ay);
  const float* p = img + (size_t)z * G.res * G.res;
  const float v = lerp1(lerp1(p[y0i * G.res + x0i], p[y0i * G.res + x1i], ax),
                        lerp1(p[y1i * G.res + x0i], p[y1i * G.res + x1i], ax), ay);
  R[(((size_t)a * nl + li) * G.nzs + z) * G.nrad + r] = v;
}

// Where the forward kernel writes.
struct FwdArgs {
  float* proj;           // full sinogram layout, or NULL
  float* corr;           // full layout ratio, or NULL
  float2* corr2;         // packed [angle of subset][s][r] {ratio, calib}, or NULL
  const uint8_t* mask;   // calibration mask (full layout) for corr2
  const void* y;         // measured prompts (storage format), unscaled
  const void* bg;        // background or NULL
  float yscale;          // RES (GFIRST.cu:224-260)
  double* llh_part;      // per-block FP64 partial log-likelihoods, or NULL
};

// ---------------------------------------------------------------------------
// k_zsum: one wave per (angle a, SG consecutive sinograms, 128 radial bins);
// lane l owns bins 2l and 2l+1 of each sinogram of the group, so a row load is
// one contiguous 512-byte dwordx2 access. The z interpolation of a sample is
// wave-uniform (scalar loads of the transposed z table). When the group's
// samples at l blend planes z0, z0+1, ..., z0+SG (consecutive planes: the
// usual case inside a segment), the SG+1 rows are loaded once for the group;
// otherwise each LOR loads its own two rows. Every LOR still sums its samples
// in l order with its own weights. Epilogue: ratio (GFIRST.cu:415-430) and the
// FP64 Poisson log-likelihood partial of the block.
// ---------------------------------------------------------------------------
template <int S, bool RATIO, bool LLH, int SG>
__global__ __launch_bounds__(256) void k_zsum(DevGeom G, const float* __restrict__ R,
                                              unsigned r_bytes, FwdArgs o, int subset, int nsub,
                                              const int* __restrict__ dsub, int a0, int nl,
                                              unsigned llh_off) {
  if (dsub) subset = *dsub;
  const unsigned b = xcd_remap(blockIdx.x, gridDim.x);
  const unsigned nrb = (G.nrad + 127) / 128, ngb = (G.nsinogs + 4 * SG - 1) / (4 * SG);
  const unsigned rb = b % nrb;
  const unsigned q = b / nrb;
  const unsigned gb = q % ngb;
  const int a = (int)(q / ngb);
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int s0 = __builtin_amdgcn_readfirstlane((int)((gb * 4 + wv) * SG));
  const int r = (int)rb * 128 + 2 * lane;
  const int ag = a0 + a;
  const int it = subset + ag * nsub;
  if (ag >= (G.nang - subset + nsub - 1) / nsub) return;  // whole block: angle not in subset
  double ll = 0.0;
  if (s0 < G.nsinogs) {
    const int ns = min(SG, G.nsinogs - s0);
    const __amdgpu_buffer_rsrc_t rr = rsrc(R, r_bytes);
    const unsigned rowb = (unsigned)G.nrad * 4u;
    const unsigned planeb = rowb * (unsigned)G.nzs;
    const unsigned voff = (unsigned)r * 4u;
    float acc0[SG], acc1[SG];
#pragma unroll
    for (int i = 0; i < SG; i++) acc0[i] = acc1[i] = 0.f;
    unsigned soff = (unsigned)a * (unsigned)nl * planeb;
    for (int li = 0; li < nl; li++, soff += planeb) {
      const uint2* zt = G.ztab + (size_t)li * G.nsinogs + s0;
      uint2 e[SG];
#pragma unroll
      for (int i = 0; i < SG; i++) e[i] = zt[i < ns ? i : ns - 1];
      const unsigned z00 = e[0].x & 0xFFFFu;
      bool reg = true;
#pragma unroll
      for (int i = 0; i < SG; i++)
        reg = reg && (e[i].x == ((z00 + i) | ((z00 + i + 1) << 16)));
      if (reg) {
        float2 row[SG + 1];
#pragma unroll
        for (int k = 0; k <= SG; k++) row[k] = ld_f2(rr, voff, soff + (z00 + k) * rowb);
#pragma unroll
        for (int i = 0; i < SG; i++) {
          const float al = __uint_as_float(e[i].y);
          acc0[i] += lerp1(row[i].x, row[i + 1].x, al);
          acc1[i] += lerp1(row[i].y, row[i + 1].y, al);
        }
      } else {
#pragma unroll
        for (int i = 0; i < SG; i++) {
          const float al = __uint_as_float(e[i].y);
          const float2 v0 = ld_f2(rr, voff, soff + (e[i].x & 0xFFFFu) * rowb);
          const float2 v1 = ld_f2(rr, voff, soff + (e[i].x >> 16) * rowb);
          acc0[i] += lerp1(v0.x, v1.x, al);
          acc1[i] += lerp1(v0.y, v1.y, al);
        }
      }
    }
#pragma unroll
    for (int i = 0; i < SG; i++) {
      if (i >= ns) break;
      const int s = s0 + i;
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const int rr2 = r + h;
        if (rr2 >= G.nrad) break;
        const float acc = h ? acc1[i] : acc0[i];
        const size_t bin = ((size_t)s * G.nang + it) * G.nrad + rr2;
        if (o.proj) o.proj[bin] = acc;
        if constexpr (RATIO) {
          const float yv = o.yscale * load_store<S>(o.y, bin);
          float bv = o.bg ? o.yscale * load_store<S>(o.bg, bin) : 0.f;
          if (bv < 0.f) bv = 0.f;
          const float den = acc + bv;
          const float c = (den >= 0.01f) ? yv / den : -1000.f;
          if (o.corr) o.corr[bin] = c;
          if (o.corr2)
            store_packed(G, o.corr2, ag, s, rr2, make_float2(c, o.mask[bin] ? 1.f : -1000.f));
          if constexpr (LLH) {
            if (den > 0.f) ll += (double)yv * log((double)den) - (double)den;
          }
        }
      }
    }
  }
  if constexpr (LLH) {
    __shared__ double red[256];
    const int t = threadIdx.x;
    red[t] = ll;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if (t < w) red[t] += red[t + w];
      __syncthreads();
    }
    if (t == 0) o.llh_part[llh_off + b] = red[0];
  }
}

// Full-layout ratio + byte mask -> packed {ratio, calib} rows of a subset.
__global__ void k_pack(DevGeom G, const float* __restrict__ corr, const uint8_t* __restrict__ mask,
                       float2* __restrict__ c2, int subset, int nsub) {
  const int r = blockIdx.x * kWave + threadIdx.x;
  const int s = blockIdx.y * 4 + threadIdx.y;
  const int a = blockIdx.z;
  if (r >= G.nrad || s >= G.nsinogs) return;
  const size_t bin = ((size_t)s * G.nang + subset + a * nsub) * G.nrad + r;
  store_packed(G, c2, a, s, r, make_float2(corr[bin], mask[bin] ? 1.f : -1000.f));
}


// Block: 4 waves, each an 8x8 tile of voxel columns (16x16 per block) times
// ZB slices. Across an 8x8 tile the z tilt of a view moves by a few planes
// only, so a wave with no slice inside the view's plane range skips the view
// (wave-uniform test); the tile's radial bins are a few neighbouring bins, so
// row loads stay within a few cache lines.
template <int ZB>
__global__ __launch_bounds__(256) void k_back(DevGeom G, const float2* __restrict__ c2,
                                              unsigned c2_bytes, float* __restrict__ numer,
                                              float* __restrict__ denom, int subset, int nsub,
                                              const int* __restrict__ dsub) {
  if (dsub) subset = *dsub;
  const int nav = (G.nang - subset + nsub - 1) / nsub;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int ix = blockIdx.x * 16 + (wv & 1) * 8 + (lane & 7);
  const int iy = blockIdx.y * 16 + (wv >> 1) * 8 + (lane >> 3);
  const int zb = blockIdx.z * ZB;
  const bool act = ix < G.res && iy < G.res;
  const __amdgpu_buffer_rsrc_t rc2 = rsrc(c2, c2_bytes);
  const float xv = (float)ix - G.RESM, yv = (float)iy - G.RESM;
  const unsigned rowb = (unsigned)G.nrad * 8u;
  const unsigned nr = (unsigned)(G.nsinogs + 2 * G.nseg);
  const float fz0 = (float)zb, fzl = (float)(zb + ZB - 1);
  float n[ZB], d[ZB];
#pragma unroll
  for (int j = 0; j < ZB; j++) n[j] = d[j] = 0.f;
  for (int si = 0; si < G.nseg; si++) {
    const SegEntry se = G.seg[si];
    const int zm = se.zmin;
    const int zhi = G.nzs - zm;
    const float fzm = (float)zm, fzhi = (float)zhi;
    // row of plane p of this segment within an angle: segrow + p (planes
    // zm-1 .. zhi are the padded rows)
    const unsigned segrow = (unsigned)(se.nza + 2 * si + 1 - zm);
    for (int a = 0; a < nav; a++) {
      const int it = subset + a * nsub;
      const float st = G.sin_t[it], ct = G.cos_t[it];
      const float yr = fmaf(yv, ct, -(xv * st));
      const float w0 = fmaf(yr, se.fzb, fz0) + 0.5f;
      const float wl = fmaf(yr, se.fzb, fzl) + 0.5f;
      const bool any = act && wl >= fzm && w0 < fzhi;
      if (__builtin_amdgcn_ballot_w64(any) == 0) continue;
      const float xr = fmaf(yv, st, xv * ct);
      const float u = fmaf(xr, G.fxb, G.RADM) + 0.5f;
      int r0, r1;
      float au;
      axis(u, G.nrad, r0, r1, au);
      const unsigned o0 = (unsigned)r0 * 8u, o1 = (unsigned)r1 * 8u;
      const unsigned vrow = (unsigned)a * nr + segrow;
      float al[ZB];
      int ii[ZB];
      bool inr[ZB];
#pragma unroll
      for (int j = 0; j < ZB; j++) {
        const float w = (j == 0) ? w0 : (j == ZB - 1) ? wl : fmaf(yr, se.fzb, (float)(zb + j)) + 0.5f;
        inr[j] = act && (zb + j < G.nzs) && w >= fzm && w < fzhi;
        const float t = w - 0.5f;
        const float f = floorf(t);
        al[j] = t - f;
        ii[j] = (int)f;
      }
      bool reg = true;
#pragma unroll
      for (int j = 1; j < ZB; j++) reg = reg && (!inr[j] || ii[j] == ii[0] + j);
      float av[ZB], bv[ZB];
      if (reg) {
        // planes ii0 + k, k = 0..ZB: slice j blends planes j and j+1
        const unsigned rb0 = (vrow + (unsigned)ii[0]) * rowb;
        float c[ZB + 1], q[ZB + 1];
        float2 v0[ZB + 1], v1[ZB + 1];
#pragma unroll
        for (int k = 0; k <= ZB; k++) {
          v0[k] = ld_f2(rc2, rb0 + o0, (unsigned)k * rowb);
          v1[k] = ld_f2(rc2, rb0 + o1, (unsigned)k * rowb);
        }
#pragma unroll
        for (int k = 0; k <= ZB; k++) {
          c[k] = lerp1(v0[k].x, v1[k].x, au);
          q[k] = lerp1(v0[k].y, v1[k].y, au);
        }
#pragma unroll
        for (int j = 0; j < ZB; j++) {
          av[j] = inr[j] ? lerp1(c[j], c[j + 1], al[j]) : 0.f;
          bv[j] = inr[j] ? lerp1(q[j], q[j + 1], al[j]) : 0.f;
        }
      } else {
#pragma unroll
        for (int j = 0; j < ZB; j++) {
          av[j] = 0.f;
          bv[j] = 0.f;
          if (inr[j]) {
            const unsigned ra = (vrow + (unsigned)ii[j]) * rowb;
            const float2 w00 = ld_f2(rc2, ra + o0, 0), w01 = ld_f2(rc2, ra + o1, 0);
            const float2 w10 = ld_f2(rc2, ra + o0, rowb), w11 = ld_f2(rc2, ra + o1, rowb);
            av[j] = lerp1(lerp1(w00.x, w01.x, au), lerp1(w10.x, w11.x, au), al[j]);
            bv[j] = lerp1(lerp1(w00.y, w01.y, au), lerp1(w10.y, w11.y, au), al[j]);
          }
        }
      }
#pragma unroll
      for (int j = 0; j < ZB; j++) {
        float A = av[j], B = bv[j];
        if (B < 0.f || A < 0.f) {  // gaps (GFIRST_GPU_Kernels.cu:157)
          A = 0.f;
          B = 0.f;
        }
        n[j] += A;
        d[j] += B;
      }
    }
  }
  if (!act) return;
#pragma unroll
  for (int j = 0; j < ZB; j++) {
    const int iz = zb + j;
    if (iz < G.nzs) {
      const size_t iV = ((size_t)iz * G.res + iy) * G.res + ix;
      numer[iV] = n[j];
      denom[iV] = d[j];
    }
  }
}

inline int nav_of(const DevGeom& G, int subset, int nsub) {
  return (G.nang - subset + nsub - 1) / nsub;
}

}  // namespace

size_t rotate_floats_per_angle(const DevGeom& G) {
  return (size_t)(G.l1 - G.l0 + 1) * G.nzs * G.nrad;
}

int forward_llh_blocks(const DevGeom& G, int subset, int nsub, int sg) {
  return ((G.nrad + 127) / 128) * ((G.nsinogs + 4 * sg - 1) / (4 * sg)) * nav_of(G, subset, nsub);
}

namespace {

template <int SG>
hipError_t zsum_launch(const DevGeom& G, const float* R, unsigned rbytes, const FwdArgs& o,
                       bool ratio, int storage, int subset, int nsub, const int* dsub, int a0,
                       int na, int nl, hipStream_t st) {
  const unsigned zrb = (G.nrad + 127) / 128, zgb = (G.nsinogs + 4 * SG - 1) / (4 * SG);
  const unsigned gz = zrb * zgb * na;
  const unsigned off = zrb * zgb * a0;
  if (!ratio) {
    hipLaunchKernelGGL((k_zsum<GFR_STORE_F32, false, false, SG>), dim3(gz), dim3(256), 0, st, G,
                       R, rbytes, o, subset, nsub, dsub, a0, nl, off);
    return hipGetLastError();
  }
#define GFR_ZS(S)                                                                             \
  if (o.llh_part)                                                                             \
    hipLaunchKernelGGL((k_zsum<S, true, true, SG>), dim3(gz), dim3(256), 0, st, G, R, rbytes, \
                       o, subset, nsub, dsub, a0, nl, off);                                   \
  else                                                                                        \
    hipLaunchKernelGGL((k_zsum<S, true, false, SG>), dim3(gz), dim3(256), 0, st, G, R,       \
                       rbytes, o, subset, nsub, dsub, a0, nl, off);
  if (storage == GFR_STORE_F16) {
    GFR_ZS(GFR_STORE_F16)
  } else if (storage == GFR_STORE_BF16) {
    GFR_ZS(GFR_STORE_BF16)
  } else {
    GFR_ZS(GFR_STORE_F32)
  }
#undef GFR_ZS
  return hipGetLastError();
}

// Deterministic sum of the first n(subset) partials; with dsub the count
// follows the device-side subset index.
__global__ void k_llh_sum(const double* __restrict__ part, int per_angle, int nang, int subset,
                          int nsub, const int* __restrict__ dsub, double* __restrict__ out) {
  if (dsub) subset = *dsub;
  const int n = per_angle * ((nang - subset + nsub - 1) / nsub);
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

}  // namespace

hipError_t launch_forward(const DevGeom& G, const float* img, float* R, size_t r_cap_floats,
                          const FwdOut& out, int subset, int nsub, const int* dsub,
                          const Tune& tune, hipStream_t st) {
  // with a device-side subset index the launch covers the largest subset
  const int nav = dsub ? nav_of(G, 0, nsub) : nav_of(G, subset, nsub);
  const int nl = G.l1 - G.l0 + 1;
  const size_t per = rotate_floats_per_angle(G);
  int chunk = (int)(r_cap_floats / per);
  if (chunk < 1) return hipErrorInvalidValue;
  if (chunk > nav) chunk = nav;
  if ((size_t)chunk * per * sizeof(float) >= 0xFFFFFFF0u) chunk = (int)(0xFFFFFFF0u / (per * 4));
  if (chunk < 1) return hipErrorInvalidValue;
  FwdArgs o{out.proj, out.corr, out.corr2, out.mask, out.y, out.bg, out.yscale, out.llh_part};
  const bool ratio = out.corr || out.corr2;
  const unsigned nrb = (G.nrad + kWave - 1) / kWave;
  for (int a0 = 0; a0 < nav; a0 += chunk) {
    const int na = min(chunk, nav - a0);
    const unsigned rbytes = (unsigned)((size_t)na * per * sizeof(float));
    const unsigned grot = nrb * ((nl + 3) / 4) * G.nzs * na;
    hipLaunchKernelGGL(k_rotate, dim3(grot), dim3(kWave, 4), 0, st, G, img, R, subset, nsub, dsub,
                       a0, nl);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    switch (tune.zsum_sg) {
      case 4: e = zsum_launch<4>(G, R, rbytes, o, ratio, out.storage, subset, nsub, dsub, a0, na, nl, st); break;
      case 16: e = zsum_launch<16>(G, R, rbytes, o, ratio, out.storage, subset, nsub, dsub, a0, na, nl, st); break;
      default: e = zsum_launch<8>(G, R, rbytes, o, ratio, out.storage, subset, nsub, dsub, a0, na, nl, st); break;
    }
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_llh_sum(const DevGeom& G, const double* part, int subset, int nsub,
                          const int* dsub, int sg, double* out, hipStream_t st) {
  const int per_angle = ((G.nrad + 127) / 128) * ((G.nsinogs + 4 * sg - 1) / (4 * sg));
  hipLaunchKernelGGL(k_llh_sum, dim3(1), dim3(256), 0, st, part, per_angle, G.nang, subset, nsub,
                     dsub, out);
  return hipGetLastError();
}

size_t packed_bytes(const DevGeom& G, int subset, int nsub) {
  return sizeof(float2) * (size_t)nav_of(G, subset, nsub) * (G.nsinogs + 2 * (size_t)G.nseg) *
         G.nrad;
}

hipError_t launch_pack(const DevGeom& G, const float* corr, const uint8_t* mask, float2* c2,
                       int subset, int nsub, hipStream_t st) {
  dim3 grid((G.nrad + kWave - 1) / kWave, (G.nsinogs + 3) / 4, nav_of(G, subset, nsub));
  hipLaunchKernelGGL(k_pack, grid, dim3(kWave, 4), 0, st, G, corr, mask, c2, subset, nsub);
  return hipGetLastError();
}

hipError_t launch_back_packed(const DevGeom& G, const float2* corr2, float* numer, float* denom,
                              int subset, int nsub, const int* dsub, const Tune& tune,
                              hipStream_t st) {
  const size_t bytes = packed_bytes(G, dsub ? 0 : subset, nsub);
  if (bytes >= 0xFFFFFFF0u) return hipErrorInvalidValue;
#define GFR_BK(ZB)                                                                                hipLaunchKernelGGL(k_back<ZB>, dim3((G.res + 15) / 16, (G.res + 15) / 16, (G.nzs + ZB - 1) / ZB),                      dim3(256), 0, st, G, corr2, (unsigned)bytes, numer, denom, subset, nsub, dsub)
  switch (tune.back_zb) {
    case 4: GFR_BK(4); break;
    case 16: GFR_BK(16); break;
    default: GFR_BK(8); break;
  }
#undef GFR_BK
  return hipGetLastError();
}

int llh_capacity(const GeomHost& h, int nsub) {
  const int nav = (h.g.nang + nsub - 1) / nsub;
  return ((h.g.nrad + 127) / 128) * ((h.info.nsinogs + 15) / 16) * nav;  // zsum_sg = 4
}

hipError_t tune_shapes(const DevGeom& G, int nsub, const FwdOut& fo, const float* img, float* R,
                       size_t r_floats, const float2* corr2, float* num, float* den,
                       hipStream_t st, Tune* out) {
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return hipErrorUnknown;
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return hipErrorUnknown;
  }
  hipError_t e = hipSuccess;
  // best of two timed runs after one warm-up
  auto time_it = [&](auto&& launch) -> float {
    float best = 1e30f;
    for (int rep = 0; rep < 3 && e == hipSuccess; rep++) {
      e = hipEventRecord(e0, st);
      if (e == hipSuccess) e = launch();
      if (e == hipSuccess) e = hipEventRecord(e1, st);
      if (e == hipSuccess) e = hipEventSynchronize(e1);
      float ms = 0.f;
      if (e == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess && rep > 0)
        best = ms < best ? ms : best;
    }
    return best;
  };
  Tune t;
  const int cand[3] = {4, 8, 16};
  float bz = 1e30f, bb = 1e30f;
  for (int c : cand) {
    Tune tt = t;
    tt.zsum_sg = c;
    const float ms = time_it([&] {
      return launch_forward(G, img, R, r_floats, fo, 0, nsub, nullptr, tt, st);
    });
    if (ms < bz) {
      bz = ms;
      t.zsum_sg = c;
    }
  }
  for (int c : cand) {
    Tune tt = t;
    tt.back_zb = c;
    const float ms = time_it([&] {
      return launch_back_packed(G, corr2, num, den, 0, nsub, nullptr, tt, st);
    });
    if (ms < bb) {
      bb = ms;
      t.back_zb = c;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (e != hipSuccess) return e;
  t.tuned = true;
  t.ms_zsum = bz;
  t.ms_back = bb;
  *out = t;
  return hipSuccess;
}

}  // namespace gfr
