// gfr_project.hip — the two projectors of the OSEM path on gfx950.
//
// Forward projection (FORWPROJ_K, GFIRST_GPU_Kernels.cu:40-102): a LOR of view
// (segment, angle theta), plane p, radial bin r sums RES+1 trilinear samples of
// the blurred image at (xc + l dx, yc + l dy, p + 0.5 + l dz). The x/y sample
// positions depend only on (theta, r, l); only z depends on the segment and
// plane. Trilinear = lerp_z(bilinear_xy(plane z0), bilinear_xy(plane z1)), so
// the kernel pair below computes every bilinear value once per angle:
//   k_rotate_tile: R[a][l][z][r] = bilinear_xy(image plane z, x(theta,r,l), y(theta,r,l))
//   k_zsum       : P(s, r) = sum_l lerp(R[a][l][z0][r], R[a][l][z1][r], alpha_z)
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

// ---------------------------------------------------------------------------
// k_rotate_tile: R[a][l][z][r], the bilinear sample of image plane z at the
// position of (bin r, sample l) of angle a. A block takes one angle, a tile
// of 32 bins x 32 samples and ZC planes (autotuned). The tile's sample
// points fall in a rotated rectangle of the image; since x and y (and then
// the clamped texel indices) are monotone in r and l, the texels the tile
// touches lie in the bounding box of its four corner samples. Per plane
// the block copies that box into LDS with coalesced row loads (the next
// plane's rows are in flight while the current one is used) and every
// output reads its four texels from LDS. A box larger than the LDS tile
// (coarse geometries) falls back to gathers from global memory.
// ---------------------------------------------------------------------------
constexpr int kRotTile = 64;         // LDS box capacity: texels per row and rows
constexpr int kRotStride = kRotTile + 1;  // LDS row stride (odd: rows on different banks)

template <int ZC>
__global__ __launch_bounds__(256) void k_rotate_tile(DevGeom G, const float* __restrict__ img,
                                                     float* __restrict__ R, int subset, int nsub,
                                                     const int* __restrict__ dsub, int a0,
                                                     int nl, int zouter) {
  __shared__ float box[kRotTile * kRotStride];
  if (dsub) subset = *dsub;
  // Logical order: bin tile, sample tile, then angle and plane chunk; every
  // XCD walks one contiguous range of logical blocks. With zouter the plane
  // chunk is outermost: an XCD rotates a few plane chunks for every angle
  // and reads each image plane once into its L2 (rotation outputs larger than
  // an L2). Otherwise the angle is: an XCD writes the angles whose pairs its
  // z-sum blocks read next, while they are still in its L2.
  const unsigned b = xcd_remap(blockIdx.x, gridDim.x);
  const unsigned nrt = (G.nrad + 31) / 32, nlt = (nl + 31) / 32;
  const unsigned nzc = (G.nzs + ZC - 1) / ZC;
  const unsigned na = gridDim.x / (nrt * nlt * nzc);
  const unsigned rt = b % nrt;
  unsigned q = b / nrt;
  const unsigned lt = q % nlt;
  q /= nlt;
  const int a = (int)(zouter ? q % na : q / nzc);
  const int zc = (int)(zouter ? q / na : q % nzc) * ZC;
  if (a0 + a >= (G.nang - subset + nsub - 1) / nsub) return;  // whole block
  const int it = subset + (a0 + a) * nsub;
  const float st = G.sin_t[it], ct = G.cos_t[it];
  const int tid = threadIdx.x;
  // texels and weights of sample (r, li)
  auto texels = [&](int r, int li, int& x0i, int& x1i, float& ax, int& y0i, int& y1i,
                    float& ay) {
    const float x0 = G.fxf * ((float)r - G.RADM);
    const float xc = fmaf(x0, ct, G.RESM) + 0.5f;
    const float yc = fmaf(x0, st, G.RESM) + 0.5f;
    const float fl = (float)(G.l0 + li);
    axis(fmaf(fl, st, xc), G.res, x0i, x1i, ax);   // dx = sin(theta) (GFIRST_GPU_Kernels.cu:81)
    axis(fmaf(fl, -ct, yc), G.res, y0i, y1i, ay);  // dy = -cos(theta) (:82)
  };
  const int rlo = (int)rt * 32, rhi = min(rlo + 31, G.nrad - 1);
  const int llo = (int)lt * 32, lhi = min(llo + 31, nl - 1);
  int bx0 = G.res, bx1 = -1, by0 = G.res, by1 = -1;
#pragma unroll
  for (int c = 0; c < 4; c++) {
    int x0i, x1i, y0i, y1i;
    float ax, ay;
    texels(c & 1 ? rhi : rlo, c & 2 ? lhi : llo, x0i, x1i, ax, y0i, y1i, ay);
    bx0 = min(bx0, x0i);
    bx1 = max(bx1, x1i);
    by0 = min(by0, y0i);
    by1 = max(by1, y1i);
  }
  const int bw = bx1 - bx0 + 1, bh = by1 - by0 + 1;
  const bool staged = bw <= kRotTile && bh <= kRotTile;
  // this thread's four outputs: bin rlo + (tid & 31), samples llo + (tid >> 5) + 8k
  const int r = rlo + (tid & 31);
  int o00[4], dx[4], dy[4];
  float ax[4], ay[4];
  bool ok[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int li = llo + (tid >> 5) + 8 * k;
    ok[k] = r < G.nrad && li < nl;
    int x0i, x1i, y0i, y1i;
    texels(r, li, x0i, x1i, ax[k], y0i, y1i, ay[k]);
    if (staged) {
      o00[k] = (y0i - by0) * kRotStride + (x0i - bx0);
      dx[k] = x1i - x0i;
      dy[k] = (y1i - y0i) * kRotStride;
    } else {
      o00[k] = y0i * G.res + x0i;
      dx[k] = x1i - x0i;
      dy[k] = (y1i - y0i) * G.res;
    }
  }
  const size_t plane = (size_t)G.res * G.res;
  const int nz = min(ZC, G.nzs - zc);
  float* out = R + (((size_t)a * nl + llo + (tid >> 5)) * G.nzs + zc) * G.nrad + r;
  const size_t lstep = (size_t)8 * G.nzs * G.nrad;  // 8 samples further
  if (!staged) {
    for (int zi = 0; zi < nz; zi++) {
      const float* pk = img + (size_t)(zc + zi) * plane;
#pragma unroll
      for (int k = 0; k < 4; k++) {
        if (ok[k]) {
          const float* p = pk + o00[k];
          out[k * lstep + (size_t)zi * G.nrad] =
              lerp1(lerp1(p[0], p[dx[k]], ax[k]), lerp1(p[dy[k]], p[dy[k] + dx[k]], ax[k]), ay[k]);
        }
      }
    }
    return;
  }
  // box loader: thread copies column tid & 63 of rows tid / 64 + 4 j
  const int lc = tid & 63, lr = tid >> 6;
  const int nj = (bh + 3) >> 2;
  const float* src = img + (size_t)(by0 + lr) * G.res + bx0 + lc;
  float pre[kRotTile / 4];
  auto fetch = [&](int z) {
    const float* pz = src + (size_t)z * plane;
#pragma unroll
    for (int j = 0; j < kRotTile / 4; j++)
      if (j < nj && lc < bw && lr + 4 * j < bh) pre[j] = pz[(size_t)(4 * j) * G.res];
  };
  fetch(zc);
  for (int zi = 0; zi < nz; zi++) {
    __syncthreads();  // the previous plane's reads are done
#pragma unroll
    for (int j = 0; j < kRotTile / 4; j++)
      if (j < nj && lc < bw && lr + 4 * j < bh) box[(lr + 4 * j) * kRotStride + lc] = pre[j];
    __syncthreads();
    if (zi + 1 < nz) fetch(zc + zi + 1);
#pragma unroll
    for (int k = 0; k < 4; k++) {
      if (ok[k]) {
        const float* p = box + o00[k];
        out[k * lstep + (size_t)zi * G.nrad] =
            lerp1(lerp1(p[0], p[dx[k]], ax[k]), lerp1(p[dy[k]], p[dy[k] + dx[k]], ax[k]), ay[k]);
      }
    }
  }
}

// Where the forward kernel writes.
struct FwdArgs {
  float* proj;           // full sinogram layout, or NULL
  float* corr;           // full layout ratio, or NULL
  float2* corr2;         // packed [angle of subset][s][r] {ratio, calib}, or NULL
  float* corr4;          // interleaved rows of a frame group (channel chan), or NULL
  int chan;
  const uint8_t* mask;   // calibration mask (full layout) for corr2 / corr4
  const void* y;         // measured prompts (storage format), unscaled
  const void* bg;        // background or NULL
  float yscale;          // RES (GFIRST.cu:224-260)
  double* llh_part;      // per-block FP64 partial log-likelihoods, or NULL
};

// ---------------------------------------------------------------------------
// k_zsum: one wave per (pair of angles, SG consecutive sinograms, 128 radial
// bins). Half-wave h takes angle 2p + h of the chunk and lane l of the half
// owns bins 4l .. 4l+3, so a row load is one 16-byte access per lane and one
// instruction fetches a row of both angles (the z interpolation depends on
// the sinogram and the sample, not on the angle). The z interpolation of a
// sample is wave-uniform (scalar loads of the transposed z table). When the
// group's samples at l blend planes z0, z0+1, ..., z0+SG (consecutive planes:
// the usual case inside a segment), the SG+1 rows are loaded once for the
// group; otherwise each LOR loads its own two rows. Every LOR still sums its
// samples in l order with its own weights. Epilogue: ratio
// (GFIRST.cu:415-430) and the FP64 Poisson log-likelihood partial of the
// block (bins with valid calibration and den >= 0.01: sum of (y/RES)
// ln(den/RES) - den/RES; the reference computes no likelihood).
// ---------------------------------------------------------------------------
template <int S, bool RATIO, bool LLH, int SG>
__global__ __launch_bounds__(256) void k_zsum(DevGeom G, const float* __restrict__ R,
                                              unsigned r_bytes, FwdArgs o, int subset, int nsub,
                                              const int* __restrict__ dsub, int a0, int na,
                                              int nl, unsigned llh_off) {
  if (dsub) subset = *dsub;
  const unsigned b = xcd_remap(blockIdx.x, gridDim.x);
  const unsigned nrb = (G.nrad + 127) / 128, ngb = (G.nsinogs + 4 * SG - 1) / (4 * SG);
  const unsigned rb = b % nrb;
  const unsigned q = b / nrb;
  const unsigned gb = q % ngb;
  const int ap = (int)(q / ngb);
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int s0 = __builtin_amdgcn_readfirstlane((int)((gb * 4 + wv) * SG));
  const int a = 2 * ap + (lane >> 5);  // angle of the chunk
  const int r = (int)rb * 128 + 4 * (lane & 31);
  const int nav = (G.nang - subset + nsub - 1) / nsub;
  const int ag = a0 + a;
  const int it = subset + ag * nsub;
  if (a0 + 2 * ap >= nav) return;  // whole block: no angle of the pair in the subset
  const bool aok = a < na && ag < nav;
  double ll = 0.0;
  if (s0 < G.nsinogs) {
    const int ns = min(SG, G.nsinogs - s0);
    // Lanes of an angle past the chunk read past r_bytes: zeros, unused.
    const __amdgpu_buffer_rsrc_t rr = rsrc(R, r_bytes);
    const unsigned rowb = (unsigned)G.nrad * 4u;
    const unsigned planeb = rowb * (unsigned)G.nzs;
    const unsigned voff = (unsigned)a * (unsigned)nl * planeb + (unsigned)r * 4u;
    float4 acc[SG];
#pragma unroll
    for (int i = 0; i < SG; i++) acc[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    unsigned soff = 0;
    // z-table entries of sample li, loaded one sample ahead
    uint2 en[SG];
#pragma unroll
    for (int i = 0; i < SG; i++) en[i] = G.ztab[s0 + (i < ns ? i : ns - 1)];
    for (int li = 0; li < nl; li++, soff += planeb) {
      uint2 e[SG];
#pragma unroll
      for (int i = 0; i < SG; i++) e[i] = en[i];
      if (li + 1 < nl) {
        const uint2* zt = G.ztab + (size_t)(li + 1) * G.nsinogs + s0;
#pragma unroll
        for (int i = 0; i < SG; i++) en[i] = zt[i < ns ? i : ns - 1];
      }
      const unsigned z00 = e[0].x & 0xFFFFu;
      bool reg = true;
#pragma unroll
      for (int i = 0; i < SG; i++)
        reg = reg && (e[i].x == ((z00 + i) | ((z00 + i + 1) << 16)));
      if (reg) {
        float4 row[SG + 1];
#pragma unroll
        for (int k = 0; k <= SG; k++) row[k] = ld_f4(rr, voff, soff + (z00 + k) * rowb);
#pragma unroll
        for (int i = 0; i < SG; i++) {
          const float al = __uint_as_float(e[i].y);
          acc[i].x += lerp1(row[i].x, row[i + 1].x, al);
          acc[i].y += lerp1(row[i].y, row[i + 1].y, al);
          acc[i].z += lerp1(row[i].z, row[i + 1].z, al);
          acc[i].w += lerp1(row[i].w, row[i + 1].w, al);
        }
      } else {
#pragma unroll
        for (int i = 0; i < SG; i++) {
          const float al = __uint_as_float(e[i].y);
          const float4 v0 = ld_f4(rr, voff, soff + (e[i].x & 0xFFFFu) * rowb);
          const float4 v1 = ld_f4(rr, voff, soff + (e[i].x >> 16) * rowb);
          acc[i].x += lerp1(v0.x, v1.x, al);
          acc[i].y += lerp1(v0.y, v1.y, al);
          acc[i].z += lerp1(v0.z, v1.z, al);
          acc[i].w += lerp1(v0.w, v1.w, al);
        }
      }
    }
    if (aok) {
#pragma unroll
      for (int i = 0; i < SG; i++) {
        if (i >= ns) break;
        const int s = s0 + i;
        const float av[4] = {acc[i].x, acc[i].y, acc[i].z, acc[i].w};
#pragma unroll
        for (int h = 0; h < 4; h++) {
          const int rr2 = r + h;
          if (rr2 >= G.nrad) break;
          const float accv = av[h];
          const size_t bin = ((size_t)s * G.nang + it) * G.nrad + rr2;
          if (o.proj) o.proj[bin] = accv;
          if constexpr (RATIO) {
            const float yv = o.yscale * load_store<S>(o.y, bin);
            float bv = o.bg ? o.yscale * load_store<S>(o.bg, bin) : 0.f;
            if (bv < 0.f) bv = 0.f;
            const float den = accv + bv;
            const float c = (den >= 0.01f) ? yv / den : -1000.f;
            if (o.corr) o.corr[bin] = c;
            if (o.corr2)
              store_packed(G, o.corr2, ag, s, rr2, make_float2(c, o.mask[bin] ? 1.f : -1000.f));
            if (o.corr4)
              store_packed4(G, o.corr4, ag, s, rr2, o.chan, c, o.mask[bin] ? 1.f : -1000.f);
            if constexpr (LLH) {
              // Poisson log-likelihood in count units over the bins whose
              // ratio enters the backprojection: model mean den/RES, counts
              // yv/RES (yv and den as the ratio sees them).
              if (den >= 0.01f && (!o.mask || o.mask[bin])) {
                const double m = (double)den / (double)o.yscale;
                ll += ((double)yv / (double)o.yscale) * log(m) - m;
              }
            }
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

// Full-layout ratio (NULL: 1 everywhere) + byte mask -> packed {ratio,
// calib} rows of a subset.
__global__ void k_pack(DevGeom G, const float* __restrict__ corr, const uint8_t* __restrict__ mask,
                       float2* __restrict__ c2, int subset, int nsub) {
  const int r = blockIdx.x * kWave + threadIdx.x;
  const int s = blockIdx.y * 4 + threadIdx.y;
  const int a = blockIdx.z;
  if (r >= G.nrad || s >= G.nsinogs) return;
  const size_t bin = ((size_t)s * G.nang + subset + a * nsub) * G.nrad + r;
  store_packed(G, c2, a, s, r, make_float2(corr ? corr[bin] : 1.f, mask[bin] ? 1.f : -1000.f));
}


// ---------------------------------------------------------------------------
// Backprojector (BACKWPROJ_K, GFIRST_GPU_Kernels.cu:112-160): per voxel and
// view, bilinear (radial bin, plane) interpolation of the packed {ratio,
// calib} rows, the -1000 gap rule, sums over views in the reference's order
// (segments outer, angles inner, GFIRST.cu:439-467).
//
// Block: 4 waves, each an 8x8 tile of voxel columns (16x16 per block) times
// ZB slices. Per view the block needs a small window of the view's rows: the
// bins its 16x16 columns project to (<= 16*sqrt2*fxb + 2) times the planes
// its slices tilt to (<= ZB + 16*sqrt2*|fzb| + 2). The window is copied into
// LDS with coalesced loads, double-buffered so that the next view's window is
// in flight while the current one is used; the per-voxel reads then come from
// LDS. Window bounds are exact: every per-lane coordinate is a monotone
// function (correctly rounded fma/add, floor) of the column position and the
// slice, so its extremes over the block are at the corner columns and the
// first/last slice.
// ---------------------------------------------------------------------------
constexpr int kBW = 33;  // LDS row stride in float2 (odd: rows fall on different banks)
constexpr int kBinsMax = 32;

// Per-lane state of one view.
struct LaneView {
  float yr, w0, wl, s0, au;
  int r0, r1;  // clamped bins (GFIRST_GPU_Kernels.cu texture clamp)
  int iq;      // clamp(floor(u - 0.5), -1, NRAD-1): packed entry pair at iq + 1
  bool any;
};

__device__ __forceinline__ LaneView lane_view(const DevGeom& G, float xv, float yv, float st,
                                              float ct, float fzb, float fz0, float fzl,
                                              float fzm, float fzhi, bool act) {
  LaneView v;
  v.yr = fmaf(yv, ct, -(xv * st));
  v.s0 = fmaf(v.yr, fzb, fz0);
  const float sl = fmaf(v.yr, fzb, fzl);
  v.w0 = v.s0 + 0.5f;
  v.wl = sl + 0.5f;
  v.any = act && v.wl >= fzm && v.w0 < fzhi;
  const float xr = fmaf(yv, st, xv * ct);
  const float u = fmaf(xr, G.fxb, G.RADM) + 0.5f;
  const float t = u - 0.5f;
  const float f = floorf(t);
  const int i = (int)f;
  v.au = t - f;
  v.r0 = min(max(i, 0), G.nrad - 1);
  v.r1 = min(max(i + 1, 0), G.nrad - 1);
  v.iq = min(max(i, -1), G.nrad - 1);
  return v;
}

// Accumulates one view into n[], d[]. The fetcher returns the {ratio, calib}
// entries of this lane's two bins (r0, r1) of a plane: rows(i0) selects
// planes i0 .. i0 + ZB, pair(k, a0, a1) reads plane i0 + k, get2(ii, a0, a1)
// any plane.
template <int ZB, class F, class T>
__device__ __forceinline__ void accumulate_view(const LaneView& v, float fzb, float yr, int zb,
                                                int nzs, float fzm, float fzhi, F& fetch, T* n,
                                                T* d) {
  // Shared z axis. Slice j samples plane position w_j = fl(fl(P + zb + j) + 0.5)
  // with P = yr * fzb exact inside the fma. When s0 = fl(P + zb) >= 1 and
  // w_{ZB-1} lie in one binade [2^E, 2^(E+1)), every exact sum P + zb + j
  // (j >= 1) lies in that binade too and rounds on its grid (ulp <= 2^-16, so
  // integer shifts keep the nearest point and the tie parity): fl(P + zb + j) =
  // s0 + j, w_j = s0 + j + 0.5 and t_j = w_j - 0.5 = s0 + j exactly. Then all
  // slices share alpha = s0 - floor(s0) and take planes floor(s0) + j, j + 1.
  const bool slow = v.any && !(v.s0 >= 1.f && (__float_as_uint(v.s0) >> 23) ==
                                                   (__float_as_uint(v.wl) >> 23));
  if (__builtin_amdgcn_ballot_w64(slow) == 0) {
    const float f0 = floorf(v.s0);
    const float al = v.s0 - f0;
    const int i0 = (int)f0;
    float c[ZB + 1], q[ZB + 1];
    fetch.rows(i0);
#pragma unroll
    for (int k = 0; k <= ZB; k++) {
      float2 a0, a1;
      fetch.pair(k, a0, a1);
      c[k] = lerp1(a0.x, a1.x, v.au);
      q[k] = lerp1(a0.y, a1.y, v.au);
    }
#pragma unroll
    for (int j = 0; j < ZB; j++) {
      const float wj = v.w0 + (float)j;  // exact, see above
      const float A = lerp1(c[j], c[j + 1], al);
      const float B = lerp1(q[j], q[j + 1], al);
      // in range and not a gap (GFIRST_GPU_Kernels.cu:157)
      const bool ok =
          v.any && (zb + j < nzs) && wj >= fzm && wj < fzhi && !(B < 0.f) && !(A < 0.f);
      n[j] += ok ? (T)A : (T)0;
      d[j] += ok ? (T)B : (T)0;
    }
    return;
  }
  // General case: one plane position per slice, as the reference computes it.
  float al[ZB];
  int ii[ZB];
  bool inr[ZB];
#pragma unroll
  for (int j = 0; j < ZB; j++) {
    const float w = (j == 0) ? v.w0 : (j == ZB - 1) ? v.wl : fmaf(yr, fzb, (float)(zb + j)) + 0.5f;
    inr[j] = v.any && (zb + j < nzs) && w >= fzm && w < fzhi;
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
    float c[ZB + 1], q[ZB + 1];
    fetch.rows(ii[0]);
#pragma unroll
    for (int k = 0; k <= ZB; k++) {
      float2 a0, a1;
      fetch.pair(k, a0, a1);
      c[k] = lerp1(a0.x, a1.x, v.au);
      q[k] = lerp1(a0.y, a1.y, v.au);
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
        float2 w00, w01, w10, w11;
        fetch.get2(ii[j], w00, w01);
        fetch.get2(ii[j] + 1, w10, w11);
        av[j] = lerp1(lerp1(w00.x, w01.x, v.au), lerp1(w10.x, w11.x, v.au), al[j]);
        bv[j] = lerp1(lerp1(w00.y, w01.y, v.au), lerp1(w10.y, w11.y, v.au), al[j]);
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
    n[j] += (T)A;
    d[j] += (T)B;
  }
}

// Rows straight from the packed array: one 16-byte buffer load per plane
// gives both bins (32-bit offsets).
struct FetchGlobal {
  __amdgpu_buffer_rsrc_t rc;
  unsigned vrow, rowb, oq, rb0 = 0;
  __device__ __forceinline__ void rows(int i0) { rb0 = (vrow + (unsigned)i0) * rowb; }
  __device__ __forceinline__ void pair(int k, float2& a0, float2& a1) const {
    const float4 q = ld_f4(rc, rb0 + oq, (unsigned)k * rowb);
    a0 = make_float2(q.x, q.y);
    a1 = make_float2(q.z, q.w);
  }
  __device__ __forceinline__ void get2(int ii, float2& a0, float2& a1) const {
    const float4 q = ld_f4(rc, (vrow + (unsigned)ii) * rowb + oq, 0);
    a0 = make_float2(q.x, q.y);
    a1 = make_float2(q.z, q.w);
  }
};

// Rows from the block's LDS window (row stride kBW, entries of bins
// rlo - 1 ...); every index is clamped into the window, which only changes
// lanes outside the view.
template <int ZB>
struct FetchLds {
  const float2* t;
  int lo, hr, c0, c1;
  const float2 *p0 = nullptr, *p1 = nullptr;
  __device__ __forceinline__ void rows(int i0) {
    const int r = min(max(i0 - lo, 0), hr - ZB);
    p0 = t + r * kBW + c0;
    p1 = t + r * kBW + c1;
  }
  __device__ __forceinline__ void pair(int k, float2& a0, float2& a1) const {
    a0 = p0[k * kBW];
    a1 = p1[k * kBW];
  }
  __device__ __forceinline__ void get2(int ii, float2& a0, float2& a1) const {
    const int r = min(max(ii - lo, 0), hr);
    a0 = t[r * kBW + c0];
    a1 = t[r * kBW + c1];
  }
};

// Block-uniform window of one view: planes [lo, lo + nr) and bins
// [rlo, rlo + nb) cover every access of the block's lanes.
struct Window {
  int lo, nr, rlo, nb;
  bool used;  // some column and slice of the block can fall in the segment's planes
};

__device__ __forceinline__ Window view_window(const DevGeom& G, float x0, float x1, float y0,
                                              float y1, float st, float ct, float fzb, float fz0,
                                              float fzl, float fzm, float fzhi) {
  // yr = fma(yv, ct, -(xv*st)) and xr = fma(yv, st, xv*ct) are monotone in
  // xv and yv: extremes at the corners.
  const float ya = fmaf(y0, ct, -(x0 * st)), yb = fmaf(y0, ct, -(x1 * st));
  const float yc = fmaf(y1, ct, -(x0 * st)), yd = fmaf(y1, ct, -(x1 * st));
  const float ymn = fminf(fminf(ya, yb), fminf(yc, yd));
  const float ymx = fmaxf(fmaxf(ya, yb), fmaxf(yc, yd));
  const float pa = fmaf(ymn, fzb, fz0), pb = fmaf(ymx, fzb, fz0);
  const float qa = fmaf(ymn, fzb, fzl), qb = fmaf(ymx, fzb, fzl);
  const float smn = fminf(pa, pb), smx = fmaxf(qa, qb);
  const float wmn = smn + 0.5f, wmx = smx + 0.5f;
  Window wd;
  wd.used = wmx >= fzm && wmn < fzhi;
  const int lo = (int)floorf(wmn - 0.5f);
  const int hi = (int)floorf(wmx - 0.5f) + 1;
  wd.lo = lo;
  wd.nr = hi - lo + 1;
  const float xa = fmaf(y0, st, x0 * ct), xb = fmaf(y0, st, x1 * ct);
  const float xc = fmaf(y1, st, x0 * ct), xd = fmaf(y1, st, x1 * ct);
  const float umn = fmaf(fminf(fminf(xa, xb), fminf(xc, xd)), G.fxb, G.RADM) + 0.5f;
  const float umx = fmaf(fmaxf(fmaxf(xa, xb), fmaxf(xc, xd)), G.fxb, G.RADM) + 0.5f;
  int ra, rb, rc, rd;
  float dummy;
  axis(umn, G.nrad, ra, rb, dummy);
  axis(umx, G.nrad, rc, rd, dummy);
  wd.rlo = ra;
  wd.nb = rd - ra + 1;
  return wd;
}

template <int ZB>
__global__ __launch_bounds__(256) void k_back(DevGeom G, const float2* __restrict__ c2,
                                              unsigned c2_bytes, float* __restrict__ numer,
                                              float* __restrict__ denom, int subset, int nsub,
                                              const int* __restrict__ dsub, int rmax) {
  extern __shared__ float2 tile[];  // [2][rmax][kBW]
  __shared__ int4 wtab[256];        // windows of a chunk of 256 views
  __shared__ short wlist[256];      // chunk entries the block touches, in view order
  __shared__ int wcnt[4];
  if (dsub) subset = *dsub;
  const int nav = (G.nang - subset + nsub - 1) / nsub;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wv = tid >> 6;
  // blocks of one z-slab are consecutive logical ids, and each XCD walks a
  // contiguous range of logical ids
  const unsigned nbx = (G.res + 15) / 16, nxy = nbx * nbx;
  const unsigned b = xcd_remap(blockIdx.x, gridDim.x);
  const int bz = (int)(b / nxy), byx = (int)(b - (unsigned)bz * nxy);
  const int bx = byx % (int)nbx, by = byx / (int)nbx;
  const int ix = bx * 16 + (wv & 1) * 8 + (lane & 7);
  const int iy = by * 16 + (wv >> 1) * 8 + (lane >> 3);
  const int zb = bz * ZB;
  const bool act = ix < G.res && iy < G.res;
  const __amdgpu_buffer_rsrc_t rc2 = rsrc(c2, c2_bytes);
  const float xv = (float)ix - G.RESM, yv = (float)iy - G.RESM;
  // active corner columns of the block
  const float cx0 = (float)(bx * 16) - G.RESM, cx1 = (float)min(bx * 16 + 15, G.res - 1) - G.RESM;
  const float cy0 = (float)(by * 16) - G.RESM, cy1 = (float)min(by * 16 + 15, G.res - 1) - G.RESM;
  const unsigned rowb = (unsigned)packed_row_len(G.nrad) * 8u;
  const unsigned nrows = (unsigned)(G.nsinogs + 2 * G.nseg);
  const float fz0 = (float)zb, fzl = (float)(zb + ZB - 1);
  float n[ZB], d[ZB];
#pragma unroll
  for (int j = 0; j < ZB; j++) n[j] = d[j] = 0.f;
  // window loader: thread t copies column t % 32 of rows t / 32 + 8 k
  const int lc = tid & 31, lr = tid >> 5;
  constexpr int kPre = 6;  // rows per thread: rmax <= 48
  float2 pre[kPre];
  // The host launches this kernel only when its row bound fits rmax and the
  // bin bound fits kBinsMax; the clamps keep every LDS access inside its
  // buffer regardless.
  auto issue = [&](const int4 w) {
    const int lo = w.x, nrr = min(w.y & 0xffff, rmax), nbb = min(w.y >> 16, kBinsMax);
    const int si = w.w & 0xffff, a = w.w >> 16;
    const SegEntry se = G.seg[si];
    const unsigned vrow = (unsigned)a * nrows + (unsigned)(se.nza + 2 * si + 1 - se.zmin);
#pragma unroll
    for (int k = 0; k < kPre; k++) {
      const int r = lr + 8 * k;
      pre[k] = make_float2(0.f, 0.f);
      if (r < nrr && lc < nbb)
        pre[k] = ld_f2(rc2, (vrow + (unsigned)(lo + r)) * rowb + (unsigned)(w.z + 1 + lc) * 8u, 0);
    }
  };
  auto commit = [&](const int4 w, int buf) {
    const int nrr = min(w.y & 0xffff, rmax), nbb = min(w.y >> 16, kBinsMax);
    float2* t = tile + buf * rmax * kBW;
#pragma unroll
    for (int k = 0; k < kPre; k++) {
      const int r = lr + 8 * k;
      if (r < nrr && lc < nbb) t[r * kBW + lc] = pre[k];
    }
  };
  const int nviews = G.nseg * nav;  // view v: segment v / nav, angle v % nav
  for (int c0 = 0; c0 < nviews; c0 += 256) {
    __syncthreads();  // the previous chunk's tables and tiles are no longer read
    // windows of this chunk's views, one per thread, and the list of the
    // ones the block touches (ballot compaction keeps view order)
    bool used = false;
    {
      const int vi = c0 + tid;
      if (vi < nviews) {
        const int si = vi / nav, a = vi - si * nav;
        const SegEntry se = G.seg[si];
        const int it = subset + a * nsub;
        const Window w = view_window(G, cx0, cx1, cy0, cy1, G.sin_t[it], G.cos_t[it], se.fzb,
                                     fz0, fzl, (float)se.zmin, (float)(G.nzs - se.zmin));
        used = w.used;
        wtab[tid] = make_int4(w.lo, (w.nr & 0xffff) | (w.nb << 16), w.rlo, si | (a << 16));
      }
    }
    const unsigned long long m = __builtin_amdgcn_ballot_w64(used);
    if (lane == 0) wcnt[wv] = __popcll(m);
    __syncthreads();
    int off = 0;
    for (int k = 0; k < wv; k++) off += wcnt[k];
    const int total = wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
    if (used) {
      const unsigned lo32 = (unsigned)m, hi32 = (unsigned)(m >> 32);
      const int below = __builtin_amdgcn_mbcnt_hi(hi32, __builtin_amdgcn_mbcnt_lo(lo32, 0u));
      wlist[off + below] = (short)tid;
    }
    __syncthreads();
    if (total == 0) continue;
    int4 cw = wtab[wlist[0]];
    issue(cw);
    commit(cw, 0);
    __syncthreads();
    int buf = 0;
    for (int e = 0; e < total; e++) {
      const bool more = e + 1 < total;
      int4 nw = make_int4(0, 0, 0, 0);
      if (more) {
        nw = wtab[wlist[e + 1]];
        issue(nw);
      }
      const int si = cw.w & 0xffff, a = cw.w >> 16;
      const SegEntry se = G.seg[si];
      const int it = subset + a * nsub;
      const float fzm = (float)se.zmin, fzhi = (float)(G.nzs - se.zmin);
      const LaneView v = lane_view(G, xv, yv, G.sin_t[it], G.cos_t[it], se.fzb, fz0, fzl, fzm,
                                   fzhi, act);
      if (__builtin_amdgcn_ballot_w64(v.any) != 0) {
        const int hc = min(cw.y >> 16, kBinsMax) - 1;
        FetchLds<ZB> f;
        f.t = tile + buf * rmax * kBW;
        f.lo = cw.x;
        f.hr = min(cw.y & 0xffff, rmax) - 1;
        f.c0 = min(max(v.r0 - cw.z, 0), hc);
        f.c1 = min(max(v.r1 - cw.z, 0), hc);
        accumulate_view<ZB>(v, se.fzb, v.yr, zb, G.nzs, fzm, fzhi, f, n, d);
      }
      if (more) commit(nw, buf ^ 1);
      __syncthreads();
      buf ^= 1;
      cw = nw;
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

// Same sums with every read from global memory: geometries whose window does
// not fit the LDS bounds.
// MINW > 0 asks the compiler for a register budget that keeps MINW waves per
// SIMD resident (an autotuner candidate; more resident waves can also thrash
// the L1). A block is BW waves of 8x8 columns: 4 (a 16x16 tile) or 1 (8x8,
// four times the blocks for small images); each voxel's sum is the same.
template <int ZB, int MINW, class T = float, int BW = 4>
__global__ __launch_bounds__(64 * BW, MINW > 0 ? MINW : 1) void k_back_direct(
    DevGeom G, const float2* __restrict__ c2, unsigned c2_bytes, T* __restrict__ numer,
    T* __restrict__ denom, int subset, int nsub, const int* __restrict__ dsub) {
  static_assert(BW == 1 || BW == 4, "blocks of 1 or 4 waves");
  constexpr int kTile = BW == 4 ? 16 : 8;
  if (dsub) subset = *dsub;
  const int nav = (G.nang - subset + nsub - 1) / nsub;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int bx, by, bz;
  xcd_region(bx, by, bz);
  const int ix = bx * kTile + (wv & 1) * 8 + (lane & 7);
  const int iy = by * kTile + (wv >> 1) * 8 + (lane >> 3);
  const int zb = bz * ZB;
  const bool act = ix < G.res && iy < G.res;
  const __amdgpu_buffer_rsrc_t rc2 = rsrc(c2, c2_bytes);
  const float xv = (float)ix - G.RESM, yv = (float)iy - G.RESM;
  const unsigned rowb = (unsigned)packed_row_len(G.nrad) * 8u;
  const unsigned nr = (unsigned)(G.nsinogs + 2 * G.nseg);
  const float fz0 = (float)zb, fzl = (float)(zb + ZB - 1);
  T n[ZB], d[ZB];
#pragma unroll
  for (int j = 0; j < ZB; j++) n[j] = d[j] = (T)0;
  for (int si = 0; si < G.nseg; si++) {
    const SegEntry se = G.seg[si];
    const float fzm = (float)se.zmin, fzhi = (float)(G.nzs - se.zmin);
    const unsigned segrow = (unsigned)(se.nza + 2 * si + 1 - se.zmin);
    for (int a = 0; a < nav; a++) {
      const int it = subset + a * nsub;
      const LaneView v = lane_view(G, xv, yv, G.sin_t[it], G.cos_t[it], se.fzb, fz0, fzl, fzm,
                                   fzhi, act);
      if (__builtin_amdgcn_ballot_w64(v.any) == 0) continue;
      FetchGlobal f;
      f.rc = rc2;
      f.vrow = (unsigned)a * nr + segrow;
      f.rowb = rowb;
      f.oq = (unsigned)(v.iq + 1) * 8u;
      accumulate_view<ZB>(v, se.fzb, v.yr, zb, G.nzs, fzm, fzhi, f, n, d);
    }
  }
  if (!act) return;
#pragma unroll
  for (int j = 0; j < ZB; j++) {
    const int iz = zb + j;
    if (iz < G.nzs) {
      const size_t iV = ((size_t)iz * G.res + iy) * G.res + ix;
      if (numer) numer[iV] = n[j];
      denom[iV] = d[j];
    }
  }
}

// ---------------------------------------------------------------------------
// Multi-frame backprojector: F frames (2 or 3) of one geometry, calibration
// mask and subset in one launch. Their packed rows are interleaved per entry
// as {ratio_0, ratio_1, ratio_2, calib} (store_packed4), so one 16-byte load
// gives a bin of every frame, and each lane's view geometry (rotation, radial
// and plane positions, the shared-z test, the skip ballot) and the
// calibration interpolation are computed once for the group. Per frame the
// sums are those k_back_direct forms from the frame's own {ratio, calib}
// rows: the same interpolants from the same values, the same gap rule (a
// negative interpolated ratio or calibration drops the view for that frame
// only) and the same view order, so the bits are identical.
// ---------------------------------------------------------------------------
struct FetchGlobal4 {
  __amdgpu_buffer_rsrc_t rc;
  unsigned vrow, rowb, oq, rb0 = 0;
  __device__ __forceinline__ void rows(int i0) { rb0 = (vrow + (unsigned)i0) * rowb; }
  __device__ __forceinline__ void pair(int k, float4& a0, float4& a1) const {
    a0 = ld_f4(rc, rb0 + oq, (unsigned)k * rowb);
    a1 = ld_f4(rc, rb0 + oq + 16u, (unsigned)k * rowb);
  }
  __device__ __forceinline__ void get2(int ii, float4& a0, float4& a1) const {
    const unsigned o = (vrow + (unsigned)ii) * rowb + oq;
    a0 = ld_f4(rc, o, 0);
    a1 = ld_f4(rc, o + 16u, 0);
  }
};

__device__ __forceinline__ float chan_of(const float4& v, int f) {
  return f == 0 ? v.x : (f == 1 ? v.y : v.z);
}

template <int ZB, int F>
__device__ __forceinline__ void accumulate_view_multi(const LaneView& v, float fzb, float yr,
                                                      int zb, int nzs, float fzm, float fzhi,
                                                      FetchGlobal4& fetch, float (&n)[F][ZB],
                                                      float (&d)[F][ZB]) {
  // shared z axis: the argument of accumulate_view
  const bool slow = v.any && !(v.s0 >= 1.f && (__float_as_uint(v.s0) >> 23) ==
                                                   (__float_as_uint(v.wl) >> 23));
  if (__builtin_amdgcn_ballot_w64(slow) == 0) {
    const float f0 = floorf(v.s0);
    const float al = v.s0 - f0;
    const int i0 = (int)f0;
    float c[F][ZB + 1], q[ZB + 1];
    fetch.rows(i0);
#pragma unroll
    for (int k = 0; k <= ZB; k++) {
      float4 a0, a1;
      fetch.pair(k, a0, a1);
      q[k] = lerp1(a0.w, a1.w, v.au);
#pragma unroll
      for (int f = 0; f < F; f++) c[f][k] = lerp1(chan_of(a0, f), chan_of(a1, f), v.au);
    }
#pragma unroll
    for (int j = 0; j < ZB; j++) {
      const float wj = v.w0 + (float)j;
      const float B = lerp1(q[j], q[j + 1], al);
      const bool base = v.any && (zb + j < nzs) && wj >= fzm && wj < fzhi && !(B < 0.f);
#pragma unroll
      for (int f = 0; f < F; f++) {
        const float A = lerp1(c[f][j], c[f][j + 1], al);
        const bool ok = base && !(A < 0.f);
        n[f][j] += ok ? A : 0.f;
        d[f][j] += ok ? B : 0.f;
      }
    }
    return;
  }
  float al[ZB];
  int ii[ZB];
  bool inr[ZB];
#pragma unroll
  for (int j = 0; j < ZB; j++) {
    const float w = (j == 0) ? v.w0 : (j == ZB - 1) ? v.wl : fmaf(yr, fzb, (float)(zb + j)) + 0.5f;
    inr[j] = v.any && (zb + j < nzs) && w >= fzm && w < fzhi;
    const float t = w - 0.5f;
    const float fl = floorf(t);
    al[j] = t - fl;
    ii[j] = (int)fl;
  }
  bool reg = true;
#pragma unroll
  for (int j = 1; j < ZB; j++) reg = reg && (!inr[j] || ii[j] == ii[0] + j);
  float av[F][ZB], bv[ZB];
  if (reg) {
    float c[F][ZB + 1], q[ZB + 1];
    fetch.rows(ii[0]);
#pragma unroll
    for (int k = 0; k <= ZB; k++) {
      float4 a0, a1;
      fetch.pair(k, a0, a1);
      q[k] = lerp1(a0.w, a1.w, v.au);
#pragma unroll
      for (int f = 0; f < F; f++) c[f][k] = lerp1(chan_of(a0, f), chan_of(a1, f), v.au);
    }
#pragma unroll
    for (int j = 0; j < ZB; j++) {
      bv[j] = inr[j] ? lerp1(q[j], q[j + 1], al[j]) : 0.f;
#pragma unroll
      for (int f = 0; f < F; f++) av[f][j] = inr[j] ? lerp1(c[f][j], c[f][j + 1], al[j]) : 0.f;
    }
  } else {
#pragma unroll
    for (int j = 0; j < ZB; j++) {
      bv[j] = 0.f;
#pragma unroll
      for (int f = 0; f < F; f++) av[f][j] = 0.f;
      if (inr[j]) {
        float4 w00, w01, w10, w11;
        fetch.get2(ii[j], w00, w01);
        fetch.get2(ii[j] + 1, w10, w11);
        bv[j] = lerp1(lerp1(w00.w, w01.w, v.au), lerp1(w10.w, w11.w, v.au), al[j]);
#pragma unroll
        for (int f = 0; f < F; f++)
          av[f][j] = lerp1(lerp1(chan_of(w00, f), chan_of(w01, f), v.au),
                           lerp1(chan_of(w10, f), chan_of(w11, f), v.au), al[j]);
      }
    }
  }
#pragma unroll
  for (int j = 0; j < ZB; j++) {
#pragma unroll
    for (int f = 0; f < F; f++) {
      const bool gap = bv[j] < 0.f || av[f][j] < 0.f;  // GFIRST_GPU_Kernels.cu:157
      n[f][j] += gap ? 0.f : av[f][j];
      d[f][j] += gap ? 0.f : bv[j];
    }
  }
}

struct MultiOut {
  float* numer[kMaxGroup];
  float* denom[kMaxGroup];
};

template <int ZB, int F>
__global__ __launch_bounds__(256) void k_back_multi(DevGeom G, const float* __restrict__ c4,
                                                    unsigned c4_bytes, MultiOut o, int subset,
                                                    int nsub, const int* __restrict__ dsub) {
  if (dsub) subset = *dsub;
  const int nav = (G.nang - subset + nsub - 1) / nsub;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int bx, by, bz;
  xcd_region(bx, by, bz);
  const int ix = bx * 16 + (wv & 1) * 8 + (lane & 7);
  const int iy = by * 16 + (wv >> 1) * 8 + (lane >> 3);
  const int zb = bz * ZB;
  const bool act = ix < G.res && iy < G.res;
  const __amdgpu_buffer_rsrc_t rc4 = rsrc(c4, c4_bytes);
  const float xv = (float)ix - G.RESM, yv = (float)iy - G.RESM;
  const unsigned rowb = (unsigned)packed_row_len(G.nrad) * 16u;
  const unsigned nr = (unsigned)(G.nsinogs + 2 * G.nseg);
  const float fz0 = (float)zb, fzl = (float)(zb + ZB - 1);
  float n[F][ZB], d[F][ZB];
#pragma unroll
  for (int f = 0; f < F; f++)
#pragma unroll
    for (int j = 0; j < ZB; j++) n[f][j] = d[f][j] = 0.f;
  for (int si = 0; si < G.nseg; si++) {
    const SegEntry se = G.seg[si];
    const float fzm = (float)se.zmin, fzhi = (float)(G.nzs - se.zmin);
    const unsigned segrow = (unsigned)(se.nza + 2 * si + 1 - se.zmin);
    for (int a = 0; a < nav; a++) {
      const int it = subset + a * nsub;
      const LaneView v = lane_view(G, xv, yv, G.sin_t[it], G.cos_t[it], se.fzb, fz0, fzl, fzm,
                                   fzhi, act);
      if (__builtin_amdgcn_ballot_w64(v.any) == 0) continue;
      FetchGlobal4 fe;
      fe.rc = rc4;
      fe.vrow = (unsigned)a * nr + segrow;
      fe.rowb = rowb;
      fe.oq = (unsigned)(v.iq + 1) * 16u;
      accumulate_view_multi<ZB, F>(v, se.fzb, v.yr, zb, G.nzs, fzm, fzhi, fe, n, d);
    }
  }
  if (!act) return;
#pragma unroll
  for (int j = 0; j < ZB; j++) {
    const int iz = zb + j;
    if (iz < G.nzs) {
      const size_t iV = ((size_t)iz * G.res + iy) * G.res + ix;
#pragma unroll
      for (int f = 0; f < F; f++) {
        o.numer[f][iV] = n[f][j];
        o.denom[f][iV] = d[f][j];
      }
    }
  }
}

// One slice per thread for small images, where the grid is a few hundred
// waves and each wave's walk over the views is a chain of memory round trips:
// the rows of NA consecutive views are all requested before the first of them
// is summed, so the chain waits once per NA views. With one slice the
// per-slice formula of accumulate_view (plane position w0, its floor and
// fraction, rows ii and ii + 1) is the whole computation, and the views are
// summed in the same order, so the sums are those of k_back_direct<1>.
template <int NA, int BW>
__global__ __launch_bounds__(64 * BW) void k_back_ahead(DevGeom G, const float2* __restrict__ c2,
                                                        unsigned c2_bytes, float* __restrict__ numer,
                                                        float* __restrict__ denom, int subset,
                                                        int nsub, const int* __restrict__ dsub) {
  static_assert(BW == 1 || BW == 4, "blocks of 1 or 4 waves");
  constexpr int kTile = BW == 4 ? 16 : 8;
  if (dsub) subset = *dsub;
  const int nav = (G.nang - subset + nsub - 1) / nsub;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int ix = blockIdx.x * kTile + (wv & 1) * 8 + (lane & 7);
  const int iy = blockIdx.y * kTile + (wv >> 1) * 8 + (lane >> 3);
  const int zb = blockIdx.z;
  const bool act = ix < G.res && iy < G.res;
  const __amdgpu_buffer_rsrc_t rc2 = rsrc(c2, c2_bytes);
  const float xv = (float)ix - G.RESM, yv = (float)iy - G.RESM;
  const unsigned rowb = (unsigned)packed_row_len(G.nrad) * 8u;
  const unsigned nr = (unsigned)(G.nsinogs + 2 * G.nseg);
  const float fz0 = (float)zb;
  float n = 0.f, d = 0.f;
  for (int si = 0; si < G.nseg; si++) {
    const SegEntry se = G.seg[si];
    const float fzm = (float)se.zmin, fzhi = (float)(G.nzs - se.zmin);
    const unsigned segrow = (unsigned)(se.nza + 2 * si + 1 - se.zmin);
    for (int a0 = 0; a0 < nav; a0 += NA) {
      float4 p0[NA], p1[NA];
      float au[NA], al[NA];
      bool inr[NA], used[NA];
#pragma unroll
      for (int u = 0; u < NA; u++) {
        const int a = a0 + u;
        used[u] = false;
        inr[u] = false;
        au[u] = al[u] = 0.f;
        if (a < nav) {
          const int it = subset + a * nsub;
          const LaneView v = lane_view(G, xv, yv, G.sin_t[it], G.cos_t[it], se.fzb, fz0, fz0,
                                       fzm, fzhi, act);
          used[u] = __builtin_amdgcn_ballot_w64(v.any) != 0;
          if (used[u]) {
            const float w = v.w0;
            inr[u] = v.any && (zb < G.nzs) && w >= fzm && w < fzhi;
            const float t = w - 0.5f;
            const float f = floorf(t);
            al[u] = t - f;
            au[u] = v.au;
            const unsigned rb = ((unsigned)a * nr + segrow + (unsigned)(int)f) * rowb +
                                (unsigned)(v.iq + 1) * 8u;
            p0[u] = ld_f4(rc2, rb, 0);
            p1[u] = ld_f4(rc2, rb, rowb);
          }
        }
      }
#pragma unroll
      for (int u = 0; u < NA; u++) {
        if (!used[u]) continue;  // wave-uniform: the view misses the wave
        const float c0 = lerp1(p0[u].x, p0[u].z, au[u]), c1 = lerp1(p1[u].x, p1[u].z, au[u]);
        const float q0 = lerp1(p0[u].y, p0[u].w, au[u]), q1 = lerp1(p1[u].y, p1[u].w, au[u]);
        float A = inr[u] ? lerp1(c0, c1, al[u]) : 0.f;
        float B = inr[u] ? lerp1(q0, q1, al[u]) : 0.f;
        if (B < 0.f || A < 0.f) {  // gaps (GFIRST_GPU_Kernels.cu:157)
          A = 0.f;
          B = 0.f;
        }
        n += A;
        d += B;
      }
    }
  }
  if (!act || zb >= G.nzs) return;
  const size_t iV = ((size_t)zb * G.res + iy) * G.res + ix;
  numer[iV] = n;
  denom[iV] = d;
}

inline int nav_of(const DevGeom& G, int subset, int nsub) {
  return (G.nang - subset + nsub - 1) / nsub;
}

}  // namespace

size_t rotate_floats_per_angle(const DevGeom& G) {
  return (size_t)(G.l1 - G.l0 + 1) * G.nzs * G.nrad;
}

// z-sum blocks (= log-likelihood partials) per pair of angles, and in all
int zsum_blocks_per_pair(const DevGeom& G, int sg) {
  return ((G.nrad + 127) / 128) * ((G.nsinogs + 4 * sg - 1) / (4 * sg));
}

int forward_llh_blocks(const DevGeom& G, int subset, int nsub, int sg) {
  return zsum_blocks_per_pair(G, sg) * ((nav_of(G, subset, nsub) + 1) / 2);
}

namespace {

template <int SG>
hipError_t zsum_launch(const DevGeom& G, const float* R, unsigned rbytes, const FwdArgs& o,
                       bool ratio, int storage, int subset, int nsub, const int* dsub, int a0,
                       int na, int nl, hipStream_t st) {
  // a0 is even (chunks hold whole pairs of angles)
  const unsigned per = (unsigned)zsum_blocks_per_pair(G, SG);
  const unsigned gz = per * (unsigned)((na + 1) / 2);
  const unsigned off = per * (unsigned)(a0 / 2);
  if (!ratio) {
    hipLaunchKernelGGL((k_zsum<GFR_STORE_F32, false, false, SG>), dim3(gz), dim3(256), 0, st, G,
                       R, rbytes, o, subset, nsub, dsub, a0, na, nl, off);
    return hipGetLastError();
  }
#define GFR_ZS(S)                                                                             \
  if (o.llh_part)                                                                             \
    hipLaunchKernelGGL((k_zsum<S, true, true, SG>), dim3(gz), dim3(256), 0, st, G, R, rbytes, \
                       o, subset, nsub, dsub, a0, na, nl, off);                                   \
  else                                                                                        \
    hipLaunchKernelGGL((k_zsum<S, true, false, SG>), dim3(gz), dim3(256), 0, st, G, R,       \
                       rbytes, o, subset, nsub, dsub, a0, na, nl, off);
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
__global__ void k_llh_sum(const double* __restrict__ part, int per_pair, int nang, int subset,
                          int nsub, const int* __restrict__ dsub, double* __restrict__ out) {
  if (dsub) subset = *dsub;
  const int n = per_pair * (((nang - subset + nsub - 1) / nsub + 1) / 2);
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
  if (chunk < nav && chunk > 1) chunk &= ~1;  // the z-sum takes angles in pairs
  if (chunk < 1) return hipErrorInvalidValue;
  // log-likelihood partials are indexed by pair: chunks must hold whole pairs
  if (out.llh_part && chunk < nav && (chunk & 1)) return hipErrorInvalidValue;
  if (out.corr4 && (out.chan < 0 || out.chan >= kMaxGroup)) return hipErrorInvalidValue;
  FwdArgs o{out.proj, out.corr,  out.corr2,  out.corr4,   out.chan,
            out.mask, out.y,     out.bg,     out.yscale,  out.llh_part};
  const bool ratio = out.corr || out.corr2 || out.corr4;
  for (int a0 = 0; a0 < nav; a0 += chunk) {
    const int na = min(chunk, nav - a0);
    const unsigned rbytes = (unsigned)((size_t)na * per * sizeof(float));
    const unsigned gxy = ((G.nrad + 31) / 32) * ((nl + 31) / 32) * na;
    // plane chunks outermost when an angle pair's rotation exceeds an XCD's
    // 4 MiB L2 (the z-sum cannot find it there anyway)
    const int zouter = 2 * per * sizeof(float) > ((size_t)4 << 20);
    if (tune.rot_z == 4)
      hipLaunchKernelGGL(k_rotate_tile<4>, dim3(gxy * ((G.nzs + 3) / 4)), dim3(256), 0, st, G, img,
                         R, subset, nsub, dsub, a0, nl, zouter);
    else if (tune.rot_z == 32)
      hipLaunchKernelGGL(k_rotate_tile<32>, dim3(gxy * ((G.nzs + 31) / 32)), dim3(256), 0, st, G,
                         img, R, subset, nsub, dsub, a0, nl, zouter);
    else
      hipLaunchKernelGGL(k_rotate_tile<16>, dim3(gxy * ((G.nzs + 15) / 16)), dim3(256), 0, st, G,
                         img, R, subset, nsub, dsub, a0, nl, zouter);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    switch (tune.zsum_sg) {
      case 4: e = zsum_launch<4>(G, R, rbytes, o, ratio, out.storage, subset, nsub, dsub, a0, na, nl, st); break;
      case 6: e = zsum_launch<6>(G, R, rbytes, o, ratio, out.storage, subset, nsub, dsub, a0, na, nl, st); break;
      default: e = zsum_launch<8>(G, R, rbytes, o, ratio, out.storage, subset, nsub, dsub, a0, na, nl, st); break;
    }
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_llh_sum(const DevGeom& G, const double* part, int subset, int nsub,
                          const int* dsub, int sg, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_llh_sum, dim3(1), dim3(256), 0, st, part, zsum_blocks_per_pair(G, sg),
                     G.nang, subset, nsub, dsub, out);
  return hipGetLastError();
}

// Rows of the backprojector's LDS window: ZB slices + 1 for the second
// interpolation plane, the tilt of a 16x16 column tile (15*sqrt2 columns
// times the largest |fzb|) and 2 for rounding; at most 48 (six rows per
// loading thread), else 0 and the direct kernel reads global memory.
int back_window_rows(const DevGeom& G, int zb) {
  const int r = zb + 3 + (int)ceilf(21.3f * G.fzb_absmax);
  return r <= 48 ? r : 0;  // 0: does not fit, read global memory directly
}

size_t packed_bytes(const DevGeom& G, int subset, int nsub) {
  return sizeof(float2) * (size_t)nav_of(G, subset, nsub) * (G.nsinogs + 2 * (size_t)G.nseg) *
         packed_row_len(G.nrad);
}

size_t packed4_bytes(const DevGeom& G, int subset, int nsub) {
  return 2 * packed_bytes(G, subset, nsub);
}

hipError_t launch_back_multi(const DevGeom& G, const float* corr4, float* const* numer,
                             float* const* denom, int nf, int subset, int nsub, const int* dsub,
                             int zb, hipStream_t st) {
  const size_t bytes = packed4_bytes(G, dsub ? 0 : subset, nsub);
  if (bytes >= 0xFFFFFFF0u || nf < 2 || nf > kMaxGroup) return hipErrorInvalidValue;
  MultiOut o{};
  for (int f = 0; f < nf; f++) {
    if (!numer[f] || !denom[f]) return hipErrorInvalidValue;
    o.numer[f] = numer[f];
    o.denom[f] = denom[f];
  }
  const unsigned nbx = (G.res + 15) / 16;
#define GFR_BM(ZB, F)                                                                          \
  hipLaunchKernelGGL((k_back_multi<ZB, F>), dim3(nbx, nbx, (G.nzs + ZB - 1) / ZB), dim3(256), \
                     0, st, G, corr4, (unsigned)bytes, o, subset, nsub, dsub)
  if (zb == 6) {
    if (nf == 3) GFR_BM(6, 3); else GFR_BM(6, 2);
  } else {
    if (nf == 3) GFR_BM(4, 3); else GFR_BM(4, 2);
  }
#undef GFR_BM
  return hipGetLastError();
}

hipError_t launch_pack(const DevGeom& G, const float* corr, const uint8_t* mask, float2* c2,
                       int subset, int nsub, hipStream_t st) {
  dim3 grid((G.nrad + kWave - 1) / kWave, (G.nsinogs + 3) / 4, nav_of(G, subset, nsub));
  hipLaunchKernelGGL(k_pack, grid, dim3(kWave, 4), 0, st, G, corr, mask, c2, subset, nsub);
  return hipGetLastError();
}

hipError_t launch_sensitivity(const DevGeom& G, const float2* corr2, double* sens, int subset,
                              int nsub, hipStream_t st) {
  const size_t bytes = packed_bytes(G, subset, nsub);
  if (bytes >= 0xFFFFFFF0u) return hipErrorInvalidValue;
  const unsigned nbx = (G.res + 15) / 16;
  hipLaunchKernelGGL((k_back_direct<4, 0, double>), dim3(nbx, nbx, (G.nzs + 3) / 4), dim3(256), 0,
                     st, G, corr2, (unsigned)bytes, (double*)nullptr, sens, subset, nsub,
                     (const int*)nullptr);
  return hipGetLastError();
}

hipError_t launch_back_packed(const DevGeom& G, const float2* corr2, float* numer, float* denom,
                              int subset, int nsub, const int* dsub, const Tune& tune,
                              hipStream_t st) {
  const size_t bytes = packed_bytes(G, dsub ? 0 : subset, nsub);
  if (bytes >= 0xFFFFFFF0u) return hipErrorInvalidValue;
  const unsigned nbx = (G.res + 15) / 16;
  const bool bins_fit = 21.3f * G.fxb + 3.f <= (float)kBinsMax;
#define GFR_BK(ZB, W)                                                                           \
  {                                                                                             \
    const int rmax = back_window_rows(G, ZB);                                                   \
    if (tune.back_lds && bins_fit && rmax > 0) {                                                \
      const unsigned nb = nbx * nbx * (unsigned)((G.nzs + ZB - 1) / ZB);                        \
      hipLaunchKernelGGL(k_back<ZB>, dim3(nb), dim3(256), sizeof(float2) * 2 * rmax * kBW, st,  \
                         G, corr2, (unsigned)bytes, numer, denom, subset, nsub, dsub, rmax);    \
    } else if (tune.back_bw == 1) {                                                             \
      const unsigned n8 = (G.res + 7) / 8;                                                      \
      hipLaunchKernelGGL((k_back_direct<ZB, W, float, 1>), dim3(n8, n8, (G.nzs + ZB - 1) / ZB), \
                         dim3(64), 0, st, G, corr2, (unsigned)bytes, numer, denom, subset,      \
                         nsub, dsub);                                                           \
    } else {                                                                                    \
      hipLaunchKernelGGL((k_back_direct<ZB, W>), dim3(nbx, nbx, (G.nzs + ZB - 1) / ZB),         \
                         dim3(256), 0, st, G, corr2, (unsigned)bytes, numer, denom, subset,     \
                         nsub, dsub);                                                           \
    }                                                                                           \
  }
  if (tune.back_ahead > 1 && tune.back_zb == 1 && !tune.back_lds && tune.back_waves == 0) {
    const unsigned n8 = (G.res + 7) / 8;
#define GFR_BA(NA, BW)                                                                            \
  hipLaunchKernelGGL((k_back_ahead<NA, BW>), BW == 4 ? dim3(nbx, nbx, G.nzs) : dim3(n8, n8, G.nzs), \
                     dim3(64 * BW), 0, st, G, corr2, (unsigned)bytes, numer, denom, subset, nsub, dsub)
    if (tune.back_ahead == 4) {
      if (tune.back_bw == 1) GFR_BA(4, 1); else GFR_BA(4, 4);
    } else {
      if (tune.back_bw == 1) GFR_BA(2, 1); else GFR_BA(2, 4);
    }
#undef GFR_BA
    return hipGetLastError();
  }
  switch (tune.back_zb) {
    case 1: GFR_BK(1, 0); break;
    case 2: GFR_BK(2, 0); break;
    case 4: GFR_BK(4, 0); break;
    case 5: GFR_BK(5, 0); break;
    case 6: GFR_BK(6, 0); break;
    case 16: GFR_BK(16, 0); break;
    default:
      if (tune.back_waves == 6) {
        GFR_BK(8, 6);
      } else {
        GFR_BK(8, 0);
      }
      break;
  }
#undef GFR_BK
  return hipGetLastError();
}

int llh_capacity(const GeomHost& h, int nsub) {
  const int nav = (h.g.nang + nsub - 1) / nsub;
  return ((h.g.nrad + 127) / 128) * ((h.info.nsinogs + 15) / 16) * ((nav + 1) / 2);  // zsum_sg = 4
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
  const int cand[3] = {4, 6, 8};
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
  const int rcand[3] = {4, 16, 32};  // planes per rotation block
  for (int c : rcand) {
    Tune tt = t;
    tt.rot_z = c;
    const float ms = time_it([&] {
      return launch_forward(G, img, R, r_floats, fo, 0, nsub, nullptr, tt, st);
    });
    if (ms < bz) {
      bz = ms;
      t.rot_z = c;
    }
  }
  // backprojector: (kind, slices, register budget, waves per block)
  // (1 and 2 slices: small images, where each wave's sequential walk over
  // the views, not the device's throughput, sets the time)
  // (views whose rows are loaded ahead: 1-slice shapes only)
  const int bcand[18][5] = {{0, 4, 0, 4, 1}, {0, 5, 0, 4, 1}, {0, 6, 0, 4, 1}, {0, 8, 0, 4, 1},
                            {0, 8, 6, 4, 1}, {0, 16, 0, 4, 1}, {0, 4, 0, 1, 1}, {0, 6, 0, 1, 1},
                            {0, 8, 0, 1, 1}, {1, 4, 0, 4, 1},  {1, 8, 0, 4, 1}, {1, 16, 0, 4, 1},
                            {0, 1, 0, 1, 1}, {0, 2, 0, 1, 1},  {0, 2, 0, 4, 1}, {0, 1, 0, 1, 2},
                            {0, 1, 0, 1, 4}, {0, 1, 0, 4, 4}};
  for (const auto& c : bcand) {
    Tune tt = t;
    tt.back_lds = c[0];
    tt.back_zb = c[1];
    tt.back_waves = c[2];
    tt.back_bw = c[3];
    tt.back_ahead = c[4];
    const float ms = time_it([&] {
      return launch_back_packed(G, corr2, num, den, 0, nsub, nullptr, tt, st);
    });
    if (ms < bb) {
      bb = ms;
      t.back_lds = tt.back_lds;
      t.back_zb = tt.back_zb;
      t.back_waves = tt.back_waves;
      t.back_bw = tt.back_bw;
      t.back_ahead = tt.back_ahead;
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
