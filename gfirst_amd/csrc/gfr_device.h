// gfr_device.h — device helpers shared by the gfx950 kernels.
//
// Rounding model (shared with the CPU oracle so results compare bit for bit):
// every kernel file is compiled with -ffp-contract=off; the only fused
// multiply-adds are explicit fmaf() calls. Interpolation weights are exact
// FP32: lerp(a,b,t) = fma(t, b-a, a), x first, then y, then z.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gfr_internal.h"

namespace gfr {

constexpr int kWave = 64;

__device__ __forceinline__ float lerp1(float a, float b, float t) { return fmaf(t, b - a, a); }

// Raw buffer resource over [p, p + bytes): 32-bit offsets, loads past the end
// return 0 without touching memory.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc(const void* p, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes,
                                           0x00020000);
}
__device__ __forceinline__ float ld_f1(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
__device__ __forceinline__ float2 ld_f2(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
  return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
__device__ __forceinline__ float4 ld_f4(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
  return __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
}

// Unnormalized coordinate -> (i0, i1, alpha) with clamp-to-edge addressing:
// texel i covers [i, i+1), so the filter blends floor(c-0.5) and the next one
// (the reference's textures use linear filtering + clamp,
// GFIRST_GPU_Kernels.cu:260-263, 270-279).
__device__ __forceinline__ void axis(float c, int n, int& i0, int& i1, float& a) {
  const float t = c - 0.5f;
  const float f = floorf(t);
  const int i = (int)f;
  a = t - f;
  i0 = min(max(i, 0), n - 1);
  i1 = min(max(i + 1, 0), n - 1);
}

template <int S>
__device__ __forceinline__ float load_store(const void* p, size_t i) {
  if constexpr (S == GFR_STORE_F32) {
    return ((const float*)p)[i];
  } else if constexpr (S == GFR_STORE_F16) {
    return (float)((const _Float16*)p)[i];
  } else {
    const uint32_t b = (uint32_t)((const uint16_t*)p)[i] << 16;
    return __uint_as_float(b);
  }
}

// Work-groups of a 1-D grid are dispatched round-robin over the 8 XCDs (each
// with its own L2). This maps the hardware block id to a logical id so that
// every XCD walks one contiguous range of logical blocks (bijective for any
// grid size).
__device__ __forceinline__ unsigned xcd_remap(unsigned b, unsigned n) {
  const unsigned nx = n >> 3, rem = n & 7u, x = b & 7u, k = b >> 3;
  return x < rem ? x * (nx + 1) + k : rem * (nx + 1) + (x - rem) * nx + k;
}

// The same round-robin for a 3-D grid of (x tile, y tile, z slab) blocks,
// linearised x fastest. With an even x count and a y count divisible by 4,
// XCD k takes every z slab of the tiles of one (nx/2) x (ny/4) region
// (x half k & 1, y quarter k >> 1), so its L2 holds only the rows that
// region projects to; other grids keep the hardware order. Bijective.
__device__ __forceinline__ void xcd_region(int& bx, int& by, int& bz) {
  bx = blockIdx.x;
  by = blockIdx.y;
  bz = blockIdx.z;
  if ((gridDim.x & 1u) == 0 && (gridDim.y & 3u) == 0) {
    const unsigned b = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    const unsigned k = b & 7u, j = b >> 3;
    const unsigned hx = gridDim.x >> 1, hy = gridDim.y >> 2, t = j % (hx * hy);
    bz = (int)(j / (hx * hy));
    bx = (int)((k & 1u) * hx + t % hx);
    by = (int)((k >> 1) * hy + t / hx);
  }
}

// Packed {ratio, calib} rows of a subset for the backprojector. Per angle of
// the subset and per segment, the segment's planes zmin..NZS-zmin-1 are
// preceded and followed by one extra row, so that every plane a voxel can
// blend (zmin-1 .. NZS-zmin) is addressable without a test:
//  - segments with zmin > 0: the extra rows hold -1000, the value the
//    reference stores outside the segment's planes (GFIRST.cu:447);
//  - the direct segment (zmin = 0): the extra rows repeat planes 0 and NZS-1,
//    which is what the clamp-to-edge filter reads for planes -1 and NZS.
// Each row holds NRAD + 2 entries: bin r at r + 1, and copies of bins 0 and
// NRAD-1 at both ends, so the pair of bins a radial position blends under
// clamp-to-edge, (clamp(i), clamp(i + 1)), is the two adjacent entries at
// clamp(i, -1, NRAD-1) + 1: one 16-byte load.
__host__ __device__ __forceinline__ size_t packed_row_len(int nrad) { return (size_t)nrad + 2; }
__device__ __forceinline__ size_t packed_rows_per_angle(const DevGeom& G) {
  return (size_t)G.nsinogs + 2 * (size_t)G.nseg;
}
__device__ __forceinline__ void store_packed(const DevGeom& G, float2* c2, int ag, int s, int r,
                                             float2 v) {
  const int2 sp = G.sino[s];
  const SegEntry se = G.seg[sp.x];
  const size_t row0 = (size_t)ag * packed_rows_per_angle(G) + se.nza + 2 * (size_t)sp.x;
  const int k = sp.y - se.zmin;
  const int nz = G.nzs - 2 * se.zmin;
  const float2 edge = se.zmin == 0 ? v : make_float2(-1000.f, -1000.f);
  const size_t L = packed_row_len(G.nrad);
  auto put = [&](size_t row, float2 val) {
    float2* p = c2 + row * L + r + 1;
    p[0] = val;
    if (r == 0) p[-1] = val;
    if (r == G.nrad - 1) p[1] = val;
  };
  put(row0 + 1 + k, v);
  if (k == 0) put(row0, edge);
  if (k == nz - 1) put(row0 + nz + 1, edge);
}

// Interleaved rows of up to three frames of one geometry, calibration mask
// and subset: the layout above with 16-byte entries {ratio_0, ratio_1,
// ratio_2, calib}. Writes frame ch's ratio and the calibration (the same
// value from every frame) of one bin, with the same halo entries and extra
// rows.
__device__ __forceinline__ void store_packed4(const DevGeom& G, float* c4, int ag, int s, int r,
                                              int ch, float ratio, float calib) {
  const int2 sp = G.sino[s];
  const SegEntry se = G.seg[sp.x];
  const size_t row0 = (size_t)ag * packed_rows_per_angle(G) + se.nza + 2 * (size_t)sp.x;
  const int k = sp.y - se.zmin;
  const int nz = G.nzs - 2 * se.zmin;
  const bool direct = se.zmin == 0;
  const size_t L = packed_row_len(G.nrad);
  auto put = [&](size_t row, float rv, float cv) {
    float* p = c4 + (row * L + r + 1) * 4;
    p[ch] = rv;
    p[3] = cv;
    if (r == 0) {
      p[ch - 4] = rv;
      p[-1] = cv;
    }
    if (r == G.nrad - 1) {
      p[4 + ch] = rv;
      p[7] = cv;
    }
  };
  put(row0 + 1 + k, ratio, calib);
  if (k == 0) put(row0, direct ? ratio : -1000.f, direct ? calib : -1000.f);
  if (k == nz - 1) put(row0 + nz + 1, direct ? ratio : -1000.f, direct ? calib : -1000.f);
}

}  // namespace gfr
