// gfr_synth.hip — seeded synthetic data on the device (the step before the
// path: measured sinograms for benchmarks and tests).
//
// Counter-based uniforms: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel
// random numbers: as easy as 1, 2, 3", SC'11), keyed by the seed, counter =
// (element index, draw index); every element is independent of the launch
// shape, so any GPU regenerates the same data.
// Poisson variates: inversion of the CDF for mean < 10; transformed rejection
// with squeeze (PTRS, W. Hormann, "The transformed rejection method for
// generating Poisson random variables", Insurance: Mathematics and Economics
// 12, 1993) for mean >= 10.
#include <math.h>

#include "gfr_device.h"

namespace gfr {

namespace {

struct U4 {
  uint32_t x, y, z, w;
};

__device__ __forceinline__ void mulhilo(uint32_t a, uint32_t b, uint32_t& hi, uint32_t& lo) {
  const uint64_t p = (uint64_t)a * (uint64_t)b;
  hi = (uint32_t)(p >> 32);
  lo = (uint32_t)p;
}

__device__ U4 philox(U4 c, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int i = 0; i < 10; i++) {
    uint32_t h0, l0, h1, l1;
    mulhilo(M0, c.x, h0, l0);
    mulhilo(M1, c.z, h1, l1);
    c = U4{h1 ^ c.y ^ k0, l1, h0 ^ c.w ^ k1, l0};
    k0 += W0;
    k1 += W1;
  }
  return c;
}

// Uniform in (0, 1) from 32 random bits (never 0 or 1).
__device__ __forceinline__ double u01(uint32_t v) { return ((double)v + 0.5) * (1.0 / 4294967296.0); }

struct Stream {
  U4 ctr;
  uint32_t k0, k1;
  U4 buf;
  int used;
  __device__ double next() {
    if (used == 4) {
      buf = philox(ctr, k0, k1);
      ctr.y++;
      used = 0;
    }
    const uint32_t v = used == 0 ? buf.x : used == 1 ? buf.y : used == 2 ? buf.z : buf.w;
    used++;
    return u01(v);
  }
};

__device__ long long poisson(double lam, Stream& rs) {
  if (!(lam > 0.0)) return 0;
  if (lam < 10.0) {
    const double u = rs.next();
    double p = exp(-lam), F = p;
    long long k = 0;
    while (u > F && k < 1000) {
      k++;
      p *= lam / (double)k;
      F += p;
    }
    return k;
  }
  const double slam = sqrt(lam), loglam = log(lam);
  const double b = 0.931 + 2.53 * slam;
  const double a = -0.059 + 0.02483 * b;
  const double invalpha = 1.1239 + 1.1328 / (b - 3.4);
  const double vr = 0.9277 - 3.6224 / (b - 2.0);
  for (int it = 0; it < 1000; it++) {
    const double U = rs.next() - 0.5;
    const double V = rs.next();
    const double us = 0.5 - fabs(U);
    const long long k = (long long)floor((2.0 * a / us + b) * U + lam + 0.43);
    if (us >= 0.07 && V <= vr) return k;
    if (k < 0 || (us < 0.013 && V > us)) continue;
    if (log(V) + log(invalpha) - log(a / (us * us) + b) <=
        -lam + (double)k * loglam - lgamma((double)k + 1.0))
      return k;
  }
  return (long long)floor(lam + 0.5);
}

__global__ void k_poisson(const float* __restrict__ mean, double scale,
                          const float* __restrict__ mean2, double scale2, uint64_t seed,
                          float* __restrict__ out, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Stream rs;
  rs.ctr = U4{(uint32_t)i, 0u, (uint32_t)((uint64_t)i >> 32), 0x50u};
  rs.k0 = (uint32_t)seed;
  rs.k1 = (uint32_t)(seed >> 32);
  rs.used = 4;
  const double lam = (double)mean[i] * scale + (mean2 ? (double)mean2[i] * scale2 : 0.0);
  out[i] = (float)poisson(lam, rs);
}

}  // namespace

hipError_t launch_poisson(const float* mean, double scale, const float* mean2, double scale2,
                          uint64_t seed, float* out, long long n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_poisson, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, mean, scale,
                     mean2, scale2, seed, out, n);
  return hipGetLastError();
}

}  // namespace gfr
