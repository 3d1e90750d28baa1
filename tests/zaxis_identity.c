/* Exhaustive-by-sampling check of the shared z axis used by the
 * backprojector: for slices zb..zb+ZB-1 of one voxel column and view, the
 * reference's plane position w_j = fmaf(yr, fzb, zb + j) + 0.5f, its floor
 * and fraction (GFIRST_GPU_Kernels.cu:134-157 via the texture's t - 0.5) equal
 * w_0 + j, floor(s0) + j and s0 - floor(s0) whenever s0 = fmaf(yr, fzb, zb)
 * >= 1 and s0 and w_{ZB-1} share a binary exponent. Prints "cases fast bad". */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static uint32_t bits(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  return u;
}
static uint64_t st = 88172645463325252ULL;
static double rnd(void) {
  st ^= st << 13;
  st ^= st >> 7;
  st ^= st << 17;
  return (double)(st >> 11) * (1.0 / 9007199254740992.0);
}

int main(int argc, char** argv) {
  const long long n = argc > 1 ? atoll(argv[1]) : 10000000LL;
  long long fast = 0, bad = 0;
  for (long long it = 0; it < n; it++) {
    const int ZB = rnd() < 0.5 ? 8 : (rnd() < 0.5 ? 4 : 16);
    float fzb = (float)(rnd() * 0.8);
    if (rnd() < 0.05) fzb = 0.f;
    const int zb = ((int)(rnd() * 170) / ZB) * ZB;
    float yr = (float)((rnd() * 2 - 1) * 100.0);
    if (rnd() < 0.5 && fzb > 0.f) {  // aim s0 at a power of two (+- a few ulps, or an integer/half below)
      const int k = (int)(rnd() * 8);
      const double tgt = ldexp(1.0, k) + (rnd() * 2 - 1) * ldexp(1.0, k - 22) -
                         (rnd() < 0.5 ? 0 : (int)(rnd() * 8)) - (rnd() < 0.3 ? 0.5 : 0);
      yr = (float)((tgt - zb) / fzb);
    }
    const float s0 = fmaf(yr, fzb, (float)zb), sl = fmaf(yr, fzb, (float)(zb + ZB - 1));
    const float w0 = s0 + 0.5f, wl = sl + 0.5f;
    if (!(s0 >= 1.f && (bits(s0) >> 23) == (bits(wl) >> 23))) continue;
    fast++;
    const float f0 = floorf(s0), al = s0 - f0;
    for (int j = 0; j < ZB; j++) {
      const float w = fmaf(yr, fzb, (float)(zb + j)) + 0.5f;
      const float t = w - 0.5f, f = floorf(t), a = t - f;
      if (bits(w) != bits(w0 + (float)j) || bits(a) != bits(al) || (int)f != (int)f0 + j) bad++;
    }
  }
  printf("%lld %lld %lld\n", n, fast, bad);
  return bad != 0;
}
