/* Checks the arithmetic of the joint bilateral filter's fast intensity
 * weight (the radius-4 kernel in gfirst_amd/csrc/gfr_kernels.hip) on the
 * CPU: the same operations in the same order, against the reference
 * expression H = (float)exp(-(double)di2 / kI), di2 = RN(RN(a / m)^2)
 * (jointbilateralfilter3D_K, GFIRST_GPU_Kernels.cu:508-539).
 *
 *   div   : the reciprocal + two FMA corrections quotient equals IEEE a/m
 *           (every divisor significand x `n` dividends, and every dividend
 *           significand of one binade x a few divisors when n is large)
 *   exp   : every `stride`-th float di2 with x >= -105: a weight kept by the
 *           fast path (not sent to the reference expression) is identical
 *   whole : random a, m over all regimes (denormal, tiny differences, huge,
 *           arbitrary bit patterns) for one sigma_I
 * Prints "<check> <cases> <slow> <bad>" per check; bad must be 0. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static double T[1024];
static const double kCN = 1477.3197218702985, kShift = 0x1.8p52;
static const double kLn2N = 0x1.62e42fefa39efp-1 / 1024;
static uint64_t st = 88172645463325252ull;
static uint64_t rnd(void) { st ^= st << 13; st ^= st >> 7; st ^= st << 17; return st; }
static float fb(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static int same(float a, float b) { return memcmp(&a, &b, 4) == 0; }

static float fast_div(float a, float m, float ym) {
  volatile float q0 = a * ym;
  float q1 = fmaf(fmaf(-q0, m, a), ym, q0);
  return fmaf(fmaf(-q1, m, a), ym, q1);
}
/* returns the weight; *slow = 1 when the kernel recomputes it */
static float fast_weight(float di2, double nki, int* slow) {
  double x = (double)di2 * nki;
  double sh = fma(x, kCN, kShift);
  uint64_t b; memcpy(&b, &sh, 8);
  int n = (int)(uint32_t)b;
  double r = fma(-(sh - kShift), kLn2N, x);
  double p = fma(r, 1.0 / 6, 0.5); p = fma(p, r, 1.0); p = fma(p, r, 1.0);
  double E = ldexp(T[n & 1023] * p, n >> 10);
  uint64_t eb; memcpy(&eb, &E, 8);
  uint32_t low = (uint32_t)eb & 0x1fffffffu;
  int amb = (uint32_t)(low - ((1u << 28) - (1u << 13))) < (1u << 14);
  int under = x < -104.5;
  *slow = !under && (!(x >= -87.0) || amb);
  return under ? 0.f : (float)E;
}
static float ref_weight(float di2, double kI) { return (float)exp(-(double)di2 / kI); }

int main(int argc, char** argv) {
  const char* what = argc > 1 ? argv[1] : "all";
  long long n = argc > 2 ? atoll(argv[2]) : 4;
  double sI = argc > 3 ? atof(argv[3]) : 0.025;
  double kI = (2.0 * sI) * (2.0 * sI), nki = -1.0 / kI;
  for (int j = 0; j < 1024; j++) T[j] = exp2((double)j / 1024);
  int all = !strcmp(what, "all");
  if (all || !strcmp(what, "div")) {
    long long c = 0, bad = 0;
    for (uint32_t sm = 0; sm < (1u << 23); sm++) {
      float m = fb((127u << 23) | sm);
      volatile float ym = 1.0f / m;
      for (long long k = 0; k < n; k++) {
        float a = fb(((100u + (uint32_t)(rnd() % 50)) << 23) | (uint32_t)(rnd() & 0x7fffff));
        volatile float q = a / m;
        c++;
        if (!same(fast_div(a, m, ym), q)) bad++;
      }
    }
    printf("div %lld 0 %lld\n", c, bad);
  }
  if (all || !strcmp(what, "exp")) {
    long long c = 0, slow = 0, bad = 0, stride = n > 0 ? n : 1;
    for (uint64_t u = 0; u < 0x7f800000u; u += (uint64_t)stride) {
      float di2 = fb((uint32_t)u);
      if (-(double)di2 / kI < -105.0) break;
      int s;
      float h = fast_weight(di2, nki, &s);
      c++;
      if (s) slow++;
      else if (!same(h, ref_weight(di2, kI))) bad++;
    }
    printf("exp %lld %lld %lld\n", c, slow, bad);
  }
  if (all || !strcmp(what, "whole")) {
    long long c = 0, slow = 0, bad = 0;
    for (long long it = 0; it < n * 1000000; it++) {
      float m = fb(((107u + (uint32_t)(rnd() % 40)) << 23) | (uint32_t)(rnd() & 0x7fffff));
      float rv = 0.f, g;
      int mode = (int)(rnd() % 4);
      if (mode == 0) g = fb((uint32_t)(rnd() & 0x7fffffff));
      else if (mode == 1) g = m * (float)((double)(rnd() % 2000000) / 1e6);
      else if (mode == 2) {
        rv = m * (float)((double)(rnd() % 1000000) / 1e6);
        g = nextafterf(rv, INFINITY);
        for (int k = (int)(rnd() % 40); k > 0; k--) g = nextafterf(g, INFINITY);
      } else g = fb((uint32_t)(rnd() % 0x01000000));
      if (!isfinite(g)) continue;
      float a = g - rv;
      volatile float ym = 1.0f / m;
      float tq = fast_div(a, m, ym);
      volatile float di2 = tq * tq;
      int s;
      float h = fast_weight(di2, nki, &s);
      volatile float te = a / m;
      volatile float de = te * te;
      c++;
      if (s) slow++;
      else if (!same(h, ref_weight(de, kI))) bad++;
    }
    printf("whole %lld %lld %lld\n", c, slow, bad);
  }
  return 0;
}
