// gfr_vlor_build.hip — the VLOR record lists on the device.
//
// From the raw records (pairs of 1-based sinogram bins) the session needs, in
// record order, the valid events (GFIRST.cu:235-237: both 0-based indices in
// (0, NDATA)) and, per touched bin, the list of events that name it, each
// list in the order the reference visits them ((i1 of event 0), (i2 of event
// 0), (i1 of event 1), ...). A stable radix sort by bin of that sequence gives
// exactly those lists; the compaction before it keeps record order (exclusive
// scan of the validity flags). Unused slots carry the bin NDATA, which sorts
// after every valid bin.
#include <algorithm>

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "gfr_internal.h"

namespace gfr {

namespace {

__global__ void k_vrec_flags(const int* __restrict__ rec, long long nrec, long long nd,
                             int* __restrict__ flag) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nrec) return;
  const long long a = (long long)rec[2 * e] - 1, b = (long long)rec[2 * e + 1] - 1;
  flag[e] = (a <= 0 || b <= 0 || a >= nd || b >= nd) ? 0 : 1;
}

__global__ void k_vrec_fill(unsigned* __restrict__ key, long long n, unsigned v) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) key[k] = v;
}

__global__ void k_vrec_compact(const int* __restrict__ rec, const int* __restrict__ flag,
                               const int* __restrict__ pos, long long nrec,
                               int* __restrict__ i1, int* __restrict__ i2,
                               unsigned* __restrict__ key, int* __restrict__ val) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nrec || !flag[e]) return;
  const int p = pos[e];
  const int a = rec[2 * e] - 1, b = rec[2 * e + 1] - 1;
  i1[p] = a;
  i2[p] = b;
  key[2 * p] = (unsigned)a;
  key[2 * p + 1] = (unsigned)b;
  val[2 * p] = p;
  val[2 * p + 1] = p;
}

// flag of the first entry of each bin's run (sentinel entries excluded)
__global__ void k_vrec_runs(const unsigned* __restrict__ key, long long n, unsigned sentinel,
                            int* __restrict__ first) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  first[k] = (key[k] != sentinel && (k == 0 || key[k] != key[k - 1])) ? 1 : 0;
}

__global__ void k_vrec_starts(const unsigned* __restrict__ key, const int* __restrict__ first,
                              const int* __restrict__ bpos, long long n,
                              long long* __restrict__ bins, int* __restrict__ start) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n || !first[k]) return;
  bins[bpos[k]] = (long long)key[k];
  start[bpos[k]] = (int)k;
}

// counts[0] = valid events, counts[1] = touched bins; start[nbins] = 2 * nev
__global__ void k_vrec_counts(const int* __restrict__ flag, const int* __restrict__ pos,
                              long long nrec, const int* __restrict__ first,
                              const int* __restrict__ bpos, long long n,
                              int* __restrict__ start, long long* __restrict__ counts) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  const long long nev = (long long)pos[nrec - 1] + flag[nrec - 1];
  const long long nb = (long long)bpos[n - 1] + first[n - 1];
  counts[0] = nev;
  counts[1] = nb;
  start[nb] = (int)(2 * nev);
}

inline unsigned blocks(long long n) { return (unsigned)((n + 255) / 256); }

}  // namespace

hipError_t build_vlor_lists(const int* d_rec, long long nrec, long long nd, int* d_i1, int* d_i2,
                            long long* d_bins, int* d_start, int* d_ev, long long* nev_out,
                            long long* nbins_out, hipStream_t st) {
  *nev_out = *nbins_out = 0;
  if (nrec == 0) return hipSuccess;
  const long long n = 2 * nrec;
  int nbits = 1;
  while (nbits < 32 && ((unsigned long long)nd >> nbits) != 0) nbits++;  // nd < 2^nbits
  int *flag = nullptr, *pos = nullptr, *val = nullptr, *first = nullptr, *bpos = nullptr;
  unsigned *key = nullptr, *key2 = nullptr;
  long long* counts = nullptr;
  void* tmp = nullptr;
  size_t t_scan1 = 0, t_sort = 0, t_scan2 = 0;
  hipError_t e = hipcub::DeviceScan::ExclusiveSum(nullptr, t_scan1, flag, pos, nrec, st);
  if (e == hipSuccess)
    e = hipcub::DeviceRadixSort::SortPairs(nullptr, t_sort, key, key2, val, d_ev, n, 0, nbits, st);
  if (e == hipSuccess) e = hipcub::DeviceScan::ExclusiveSum(nullptr, t_scan2, first, bpos, n, st);
  const size_t t_max = std::max(t_scan1, std::max(t_sort, t_scan2));
  if (e == hipSuccess) e = hipMalloc((void**)&flag, sizeof(int) * nrec);
  if (e == hipSuccess) e = hipMalloc((void**)&pos, sizeof(int) * nrec);
  if (e == hipSuccess) e = hipMalloc((void**)&val, sizeof(int) * n);
  if (e == hipSuccess) e = hipMalloc((void**)&first, sizeof(int) * n);
  if (e == hipSuccess) e = hipMalloc((void**)&bpos, sizeof(int) * n);
  if (e == hipSuccess) e = hipMalloc((void**)&key, sizeof(unsigned) * n);
  if (e == hipSuccess) e = hipMalloc((void**)&key2, sizeof(unsigned) * n);
  if (e == hipSuccess) e = hipMalloc((void**)&counts, 2 * sizeof(long long));
  if (e == hipSuccess) e = hipMalloc(&tmp, t_max ? t_max : 4);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_vrec_flags, dim3(blocks(nrec)), dim3(256), 0, st, d_rec, nrec, nd, flag);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipcub::DeviceScan::ExclusiveSum(tmp, t_scan1, flag, pos, nrec, st);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_vrec_fill, dim3(blocks(n)), dim3(256), 0, st, key, n, (unsigned)nd);
    hipLaunchKernelGGL(k_vrec_compact, dim3(blocks(nrec)), dim3(256), 0, st, d_rec, flag, pos,
                       nrec, d_i1, d_i2, key, val);
    e = hipGetLastError();
  }
  if (e == hipSuccess)
    e = hipcub::DeviceRadixSort::SortPairs(tmp, t_sort, key, key2, val, d_ev, n, 0, nbits, st);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_vrec_runs, dim3(blocks(n)), dim3(256), 0, st, key2, n, (unsigned)nd,
                       first);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipcub::DeviceScan::ExclusiveSum(tmp, t_scan2, first, bpos, n, st);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_vrec_starts, dim3(blocks(n)), dim3(256), 0, st, key2, first, bpos, n,
                       d_bins, d_start);
    hipLaunchKernelGGL(k_vrec_counts, dim3(1), dim3(64), 0, st, flag, pos, nrec, first, bpos, n,
                       d_start, counts);
    e = hipGetLastError();
  }
  long long hc[2] = {0, 0};
  if (e == hipSuccess) e = hipMemcpyAsync(hc, counts, sizeof(hc), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e == hipSuccess) {
    *nev_out = hc[0];
    *nbins_out = hc[1];
  }
  for (void* p : {(void*)flag, (void*)pos, (void*)val, (void*)first, (void*)bpos, (void*)key,
                  (void*)key2, (void*)counts, tmp})
    if (p) (void)hipFree(p);
  return e;
}

}  // namespace gfr
