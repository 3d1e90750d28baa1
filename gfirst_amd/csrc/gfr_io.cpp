// gfr_io.cpp — the reference's raw-file semantics (host only).
#include <stdio.h>
#include <string.h>

#include <vector>

#include "gfirst.h"

extern "C" {

// ReadSinogram_Type (GFIRST_CPU_Kernels.cpp:104-142): a zeroed 2*NDATA buffer
// (prompts, background) receives filesize/4 words read as int32; if the max of
// that int view exceeds 1e6 the words are reinterpreted as float32. A missing
// file leaves zeros (the reference prints "Sinogram Not Found!"). Words past
// 2*NDATA are dropped and reported (the reference writes past its buffer).
int gfr_read_sinogram_file(const char* path, long long ndata, float* out, int* is_float,
                           int* truncated) {
  if (!path || !out || ndata <= 0) return GFR_EINVAL;
  const long long n = 2 * ndata;
  memset(out, 0, sizeof(float) * (size_t)n);
  if (is_float) *is_float = 0;
  if (truncated) *truncated = 0;
  FILE* f = fopen(path, "rb");
  if (!f) return 1;
  if (fseek(f, 0L, SEEK_END) != 0) {
    fclose(f);
    return GFR_EIO;
  }
  const long long words = ftell(f) / 4;
  rewind(f);
  const long long m = words < n ? words : n;
  if (truncated && words > n) *truncated = 1;
  std::vector<int32_t> buf((size_t)m);
  if (m > 0 && fread(buf.data(), 4, (size_t)m, f) != (size_t)m) {
    fclose(f);
    return GFR_EIO;
  }
  fclose(f);
  float mx = 0.f;
  for (long long i = 0; i < m; i++) {
    out[i] = (float)buf[i];
    if (out[i] > mx) mx = out[i];
  }
  if (mx > 1.0e6) {
    memcpy(out, buf.data(), sizeof(float) * (size_t)m);
    if (is_float) *is_float = 1;
  }
  return 0;
}

// VLOR file -> records as the reference's while(!feof) loop sees them
// (GFIRST.cu:233-240): every full (i1+1, i2+1) int32 pair, then the buffer
// once more as the final fread left it (the last record again, or for an odd
// word count the last word with the previous second index). An empty file
// yields no record (the reference would process uninitialised values).
long long gfr_read_vlor_file(const char* path, int32_t* rec, long long cap) {
  if (!path) return GFR_EINVAL;
  FILE* f = fopen(path, "rb");
  if (!f) return GFR_EIO;
  if (fseek(f, 0L, SEEK_END) != 0) {
    fclose(f);
    return GFR_EIO;
  }
  const long long words = ftell(f) / 4;
  rewind(f);
  if (words == 0) {
    fclose(f);
    return 0;
  }
  const long long nrec = words / 2 + 1;
  if (!rec) {
    fclose(f);
    return nrec;
  }
  if (cap < nrec) {
    fclose(f);
    return GFR_EINVAL;
  }
  std::vector<int32_t> buf((size_t)words);
  if (fread(buf.data(), 4, (size_t)words, f) != (size_t)words) {
    fclose(f);
    return GFR_EIO;
  }
  fclose(f);
  int32_t a = 0, b = 0;
  long long k = 0;
  for (long long e = 0; e < words / 2; e++, k++) {
    a = buf[2 * e];
    b = buf[2 * e + 1];
    rec[2 * k] = a;
    rec[2 * k + 1] = b;
  }
  if (words % 2) a = buf[words - 1];
  rec[2 * k] = a;
  rec[2 * k + 1] = b;
  return k + 1;
}

}  // extern "C"
