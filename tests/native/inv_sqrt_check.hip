// Checks on the GPU that rflu::inv_sqrt_cr (the stencil weights' 1/sqrt) returns the
// bits of 1.0 / sqrt(x): log-uniform arguments over its fast range [2^-700, 2^700],
// arguments in [2^-27, 2^27] (cell distances), perfect squares k^2 with their two
// neighbouring doubles (the hard cases of a square root's rounding), arguments outside
// the fast range (the library path), and squares of values whose significand is near
// all ones or near one, with their neighbours (the hard cases of the quotient 1/y).  Usage: inv_sqrt_check [log2 samples
// per set]; exit status 0 when every result matches.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>

#include "kernels.hpp"

__device__ __forceinline__ uint64_t splitmix(uint64_t z) {
  z += 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

__global__ void k_check(int set, uint64_t n, uint64_t seed, unsigned long long* bad,
                        double* example) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint64_t r = splitmix(seed ^ (i * 0x632be59bd9b4e019ULL));
    const double m = 1.0 + (double)(r >> 11) * 0x1p-53;
    double x;
    if (set == 0) {
      x = ldexp(m, (int)(r % 1401) - 700);
    } else if (set == 1) {
      x = ldexp(m, (int)(r % 55) - 27);
    } else if (set == 2) {
      // k < 2^26 keeps k*k below 2^52, so the product is exact: a perfect square
      const double k = (double)(1 + (r % ((1ULL << 26) - 1)));
      const uint64_t b = (uint64_t)__double_as_longlong(k * k);
      const int w = (int)(i % 3);
      x = __longlong_as_double((long long)(w == 0 ? b : w == 1 ? b + 1 : b - 1));
      x = ldexp(x, 2 * (int)((r >> 40) % 201) - 200);
    } else if (set == 4) {
      // the quotient's hard cases: square roots whose significand is within 1024 ulps of
      // all ones or of one, with the squares' two neighbouring doubles
      const double j = (double)(r % 1024);
      const double q = (r >> 10) & 1 ? 2.0 - j * 0x1p-52 : 1.0 + j * 0x1p-52;
      const uint64_t b = (uint64_t)__double_as_longlong(q * q);
      const int w = (int)((r >> 11) % 3);
      x = __longlong_as_double((long long)(w == 0 ? b : w == 1 ? b + 1 : b - 1));
      x = ldexp(x, 2 * (int)((r >> 20) % 601) - 600);
    } else {
      x = ldexp(m, (r & 1) ? 700 + (int)((r >> 1) % 300) : -700 - (int)((r >> 1) % 370));
    }
    const double got = rflu::inv_sqrt_cr(x), want = 1.0 / sqrt(x);
    if (__double_as_longlong(got) != __double_as_longlong(want)) {
      atomicAdd(bad, 1ull);
      *example = x;
    }
  }
}

int main(int argc, char** argv) {
  const int lg = argc > 1 ? atoi(argv[1]) : 26;
  if (lg < 10 || lg > 32) return 2;
  unsigned long long* bad = nullptr;
  double* ex = nullptr;
  if (hipMalloc(&bad, sizeof(unsigned long long)) != hipSuccess) return 2;
  if (hipMalloc(&ex, sizeof(double)) != hipSuccess) return 2;
  int rc = 0;
  for (int set = 0; set < 5; ++set) {
    const uint64_t n = 1ULL << lg;
    if (hipMemset(bad, 0, sizeof(unsigned long long)) != hipSuccess) return 2;
    hipLaunchKernelGGL(k_check, dim3(4096), dim3(256), 0, 0, set, n, 977ULL + set, bad, ex);
    if (hipDeviceSynchronize() != hipSuccess) return 3;
    unsigned long long h = 0;
    double e = 0.0;
    if (hipMemcpy(&h, bad, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess) return 3;
    if (hipMemcpy(&e, ex, sizeof(e), hipMemcpyDeviceToHost) != hipSuccess) return 3;
    printf("set %d: %llu arguments, %llu mismatches%s", set, (unsigned long long)n, h, h ? "" : "\n");
    if (h) printf(" (e.g. x = %a)\n", e);
    fflush(stdout);
    if (h) rc = 1;
  }
  (void)hipFree(bad);
  (void)hipFree(ex);
  return rc;
}
