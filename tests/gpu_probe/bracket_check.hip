// Test-only library: exhaustive check of the MFMA form's certified
// acceptance brackets (pysa_amd/csrc/accept.hpp) over every 32-bit Philox
// word.  PySA accepts a proposal iff log_r < beta * dn (pysa/ising.py:52),
// here log_r = det_log((w + 0.5) 2^-32); for beta > 0 that is dn > L / beta
// with L = det_log(u).  The engine's fast path takes dn > hi as an accept and
// dn < lo as a reject, so it is exact iff lo <= L / beta <= hi for every
// word, i.e. lo * beta <= L <= hi * beta, both products exact in f64.
// Compiled with the engine's flags, so det_log and the hardware logarithm
// give the engine's bits.  Every launch exits normally; violations are
// counted, never trapped on.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "accept.hpp"
#include "rng.hpp"

namespace {

constexpr uint32_t kWordsPerThread = 256;

__global__ __launch_bounds__(256) void bracket_check_kernel(float beta, int variant,
                                                            unsigned long long* bad,
                                                            double* widest) {
  const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;  // < 2^24
  const uint32_t w0 = (uint32_t)(gid * kWordsPerThread);
  float cb, c0;
  pysa::bracket_factors(beta, cb, c0);
  const float inv_b = 1.0f / beta;
  unsigned long long nb = 0;
  double wmax = 0.0;
  for (uint32_t k = 0; k < kWordsPerThread; ++k) {
    const uint32_t w = w0 + k;
    float lo, hi;
    if (variant == 0)
      pysa::accept_bounds(w, beta, inv_b, lo, hi);
    else
      pysa::accept_bounds_f(w, cb, c0, lo, hi);
    const double L = pysa::log_u32(w);
    const double lb = (double)lo * (double)beta, hb = (double)hi * (double)beta;
    if (!(lb <= L && L <= hb)) ++nb;
    const double width = (hb - lb) / -L;  // L < 0 for every word
    wmax = width > wmax ? width : wmax;
  }
  if (nb) atomicAdd(bad, nb);
  // relative bracket width, for the record (max over the wave)
  for (int o = 32; o >= 1; o >>= 1) {
    const double t = __shfl_xor(wmax, o, 64);
    wmax = t > wmax ? t : wmax;
  }
  if ((threadIdx.x & 63) == 0) {
    unsigned long long* wb = reinterpret_cast<unsigned long long*>(widest);
    atomicMax(wb, (unsigned long long)__double_as_longlong(wmax));  // positive doubles order as bits
  }
}

}  // namespace

extern "C" int bracket_check(float beta, int variant, unsigned long long* bad_out,
                             double* widest_out) {
  unsigned long long* d = nullptr;
  if (hipMalloc(&d, 2 * sizeof(unsigned long long)) != hipSuccess) return -1;
  int rc = 0;
  if (hipMemset(d, 0, 2 * sizeof(unsigned long long)) != hipSuccess) rc = -2;
  if (!rc) {
    const unsigned threads = 256, blocks = (unsigned)((1ull << 32) / kWordsPerThread / threads);
    hipLaunchKernelGGL(bracket_check_kernel, dim3(blocks), dim3(threads), 0, 0, beta, variant, d,
                       reinterpret_cast<double*>(d + 1));
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) rc = -3;
  }
  unsigned long long h[2] = {0, 0};
  if (!rc && hipMemcpy(h, d, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess) rc = -4;
  (void)hipFree(d);
  *bad_out = h[0];
  double wd;
  __builtin_memcpy(&wd, &h[1], 8);
  *widest_out = wd;
  return rc;
}
