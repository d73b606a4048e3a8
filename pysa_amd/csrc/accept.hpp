// Certified acceptance brackets of the MFMA form.  Shared by the engine
// (blocked.hip) and the test-only exhaustive check of the brackets.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "rng.hpp"

namespace pysa {

// Certified acceptance thresholds.  PySA accepts a proposal iff
// dn >= 0 or det_log(u) < beta*dn (product exact in f64), with
// u = (w + 0.5) 2^-32.  For beta > 0 that is dn > q with q = det_log(u)/beta
// (q < 0).  `accept_bounds` brackets q from the hardware logarithm:
// dn > hi certainly accepts, dn < lo certainly rejects, and only a dn inside
// [lo, hi] (a relative width of a few 1e-6) needs the exact test.  The bound
// on |L - log(u)| covers the rounding of u to f32, the hardware log2 and the
// scaling by ln 2; the bracket adds the rounding of q = L/beta and of hi, lo.
// beta <= 0 accepts every proposal (hi = lo = -inf).
__device__ __forceinline__ void accept_bounds(uint32_t w, float beta, float inv_b, float& lo,
                                              float& hi) {
  if (!(beta > 0.f)) {
    lo = hi = -INFINITY;
    return;
  }
  if (!(inv_b < 1e30f)) {  // beta too small to scale: every decision goes exact
    lo = -INFINITY;
    hi = INFINITY;
    return;
  }
  const float uf = ((float)w + 0.5f) * 2.3283064365386963e-10f;
  const float L = __builtin_amdgcn_logf(uf) * 0.69314718055994531f;
  const float q = L * inv_b;
  const float m = (3.9e-6f * fabsf(L) + 9.6e-7f) * inv_b * 1.001f + 3.6e-7f * fabsf(q);
  hi = q + m;
  lo = q - m;
}

// The same bracket with the per-replica factors taken out of the per-row
// work: q = log2(uf) * cb with cb = ln2 / beta (one rounding fewer than
// L = log2(uf) ln2 followed by q = L / beta, the same count overall), and
// m = |q| A + c0 with A = 4.27e-6 >= 1.001 * 3.9e-6 (1 + 2^-22) + 3.6e-7,
// c0 = 9.62e-7 / beta >= 1.001 * 9.6e-7 / beta, both rounded up so that the
// fma's own rounding stays inside the margin of accept_bounds.  beta <= 0
// gives cb = 0, c0 = -inf: hi = -inf, every proposal taken; beta too small
// to scale gives cb = 0, c0 = +inf: every decision exact.
__device__ __forceinline__ void bracket_factors(float beta, float& cb, float& c0) {
  if (!(beta > 0.f)) {
    cb = 0.f;
    c0 = -INFINITY;
    return;
  }
  const float inv_b = 1.0f / beta;
  if (!(inv_b < 1e30f)) {
    cb = 0.f;
    c0 = INFINITY;
    return;
  }
  cb = 0.69314718055994531f * inv_b;
  c0 = 9.62e-7f * inv_b;
}
__device__ __forceinline__ void accept_bounds_f(uint32_t w, float cb, float c0, float& lo,
                                                float& hi) {
  // (w + 0.5) 2^-32 rounded as in accept_bounds (scaling by 2^-32 is exact)
  const float uf = __builtin_fmaf((float)w, 2.3283064365386963e-10f, 1.1641532182693481e-10f);
  const float q = __builtin_amdgcn_logf(uf) * cb;
  const float m = __builtin_fmaf(fabsf(q), 4.27e-6f, c0);
  hi = q + m;
  lo = q - m;
}

// The exact test (pysa/ising.py:52) for the rare dn inside the bracket.
__device__ __noinline__ inline bool accept_exact(uint32_t w, float beta, float dn) {
  return dn >= 0.f || log_u32(w) < (double)beta * (double)dn;
}

}  // namespace pysa
