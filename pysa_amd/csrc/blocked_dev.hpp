// Device helpers and launch parameters of the MFMA form, shared by its
// translation units (blocked.hip: encoding, the split-K kernel and the host
// side; blocked_pipe_*.hip: the pipelined kernel, one coupling encoding each).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>
#include <utility>

#include "accept.hpp"
#include "engine.hpp"
#include "rng.hpp"

namespace pysa {
namespace blk {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(16))) float f32x16;
typedef __attribute__((ext_vector_type(4))) int i32x4;
typedef __attribute__((ext_vector_type(16))) int i32x16;
typedef __attribute__((ext_vector_type(4))) unsigned u32x4v;

constexpr int kB = 32;  // block size (rows of one MFMA tile)

// x with bit k moved to bit k ^ m (m < 32, wave-uniform): the in-block
// permutation of the block-random order, an involution.  Bits 0-2 of m: one
// exchange of bit pairs d apart per bit, shifted by m & d (a shift by 0
// leaves x as it is, so no select); bits 3-4: the bytes exchanged by one
// byte permute (byte i from byte i ^ (m >> 3)).
__device__ __forceinline__ uint32_t xor_permute_bits(uint32_t x, uint32_t m) {
  const uint32_t s1 = m & 1u, s2 = m & 2u, s4 = m & 4u;
  x = ((x << s1) & 0xAAAAAAAAu) | ((x >> s1) & 0x55555555u);
  x = ((x << s2) & 0xCCCCCCCCu) | ((x >> s2) & 0x33333333u);
  x = ((x << s4) & 0xF0F0F0F0u) | ((x >> s4) & 0x0F0F0F0Fu);
  const uint32_t sel = 0x03020100u ^ (((m >> 3) & 3u) * 0x01010101u);
  return __builtin_amdgcn_perm(x, x, sel);
}

// ------------------------------------------------------------ device util --
// 16 spin bits of one lane-half of a block: bf16 form keeps positions
// {8h..8h+7} in the low byte and {16+8h..16+8h+7} in the high byte; int8
// form keeps positions 16h..16h+15.
template <bool I8>
__device__ __forceinline__ uint32_t half_of(uint32_t blk, int h) {
  if (I8) return (blk >> (16 * h)) & 0xFFFFu;
  return ((blk >> (8 * h)) & 0xFFu) | (((blk >> (16 + 8 * h)) & 0xFFu) << 8);
}
template <bool I8>
__device__ __forceinline__ uint32_t block_of(uint32_t h0, uint32_t h1) {
  if (I8) return (h0 & 0xFFFFu) | (h1 << 16);
  return (h0 & 0xFFu) | ((h1 & 0xFFu) << 8) | ((h0 >> 8) << 16) | ((h1 >> 8) << 24);
}

// Sixteen spins (bits 0-15 of y: K step 0 in bits 0-7, step 1 in 8-15) as
// bf16 +1 / -1 operand words, two spins per dword.  nz = ~(y | y << 15) puts
// the complement of spin 2e at bit 2e and of spin 2e+1 at bit 2e+16; a shift
// by 15 - 2e (step 0) or 7 - 2e (step 1) moves them to the sign bits 15 and
// 31, which a bit-field insert drops into -1.0 (0xBF80).  (The blocked form
// always multiplies +-1 spins; QUBO fields are recovered from J.s as
// (J.s + J.1) / 2.)
template <int ST>
__device__ __forceinline__ void spins_bf16(uint32_t nz, uint32_t (&d)[4]) {
#pragma unroll
  for (int e = 0; e < 4; ++e) d[e] = ((nz << (15 - 8 * ST - 2 * e)) & 0x80008000u) | 0x3F803F80u;
}
__device__ __forceinline__ uint32_t spins_nz(uint32_t y) { return ~(y | (y << 15)); }
// The same with the mask and the +-1 pattern held in registers (vm =
// 0x80008000, vc = 0x3F803F80), so each word is a shift and one AND-OR.
template <int ST>
__device__ __forceinline__ void spins_bf16_r(uint32_t nz, uint32_t vm, uint32_t vc,
                                             uint32_t (&d)[4]) {
#pragma unroll
  for (int e = 0; e < 4; ++e) d[e] = ((nz << (15 - 8 * ST - 2 * e)) & vm) | vc;
}
// Sixteen spins (bits of x) as int8: ising (mul, xr) = (0xFE, 0xFFFFFFFF)
// gives +1 / -1 per byte, qubo (1, 0) gives 1 / 0.
__device__ __forceinline__ void spins_i8(uint32_t x, uint32_t mul, uint32_t xr, uint32_t (&d)[4]) {
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const uint32_t t = (x >> (4 * e)) & 15u;
    const uint32_t sp = (t * 0x00204081u) & 0x01010101u;  // bit b -> byte b
    d[e] = (sp * mul) ^ xr;
  }
}

// f(std::integral_constant<int, 0>{}), ..., f(std::integral_constant<int, N-1>{}):
// a loop whose index is a compile-time constant in the body, so arrays indexed
// by expressions of it stay in registers.
template <typename F, int... K>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, K...>) {
  (f(std::integral_constant<int, K>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

template <int NW>
struct Bits {
  typedef uint32_t V __attribute__((ext_vector_type(NW)));
  V w;
};

// Value of sub-lane `a` of each group of LPR consecutive lanes, broadcast to
// the whole group (DPP quad permute; LPR divides 4).
template <int LPR>
__device__ __forceinline__ float group_bcast(float v, int a) {
  if (LPR == 1) return v;
  const int ctrl = (LPR == 4) ? (a * 0x55) : (a | (a << 2) | ((2 + a) << 4) | ((2 + a) << 6));
  int r;
  switch (ctrl) {  // the DPP control must be a compile-time constant
#define PYSA_DPP(C) \
  case C: r = __builtin_amdgcn_mov_dpp(__float_as_int(v), C, 0xF, 0xF, false); break;
    PYSA_DPP(0x00) PYSA_DPP(0x55) PYSA_DPP(0xAA) PYSA_DPP(0xFF) PYSA_DPP(0xA0) PYSA_DPP(0xF5)
#undef PYSA_DPP
    default: r = __float_as_int(v);
  }
  return __int_as_float(r);
}

template <int NW>
__device__ __forceinline__ uint32_t get_half(const Bits<NW>& s, int c) {
  return (s.w[c >> 1] >> (16 * (c & 1))) & 0xFFFFu;
}
template <int NW>
__device__ __forceinline__ void set_half(Bits<NW>& s, int c, uint32_t v) {
  const int sh = 16 * (c & 1);
  s.w[c >> 1] = (s.w[c >> 1] & ~(0xFFFFu << sh)) | ((v & 0xFFFFu) << sh);
}

struct BParams {
  int n, nblk, ptype, R, reads, S, rpw, np, use_pt, logw, best_sem;
  int order;  // kSequential or kBlockRandom (pipelined kernel only)
  int refresh;  // pipelined kernel: sweeps between f64 energy recomputations (0 = never)
  int kb;  // row blocks of the K range each wave of a group owns
  int hzero;  // every local field is zero (the field sums skip h)
  uint32_t k0, k1;
  uint64_t sweep_offset;
  const void* frag;
  // f32 J^T with every coupling replaced by the sum of its bf16 pieces (J
  // itself when the pieces are exact), row j = coefficients J[:][j]: the
  // in-block updates use the same couplings as the matrix products, so every
  // decision and energy change comes from one energy function
  const float* Jt;
  const float* h;
  const float* rowsum;  // n row sums of J (QUBO with bf16 operands), else null
  const float* betas;
  const int* bsort;
  int8_t* spins;
  float* energies;
  int64_t* beta_idx;
  int8_t* best_rep;     // reads*R*n (per-replica best, best_sem 0)
  float* best_rep_e;    // reads*R
  int* best_rep_sw;     // reads*R
  int8_t* best_state;   // reads*n
  float* best_energy;   // reads
  int* best_sweeps;     // reads
  double* log_omegas;   // reads*S or null
  unsigned long long* accepted;
  // batched MFMA form (pipelined kernel with BAT): one entry per workgroup,
  // the int8 coupling matrices of the batch (offsets in the entries); h,
  // betas and frag are then the batch's base pointers
  const BatWG* bat;
  const int8_t* sgn;
  // AIS (pipelined kernel with kModeAIS): the schedule temps[S] (the beta of
  // each sweep), the weight ordering (0: after the sweep, 1: before it), and
  // per chain (read g, R = 1) the f64 energy (in: initial, out: tracked) and
  // the log-weight (out)
  const float* temps;
  int ordering;
  double* ais_e;
  double* ais_w;
};

// Store the spins of one operand-role replica for the blocks this wave owns.
template <int NW, bool I8>
__device__ __forceinline__ void store_range(const Bits<NW>& s, int c0, int kb, int nblk, int n,
                                            int ptype, int8_t* dst) {
  const int h = (threadIdx.x & 63) >> 5;
  const int8_t dn = ptype == 0 ? int8_t(-1) : int8_t(0);
  for (int j = 0; j < kb; ++j) {
    const int c = c0 + j;
    if (c >= nblk) break;
    const uint32_t hv = get_half<NW>(s, j);
#pragma unroll
    for (int b = 0; b < 16; ++b) {
      const int rel = I8 ? (16 * h + b) : ((b < 8) ? (8 * h + b) : (16 + 8 * h + (b - 8)));
      const int pos = 32 * c + rel;
      if (pos < n) dst[pos] = ((hv >> b) & 1) ? int8_t(1) : dn;
    }
  }
}


}  // namespace blk
}  // namespace pysa
