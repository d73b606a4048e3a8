// MFMA-blocked form of the engine.
//
// Same Markov chain as the plain form (pysa/simulation.py:123-236 with
// ising.py:40-63 / qubo.py:40-63 and utils.py:26-58), computed differently:
// the N spins are swept in blocks of B = 32.  At the start of block c the
// local fields of its 32 spins are recomputed for 64 replicas at once as a
// matrix product on the matrix cores,
//     F[blk, replicas] = J[blk, :] . S[:, replicas]      (K = N)
// then the 32 proposals of the block run sequentially per replica (one lane
// per replica) with in-block field updates from J[blk, blk].  The product
// costs the same 2N flop per spin update as an incremental field update but
// needs no stored fields and cannot drift.
//
// Operands.  Spins stay in VGPRs as bits laid out as the MFMA B operand and
// are expanded to bf16 (+-1 / 0,1) or int8 on the fly.  Couplings are packed
// once per problem into 1 KiB fragments in the MFMA A-operand layout, one per
// (row block, K step, piece):
//   int8   J exactly, when every coupling is an integer in [-128, 127]
//          (v_mfma_i32_32x32x32_i8, exact integer accumulation);
//   bf16   J = hi + lo (+ mid): 1, 2 or 3 bf16 pieces; 3 pieces are exact for
//          every f32, 1-2 are exact for small integers, 2 is the fast
//          split used for real couplings (v_mfma_f32_32x32x16_bf16, f32
//          accumulation).
//
// Randomness is the plain form's: the Philox uniform of (chain, sweep,
// position) and the deterministic-log acceptance test, evaluated through a
// certified single-precision fast path with an exact f64 fallback, so with
// exact couplings the trajectories equal the plain form's bit for bit.
//
// Layout: a workgroup owns a group of 64 replicas = floor(64/R) reads of R
// replicas (replicas q and q+32 form MFMA column tiles 0 and 1), split over
// KW = 1, 2 or 4 waves so that small batches still occupy every SIMD: each
// wave holds the spins of one K range of the problem for all 64 replicas and
// multiplies that range; the partial block fields are summed in LDS, where
// each wave then runs the proposals of 64/KW replicas.  One launch runs all
// sweeps; parallel tempering, best tracking and log-omega run per read
// inside the workgroup after every sweep.
#include "blocked.hpp"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <type_traits>
#include <utility>
#include <vector>

#include <mutex>

#include "../../include/pysa_amd.h"
#include "accept.hpp"
#include "engine.hpp"
#include "rng.hpp"

// Best-state scratch of the allocation-free entry point
// (pysa_amd_metropolis_device_noalloc): one copy per device, held until the
// process ends.  Host code never uses this name as an address: it passes the
// device address that noalloc_scratch() looks up.
static __device__ __align__(256) uint8_t scratch[pysa::kNoAllocScratchBytes];

namespace pysa {
namespace {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(16))) float f32x16;
typedef __attribute__((ext_vector_type(4))) int i32x4;
typedef __attribute__((ext_vector_type(16))) int i32x16;
typedef __attribute__((ext_vector_type(4))) unsigned u32x4v;

constexpr int kB = 32;  // block size (rows of one MFMA tile)

// ---------------------------------------------------------------- packing --
__host__ __device__ inline uint16_t bf16_rn(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7F800000u) == 0x7F800000u) return (uint16_t)(u >> 16);  // inf / nan
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf16_to_f(uint16_t b) {
  uint32_t u = (uint32_t)b << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

// Sum of the first np bf16 pieces of v (hi, then lo of the remainder, ...):
// the coupling value the bf16 products multiply with.  Each remainder is
// exact in f32 and the partial sums carry at most 8 np significant bits, so
// the sum is exact as well.
__host__ __device__ inline float bf16_pieces_sum(float v, int np) {
  float acc = 0.f;
  for (int p = 0; p < np; ++p) {
    const float b = bf16_to_f(bf16_rn(v));
    acc = acc + b;
    v = v - b;
  }
  return acc;
}

// out = the pieces sum of every element of J (n*n).
__global__ void pieces_sum_kernel(const float* __restrict__ J, long long count, int np,
                                  float* __restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) out[i] = bf16_pieces_sum(J[i], np);
}

// Row sums of the couplings the products use (J, or its np-piece sums when
// np > 0) in f64, rounded once (QUBO fields from +-1 products).
__global__ void rowsum_kernel(const float* __restrict__ J, int n, int np, float* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int j = 0; j < n; ++j) {
    const float v = J[(size_t)i * n + j];
    s += (double)(np > 0 ? bf16_pieces_sum(v, np) : v);
  }
  out[i] = (float)s;
}

// x with bit k moved to bit k ^ m (m < 32, wave-uniform): the in-block
// permutation of the block-random order, an involution.
__device__ __forceinline__ uint32_t xor_permute_bits(uint32_t x, uint32_t m) {
  if (m & 1u) x = ((x & 0x55555555u) << 1) | ((x >> 1) & 0x55555555u);
  if (m & 2u) x = ((x & 0x33333333u) << 2) | ((x >> 2) & 0x33333333u);
  if (m & 4u) x = ((x & 0x0F0F0F0Fu) << 4) | ((x >> 4) & 0x0F0F0F0Fu);
  if (m & 8u) x = ((x & 0x00FF00FFu) << 8) | ((x >> 8) & 0x00FF00FFu);
  if (m & 16u) x = (x << 16) | (x >> 16);
  return x;
}

// Fragment (c, step, piece) of the bf16 form: lane l holds
// piece(J[32c + l%32][16*step + 8*(l/32) + e]), e < 8.
__global__ void pack_bf16_kernel(const float* __restrict__ J, int n, int nblk, int np,
                                 uint16_t* __restrict__ out) {
  const long long nfrag = (long long)nblk * (2 * nblk) * np;
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= nfrag * 64) return;
  const int lane = (int)(gid & 63);
  const long long f = gid >> 6;
  const int piece = (int)(f % np);
  const long long cs = f / np;
  const int step = (int)(cs % (2 * nblk));
  const int c = (int)(cs / (2 * nblk));
  const int row = 32 * c + (lane & 31);
  uint16_t* dst = out + gid * 8;
  for (int e = 0; e < 8; ++e) {
    const int col = 16 * step + 8 * (lane >> 5) + e;
    float v = (row < n && col < n) ? J[(size_t)row * n + col] : 0.f;
    uint16_t b = 0;
    for (int p = 0; p <= piece; ++p) {
      b = bf16_rn(v);
      v = v - bf16_to_f(b);
    }
    dst[e] = b;
  }
}

// Fragment (c, step) of the int8 form: lane l holds J[32c + l%32][32*step + 16*(l/32) + e].
__global__ void pack_i8_kernel(const float* __restrict__ J, int n, int nblk, int8_t* __restrict__ out) {
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (long long)nblk * nblk * 64) return;
  const int lane = (int)(gid & 63);
  const long long f = gid >> 6;
  const int step = (int)(f % nblk), c = (int)(f / nblk);
  const int row = 32 * c + (lane & 31);
  int8_t* dst = out + gid * 16;
  for (int e = 0; e < 16; ++e) {
    const int col = 32 * step + 16 * (lane >> 5) + e;
    dst[e] = (row < n && col < n) ? (int8_t)J[(size_t)row * n + col] : (int8_t)0;
  }
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
};

// Workgroup-shared state of one replica group (64 replicas, KW waves).
template <int KW>
struct GroupLds {
  float part[KW][2][32 * 33];  // partial block fields: [wave][tile][replica col * 33 + row]
  uint32_t bb[2][64];          // block spin bits per replica (double-buffered by block parity)
  float jb[kB * kB];           // jb[i*32 + k] = J^T[pos0+i][pos0+k] = J[pos0+k][pos0+i]
  float hb[kB];
  float rsb[kB];  // row sums J[pos0+i] . 1 of the block rows (QUBO, bf16 operands)
  float E[64];
  float rb[64];   // beta of each replica
  int bi[64];     // slot -> replica within its read, at g*R + slot
  int flag[64];   // per-replica improvement flags
  int src[64];    // per-read source replica of a new global best (-1: none)
  double lu[64];  // PT log-uniforms, at g*R + pair
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

// One workgroup = one group of 64 replicas (floor(64/R) reads) split over KW
// waves.  Wave w owns the spins of blocks [w*kb, (w+1)*kb) for all 64
// replicas (MFMA operand layout) and computes the partial block fields over
// that K range; the partials meet in LDS, where wave w then runs the block's
// proposals for replicas [w*64/KW, (w+1)*64/KW), one lane per replica.
template <int NP, bool I8, int NW, int KW>
__global__ __launch_bounds__(64 * KW) void blocked_pt_kernel(BParams p) {
  __shared__ GroupLds<KW> L;
  constexpr int Q = 64 / KW;  // replicas per wave in the proposal phase
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, tid = threadIdx.x;
  const int R = p.R, n = p.n, nblk = p.nblk, kb = p.kb;
  const int read0 = blockIdx.x * p.rpw;
  const int h = lane >> 5;
  // Operand role: tile-0 replica qa and tile-1 replica qb of the group.
  const int qa = lane & 31, qb = 32 + (lane & 31);
  const int ga = qa / R, gb = qb / R;
  const bool acta = (ga < p.rpw) && (read0 + ga < p.reads);
  const bool actb = (gb < p.rpw) && (read0 + gb < p.reads);
  const size_t base_a = acta ? ((size_t)(read0 + ga) * R + (qa - ga * R)) * n : 0;
  const size_t base_b = actb ? ((size_t)(read0 + gb) * R + (qb - gb * R)) * n : 0;
  const int c0 = w * kb;  // first block of this wave's K range
  // Proposal role: the LPR lanes a = 0..LPR-1 of group lane/LPR hold replica
  // q of the group, sub-lane a owning block rows [a*RPL, (a+1)*RPL).
  constexpr int LPR = KW, RPL = kB / KW;
  const int a = lane % LPR;
  const int q = w * Q + lane / LPR;
  const int g = q / R, rr = q - g * R;
  const int read = read0 + g;
  const bool act = (g < p.rpw) && (read < p.reads);

  Bits<NW> sa, sb;
  sa.w = 0;
  sb.w = 0;
  for (int j = 0; j < kb; ++j) {
    const int c = c0 + j;
    if (c >= nblk) break;
    uint32_t va = 0, vb = 0;
#pragma unroll
    for (int b = 0; b < 16; ++b) {
      const int rel = I8 ? (16 * h + b) : ((b < 8) ? (8 * h + b) : (16 + 8 * h + (b - 8)));
      const int pos = 32 * c + rel;
      if (pos < n) {
        if (acta && p.spins[base_a + pos] == 1) va |= 1u << b;
        if (actb && p.spins[base_b + pos] == 1) vb |= 1u << b;
      }
    }
    set_half<NW>(sa, j, va);
    set_half<NW>(sb, j, vb);
  }
  float E = act ? p.energies[(size_t)read * R + rr] : 0.f;
  if (act && a == 0) {
    L.bi[q] = rr;  // beta_idx = arange(R) (sa.py:325)
    L.rb[q] = p.betas[rr];
  }
  float bestE = E;
  if (p.best_sem == 0 && act && a == 0) {
    p.best_rep_e[(size_t)read * R + rr] = E;
    p.best_rep_sw[(size_t)read * R + rr] = 0;
  }
  float gE = E;  // simulation.py:205 (meaningful on read leaders, rr == 0)
  int gsw = 0, gimp = 0;
  const uint64_t cid = (uint64_t)read * R + rr;
  unsigned long long acc = 0;
  __syncthreads();

  // Prefetch pipeline.  Coupling fragments of the wave's K range stream
  // through a ring of PF register stages; the first PF stages of the next
  // block, and its J[blk][blk] / h[blk] share, are requested right after the
  // current block's products so their latency hides behind the proposals.
  constexpr int PF = 4;
  // Operand expansion constants of the spin encoding (see spins_bf16 / spins_i8).
  const uint32_t emul = p.ptype == 0 ? 0xFEu : 1u;
  const uint32_t exr = p.ptype == 0 ? 0xFFFFFFFFu : 0u;
  constexpr int NST = I8 ? 1 : 2;  // fragments per K step of one block column
  constexpr int NJB = (kB * kB + 64 * KW - 1) / (64 * KW);
  typedef typename std::conditional<I8, i32x4, bf16x8>::type Frag;
  Frag ring[PF][NST][NP];
  float jb_reg[NJB], h_reg = 0.f, rs_reg = 0.f;
  const int jn = max(0, min(kb, nblk - c0));
  auto frag_ptr = [&](int c, int j, int st, int pc) -> const Frag* {
    if (I8) return reinterpret_cast<const Frag*>(p.frag) + ((size_t)c * nblk + c0 + j) * 64 + lane;
    return reinterpret_cast<const Frag*>(p.frag) +
           (((size_t)c * (2 * nblk) + 2 * (c0 + j) + st) * NP + pc) * 64 + lane;
  };
  auto load_stage = [&](int c, int j, Frag (&dst)[NST][NP]) {
#pragma unroll
    for (int st = 0; st < NST; ++st)
#pragma unroll
      for (int pc = 0; pc < NP; ++pc) dst[st][pc] = *frag_ptr(c, j, st, pc);
  };
  auto prefetch_block = [&](int c) {
#pragma unroll
    for (int u = 0; u < PF; ++u)
      if (u < jn) load_stage(c, u, ring[u]);
    const int pos0 = 32 * c;
#pragma unroll
    for (int e2 = 0; e2 < NJB; ++e2) {
      const int e = tid + e2 * 64 * KW;
      const int i = e >> 5, k = e & 31;
      jb_reg[e2] = (e < kB * kB && pos0 + i < n && pos0 + k < n)
                       ? p.Jt[(size_t)(pos0 + i) * n + pos0 + k] : 0.f;
    }
    h_reg = (tid < kB && pos0 + tid < n) ? p.h[pos0 + tid] : 0.f;
    if (!I8 && p.rowsum) rs_reg = (tid < kB && pos0 + tid < n) ? p.rowsum[pos0 + tid] : 0.f;
  };
  prefetch_block(0);

  for (int s = 0; s < p.S; ++s) {
    const uint64_t sw = (uint64_t)s + p.sweep_offset;
    const float beta = act ? L.rb[q] : 0.f;
    const float cu = (p.ptype == 0) ? 2.f : 1.f;  // |dn / (F + h)|
    const float inv_b = (beta > 0.f) ? 1.0f / beta : 0.f;
    double dsum = 0.0;
    uint32_t nacc = 0;
    for (int c = 0; c < nblk; ++c) {
      const int buf = c & 1, oc = c / kb, pos0 = 32 * c;
      // ---- phase 1a: publish the prefetched J[blk][blk] (as J^T rows) and h[blk]
#pragma unroll
      for (int e2 = 0; e2 < NJB; ++e2) {
        const int e = tid + e2 * 64 * KW;
        if (e < kB * kB) L.jb[e] = jb_reg[e2];
      }
      if (tid < kB) {
        L.hb[tid] = h_reg;
        L.rsb[tid] = rs_reg;
      }
      // ---- phase 1b: the owner publishes the block's spins of all 64 replicas
      if (w == oc) {
        const int j = c - c0;
        const uint32_t mine = (lane < 32) ? get_half<NW>(sa, j) : get_half<NW>(sb, j);
        const uint32_t give = (lane < 32) ? get_half<NW>(sb, j) : get_half<NW>(sa, j);
        const uint32_t got = (uint32_t)__shfl_xor((int)give, 32, 64);
        L.bb[buf][lane] = (lane < 32) ? block_of<I8>(mine, got) : block_of<I8>(got, mine);
      }
      // ---- phase 1c: partial block fields over this wave's K range
      f32x16 acc0 = {}, acc1 = {};
      i32x16 iacc0 = {}, iacc1 = {};
      constexpr int kUnrollJ = NW <= 8 ? (2 * NW + PF - 1) / PF : 1;
#pragma unroll kUnrollJ
      for (int jb0 = 0; jb0 < 2 * NW; jb0 += PF) {
#pragma unroll
        for (int u = 0; u < PF; ++u) {
          const int j = jb0 + u;
          if (j >= jn) continue;
          Frag fa[NST][NP];
#pragma unroll
          for (int st = 0; st < NST; ++st)
#pragma unroll
            for (int pc = 0; pc < NP; ++pc) fa[st][pc] = ring[u][st][pc];
          if (j + PF < jn) load_stage(c, j + PF, ring[u]);
          const uint32_t ha = get_half<NW>(sa, j), hb = get_half<NW>(sb, j);
          if (I8) {
            uint32_t da[4], db[4];
            spins_i8(ha, emul, exr, da);
            spins_i8(hb, emul, exr, db);
            i32x4 ba, bb;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
              ba[e] = (int)da[e];
              bb[e] = (int)db[e];
            }
            const i32x4 a8 = __builtin_bit_cast(i32x4, fa[0][0]);
            iacc0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a8, ba, iacc0, 0, 0, 0);
            iacc1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a8, bb, iacc1, 0, 0, 0);
          } else {
            const uint32_t nza = spins_nz(ha), nzb = spins_nz(hb);
#pragma unroll
            for (int st = 0; st < NST; ++st) {
              uint32_t da[4], db[4];
              if (st == 0) {
                spins_bf16<0>(nza, da);
                spins_bf16<0>(nzb, db);
              } else {
                spins_bf16<1>(nza, da);
                spins_bf16<1>(nzb, db);
              }
              u32x4v ua = {da[0], da[1], da[2], da[3]}, ub = {db[0], db[1], db[2], db[3]};
              const bf16x8 ba = __builtin_bit_cast(bf16x8, ua), bb = __builtin_bit_cast(bf16x8, ub);
#pragma unroll
              for (int pc = 0; pc < NP; ++pc) {
                const bf16x8 a16 = __builtin_bit_cast(bf16x8, fa[st][pc]);
                acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a16, ba, acc0, 0, 0, 0);
                acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a16, bb, acc1, 0, 0, 0);
              }
            }
          }
        }
      }
      prefetch_block(c + 1 < nblk ? c + 1 : 0);
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int i = 8 * (v >> 2) + 4 * h + (v & 3);
        L.part[w][0][(lane & 31) * 33 + i] = I8 ? (float)iacc0[v] : acc0[v];
        L.part[w][1][(lane & 31) * 33 + i] = I8 ? (float)iacc1[v] : acc1[v];
      }
      __syncthreads();
      // ---- phase 2: the block's proposals in sequential order.  Replica q's
      // LPR lanes each hold RPL of the 32 block fields; the lane owning row
      // k decides, broadcasts dS, and every lane updates its rows from row k
      // of J^T[blk][blk].
      {
        const int tq = q >> 5, jq = q & 31, r0 = a * RPL;
        // fp = F + h of this lane's RPL rows
        float fp[RPL];
#pragma unroll
        for (int rl = 0; rl < RPL; ++rl) {
          float t = 0.f;
#pragma unroll
          for (int ww = 0; ww < KW; ++ww) t += L.part[ww][tq][jq * 33 + r0 + rl];
          // QUBO with bf16 operands: the product used +-1 spins, F = (J.s + J.1) / 2
          if (!I8 && p.ptype == 1) t = 0.5f * (t + L.rsb[r0 + rl]);
          fp[rl] = t + L.hb[r0 + rl];
        }
        const uint32_t all = L.bb[buf][q];
        uint32_t mb = (RPL == 32) ? all : ((all >> r0) & ((1u << RPL) - 1u));
        const uint32_t mb0 = mb;
        uint32_t uw[RPL];
#pragma unroll
        for (int m = 0; m < RPL / 4; ++m) {
          const u32x4 blk = sweep_block(cid, sw, (uint32_t)(pos0 + r0 + 4 * m), p.k0, p.k1);
          uw[4 * m] = blk.x;
          uw[4 * m + 1] = blk.y;
          uw[4 * m + 2] = blk.z;
          uw[4 * m + 3] = blk.w;
        }
        // Acceptance brackets of the lane's rows, off the sequential chain.
        float tlo[RPL], thi[RPL];
#pragma unroll
        for (int rl = 0; rl < RPL; ++rl) {
          accept_bounds(uw[rl], beta, inv_b, tlo[rl], thi[rl]);
          if (pos0 + r0 + rl >= n) tlo[rl] = thi[rl] = INFINITY;  // padding: never taken
        }
        float dpart = 0.f;  // this lane's accepted energy changes (exact for integer couplings)
#pragma unroll
        for (int k = 0; k < kB; ++k) {
          {
            const int ak = k / RPL, kl = k % RPL;
            const bool up = (mb >> kl) & 1u;
            // dn = 2 s (F + h) (ising) or (2x - 1)(F + h) (qubo), exact in f32
            const float sg = up ? cu : -cu;
            const float dn = fp[kl] * sg;
            const bool own = act && (a == ak);
            bool take = dn > thi[kl];
            if (__builtin_expect(own & !take & !(dn < tlo[kl]), 0))
              take = accept_exact(uw[kl], beta, dn);
            take = take & own;
            mb ^= (uint32_t)take << kl;
            dpart += take ? -dn : 0.f;
            const float dS = group_bcast<LPR>(take ? -sg : 0.f, ak);
            const float4* jr = reinterpret_cast<const float4*>(&L.jb[k * kB + r0]);
#pragma unroll
            for (int m = 0; m < RPL / 4; ++m) {
              const float4 j4 = jr[m];
              // dS is 0, +-1 or +-2, so the fused form rounds like mul + add
              fp[4 * m] = __builtin_fmaf(j4.x, dS, fp[4 * m]);
              fp[4 * m + 1] = __builtin_fmaf(j4.y, dS, fp[4 * m + 1]);
              fp[4 * m + 2] = __builtin_fmaf(j4.z, dS, fp[4 * m + 2]);
              fp[4 * m + 3] = __builtin_fmaf(j4.w, dS, fp[4 * m + 3]);
            }
          }
        }
        dsum += (double)dpart;
        nacc += (uint32_t)__builtin_popcount(mb ^ mb0);
        if (RPL == 32)
          L.bb[buf][q] = mb;
        else if (RPL == 16)
          reinterpret_cast<uint16_t*>(L.bb[buf])[q * 2 + a] = (uint16_t)mb;
        else
          reinterpret_cast<uint8_t*>(L.bb[buf])[q * 4 + a] = (uint8_t)mb;
      }
      __syncthreads();
      // ---- phase 3: the owner takes the new spins back into operand layout
      if (w == oc) {
        const int j = c - c0;
        set_half<NW>(sa, j, half_of<I8>(L.bb[buf][qa], h));
        set_half<NW>(sb, j, half_of<I8>(L.bb[buf][qb], h));
      }
    }
    // ---- end of sweep: energy (simulation.py:157-158) and best states
    acc += nacc;
#pragma unroll
    for (int o = 1; o < LPR; o <<= 1) dsum += __shfl_xor(dsum, o, 64);
    E = (float)((double)E + dsum);
    if (act && a == 0) L.E[q] = E;
    if (p.best_sem == 0) {  // simulation.py:161-164
      const bool imp = act && (E < bestE);
      if (imp) {
        bestE = E;
        if (a == 0) {
          p.best_rep_e[(size_t)read * R + rr] = E;
          p.best_rep_sw[(size_t)read * R + rr] = s;
        }
      }
      if (a == 0) L.flag[q] = imp ? 1 : 0;
    }
    // PT log-uniforms, one lane per pair of each read (drawn ahead of the walk).
    if (p.use_pt && act && a == 0 && rr < R - 1)
      L.lu[q] = log_u32(pick_word(pt_block((uint64_t)read, sw, (uint32_t)rr, p.k0, p.k1),
                                  (uint32_t)rr & 3u));
    __syncthreads();
    if (p.best_sem == 0) {
      if (acta && L.flag[qa]) store_range<NW, I8>(sa, c0, kb, nblk, n, p.ptype, p.best_rep + base_a);
      if (actb && L.flag[qb]) store_range<NW, I8>(sb, c0, kb, nblk, n, p.ptype, p.best_rep + base_b);
    }
    if (act && rr == 0 && a == 0) {  // read leader
      const int b0 = g * R;
      int src = -1;
      if (p.best_sem == 1) {  // simulation.py:222-225 in replica order
        float cur = gE;
        for (int k = 0; k < R; ++k)
          if (L.E[b0 + k] < cur) {
            cur = L.E[b0 + k];
            src = k;
          }
        if (src >= 0) {
          gE = cur;
          gsw = s;
          gimp = 1;
        }
      }
      if (p.use_pt) {  // utils.py:44-58
        for (int k = 0; k < R - 1; ++k) {
          const int a = R - k - 1, b = R - k - 2;
          const float dd = (L.E[b0 + L.bi[b0 + a]] - L.E[b0 + L.bi[b0 + b]]) *
                           (p.betas[a] - p.betas[b]);
          if (dd >= 0.f || L.lu[b0 + k] < (double)dd) {
            const int t = L.bi[b0 + a];
            L.bi[b0 + a] = L.bi[b0 + b];
            L.bi[b0 + b] = t;
          }
        }
      }
      for (int k = 0; k < R; ++k) L.rb[b0 + L.bi[b0 + k]] = p.betas[k];
      if (p.logw && p.log_omegas) {  // ais.py:99-112 over the ascending ladder
        double lw = 0.0, bold = 0.0;
        for (int r = 1; r < R; ++r) {
          const int k = p.bsort[r];
          const float be = p.betas[k];
          const float en = L.E[b0 + L.bi[b0 + k]];
          const float a = -be * en;
          lw += (double)a - (-bold * (double)en);
          bold = (double)be;
        }
        p.log_omegas[(size_t)read * p.S + s] = lw;
      }
      L.src[g] = src;
    }
    __syncthreads();
    if (p.best_sem == 1) {  // every owner wave copies its part of the new best state
      if (acta && L.src[ga] == qa - ga * R)
        store_range<NW, I8>(sa, c0, kb, nblk, n, p.ptype, p.best_state + (size_t)(read0 + ga) * n);
      if (actb && L.src[gb] == qb - gb * R)
        store_range<NW, I8>(sb, c0, kb, nblk, n, p.ptype, p.best_state + (size_t)(read0 + gb) * n);
    }
    __syncthreads();
  }

  // ---- outputs
  if (acta) store_range<NW, I8>(sa, c0, kb, nblk, n, p.ptype, p.spins + base_a);
  if (actb) store_range<NW, I8>(sb, c0, kb, nblk, n, p.ptype, p.spins + base_b);
  if (act && p.accepted) atomicAdd(p.accepted + read, acc);
  if (act && a == 0) {
    p.energies[(size_t)read * R + rr] = E;
    p.beta_idx[(size_t)read * R + rr] = L.bi[q];
    if (rr == 0 && p.best_sem == 1) {
      p.best_energy[read] = gE;
      p.best_sweeps[read] = gsw;
      L.src[g] = gimp;
    }
  }
  if (p.best_sem == 1) {
    // never improved: best_state is still the view of states[0] (simulation.py:204)
    __syncthreads();
    if (acta && qa - ga * R == 0 && !L.src[ga])
      store_range<NW, I8>(sa, c0, kb, nblk, n, p.ptype, p.best_state + (size_t)(read0 + ga) * n);
    if (actb && qb - gb * R == 0 && !L.src[gb])
      store_range<NW, I8>(sb, c0, kb, nblk, n, p.ptype, p.best_state + (size_t)(read0 + gb) * n);
  }
}

// ------------------------------------------------------- pipelined form --
// Eight waves per group of 64 replicas, two per SIMD.  Wave 0 runs the
// proposals with one lane per replica: the lane holds all 32 fields of the
// block, so a decision needs no cross-lane broadcast.  Six matrix waves, a
// third of the K range times a tile of 32 replicas each, compute the next
// block's fields over every K block except c (whose spins are changing)
// while wave 0 sweeps block c:
//     P_{c+1} = sum_{b != c} J[c+1, b] . S_b
// When wave 0 is done with block c, the owners of block c take the new spins
// and add the missing product J[c+1, c] . S_c (one K block on the matrix
// cores, from fragments kept back while computing P_{c+1}); the fields of
// block c+1 are P_{c+1} plus that correction.  The matrix products of one
// block therefore run under the proposals of the previous one.  The matrix
// waves (and wave 4, which shares wave 0's SIMD) also prepare the next
// block's acceptance brackets, so the chain only compares, flips and updates.
// The two waves of a K range share a SIMD: gfx950 does not overlap a wave's
// own VALU and MFMA instructions, but it does overlap different waves'.
// Shared state alternates between two buffers by the parity of a running
// block counter.
// K ranges of the pipelined kernel (each served by two matrix waves).
constexpr int kPipeMW = 3;
constexpr int kPipePF = 3;  // fragment stages in flight per matrix wave (slot ring)

// A replica's 32 fields take 32 floats; its 16-byte row groups are XOR-
// swizzled by the replica index so that 8 lanes of a 16-byte access hit all
// 32 banks.
__device__ __forceinline__ void swap2(float& a, float& b) {
  const float t = a;
  a = b;
  b = t;
}
// Field layout [replica][row]: the 16-byte groups of four rows are XOR
// swizzled by the replica so that a 16-byte access by consecutive replicas
// hits distinct banks in every aligned run of 8 lanes (32 banks) and of 16
// lanes (64 banks: consecutive rows of 128 bytes alternate bank halves).
__device__ __forceinline__ int fld(int q, int r) {
  return q * 32 + 4 * ((r >> 2) ^ (q & 7) ^ ((q >> 3) & 1)) + (r & 3);
}
constexpr int kHL = 36;
constexpr int kFG = 2;  // row groups per stage of the proposal wave's field sums
template <int MW>
struct __align__(16) PipeLds {
  float part[2][MW][64 * 32];  // [parity][matrix wave][fld(replica, row)]
  float corr[64 * 32];         // J[c+1, c] . S_c of the next block
  float jb[2][kB * kB];       // row i of J[blk][blk] right of the diagonal, shifted left:
                              // J^T[pos0+i][pos0+i+1+t] at i*32 + t (t < 31 - i), else 0
  float hb[2][kB];
  float rsb[2][kB];
  // acceptance brackets [parity][replica][visit index], rows padded to
  // kHL floats so that 16-byte reads of consecutive replicas hit distinct banks
  float lo[2][64][kHL];
  float hi[2][64][kHL];
  uint32_t bb[2][64];         // block spins before the block (bit r = row r)
  uint32_t bnew[2][64];       // block spins after the block (by block parity)
  float sfp[64 * 33];         // exact rerun staging
  // visit schedule by pass parity: block at step t and its in-block XOR mask
  uint8_t ord[2][64];
  uint8_t msk[2][64];
  float E[64];
  float rb[64];
  int bi[64];
  int flag[64];
  int src[64];
  double lu[64];
};

// BR: block-random visit order (schedule tables in LDS); otherwise blocks are
// visited in order with no in-block shuffle, and every schedule entry is
// plain arithmetic on the step.
template <int NP, bool I8, int NW, int MW, bool BR>
__global__ __launch_bounds__(128 * (MW + 1)) void blocked_pipe_kernel(BParams p) {
  __shared__ PipeLds<MW> L;
  const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), tid = threadIdx.x;
  const int R = p.R, n = p.n, nblk = p.nblk, kb = p.kb;
  const int read0 = blockIdx.x * p.rpw;
  const int h = lane >> 5;
  // Two waves per SIMD (SIMD w & 3).  Wave 0 runs the proposals; matrix
  // wave w = 1..MW computes the products of K range m = w - 1 for both tiles
  // of 32 replicas (each coupling fragment is loaded once for both); the
  // helper waves MW+1.. stage the next blocks' J block, h and row sums,
  // compute the acceptance brackets (those on SIMDs 1..MW) and, wave MW+1,
  // the visit schedule.
  static_assert(MW == 3, "eight waves: the proposal wave, three matrix waves, four helpers");
  const bool mw = w >= 1 && w <= MW;  // a matrix wave
  const bool hw = w > MW;             // a helper wave
  const int m = mw ? w - 1 : 0;
  const int hl = hw ? (w - MW - 1) * 64 + lane : -1;  // helper-lane index 0..64 (MW+1) - 1
  // Operand role (matrix waves): replica qtt[tt] of tile tt.
  int qtt[2], gtt[2];
  bool actt[2];
  size_t base_t[2];
#pragma unroll
  for (int tt = 0; tt < 2; ++tt) {
    qtt[tt] = tt * 32 + (lane & 31);
    gtt[tt] = qtt[tt] / R;
    actt[tt] = mw && (gtt[tt] < p.rpw) && (read0 + gtt[tt] < p.reads);
    base_t[tt] = actt[tt] ? ((size_t)(read0 + gtt[tt]) * R + (qtt[tt] - gtt[tt] * R)) * n : 0;
  }
  // Replica role (every wave): lane = replica q of the group.
  const int q = lane;
  const int g = q / R, rr = q - g * R;
  const int read = read0 + g;
  const bool act = (g < p.rpw) && (read < p.reads);
  const uint64_t cid = (uint64_t)read * R + rr;
  const int c0 = mw ? m * kb : 0;
  const int jn = mw ? max(0, min(kb, nblk - c0)) : 0;

  Bits<NW> sa[2];  // this wave's K range of each tile's spins, operand layout
#pragma unroll
  for (int tt = 0; tt < 2; ++tt) {
    sa[tt].w = 0;
    for (int j = 0; j < jn; ++j) {
      const int c = c0 + j;
      uint32_t va = 0;
#pragma unroll
      for (int b = 0; b < 16; ++b) {
        const int rel = I8 ? (16 * h + b) : ((b < 8) ? (8 * h + b) : (16 + 8 * h + (b - 8)));
        const int pos = 32 * c + rel;
        if (pos < n && actt[tt] && p.spins[base_t[tt] + pos] == 1) va |= 1u << b;
      }
      set_half<NW>(sa[tt], j, va);
    }
  }
  float E = 0.f, bestE = 0.f, gE = 0.f;
  int gsw = 0, gimp = 0;
  unsigned long long acc = 0;
  if (w == 0) {
    E = act ? p.energies[(size_t)read * R + rr] : 0.f;
    L.bi[q] = act ? rr : 0;  // beta_idx = arange(R) (sa.py:325)
    L.rb[q] = act ? p.betas[rr] : 0.f;
    bestE = E;
    gE = E;  // simulation.py:205 (meaningful on read leaders, rr == 0)
    if (p.best_sem == 0 && act) {
      p.best_rep_e[(size_t)read * R + rr] = E;
      p.best_rep_sw[(size_t)read * R + rr] = 0;
    }
  }
  for (int i = tid; i < 64 * 32; i += 128 * (MW + 1)) L.corr[i] = 0.f;

  constexpr int PF = kPipePF;
  const uint32_t emul = p.ptype == 0 ? 0xFEu : 1u;
  const uint32_t exr = p.ptype == 0 ? 0xFFFFFFFFu : 0u;
  // operand-expansion constants in vector registers (see spins_bf16_r)
  uint32_t vmask = 0x80008000u, vone = 0x3F803F80u;
  asm volatile("" : "+v"(vmask), "+v"(vone));
  constexpr int NST = I8 ? 1 : 2;
  constexpr int NHL = 64 * (MW + 1);  // helper lanes
  constexpr int NJB = (kB * kB + NHL - 1) / NHL;
  typedef typename std::conditional<I8, i32x4, bf16x8>::type Frag;
  Frag ring[PF][NST][NP];
  float jb_reg[NJB], h_reg = 0.f, rs_reg = 0.f;
  uint32_t jb_ok = 0;  // bit e2: jb_reg[e2] is a coupling; bit 31: h_reg / rs_reg are
  auto frag_ptr = [&](int c, int j, int st, int pc) -> const Frag* {
    if (I8) return reinterpret_cast<const Frag*>(p.frag) + ((size_t)c * nblk + c0 + j) * 64 + lane;
    return reinterpret_cast<const Frag*>(p.frag) +
           (((size_t)c * (2 * nblk) + 2 * (c0 + j) + st) * NP + pc) * 64 + lane;
  };
  // Stage j of this wave's K range; stages past the range reload the
  // range's last K block (recently read, so a cache hit), so every path
  // issues the same loads and the wait counts stay exact (a skipped load
  // would make the compiler wait for all of them).
  // One K block of fragments: NST * NP * 64 of them, consecutive.
  constexpr int kStage = NST * NP * 64;
  auto load_stage = [&](int c, int j, Frag (&dst)[NST][NP]) {
    // jn passes through an empty asm so that the per-stage offsets are
    // recomputed (a few scalar ops) rather than hoisted out of the block
    // loop into registers that would spill
    int jnl = __builtin_amdgcn_readfirstlane(jn);
    asm volatile("" : "+s"(jnl));
    const int jc = jnl > 0 ? min(j, jnl - 1) : nblk - 1 - c0;
    // wave-uniform base of the stage, then a per-lane 16-byte offset
    const Frag* base = frag_ptr(c, 0, 0, 0) - lane + (size_t)jc * kStage;
#pragma unroll
    for (int st = 0; st < NST; ++st)
#pragma unroll
      for (int pc = 0; pc < NP; ++pc) dst[st][pc] = base[(st * NP + pc) * 64 + lane];
  };
  // Request row block c: the first PF fragment stages of this wave's K range
  // (matrix waves).
  auto prefetch_frags = [&](int c) {
#pragma unroll
    for (int u = 0; u < PF; ++u) load_stage(c, u, ring[u]);
  };
  // Request row block c, visited with in-block mask xm, spread over the
  // helper lanes: J[blk][blk] in visit order, h and row sums.
  auto prefetch_jb = [&](int c, uint32_t xm) {
    const int pos0 = 32 * c;
    // Every lane loads (an in-bounds address when its element is padding)
    // and publish_block zeroes the padding: no register is written here
    // except by these loads, which keeps them from waiting on older loads.
    jb_ok = 0;
#pragma unroll
    for (int e2 = 0; e2 < NJB; ++e2) {
      const int e = hl + e2 * NHL;
      // slot t = e & 31 of visit row i holds visit column k = i + 1 + t;
      // visit index v is block row v ^ xm
      const int i = e >> 5, k = i + 1 + (e & 31);
      const int ri = i ^ (int)xm, rk = k ^ (int)xm;
      const bool ok = e >= 0 && e < kB * kB && k < kB && pos0 + ri < n && pos0 + rk < n;
      jb_ok |= (uint32_t)ok << e2;
      jb_reg[e2] = p.Jt[ok ? (size_t)(pos0 + ri) * n + pos0 + rk : 0];
    }
    const bool hok = hl >= 0 && hl < kB && pos0 + hl < n;
    jb_ok |= (uint32_t)hok << 31;
    h_reg = p.h[hok ? pos0 + hl : 0];
    if (!I8 && p.rowsum) rs_reg = p.rowsum[hok ? pos0 + hl : 0];
  };
  auto publish_block = [&](int pb) {
#pragma unroll
    for (int e2 = 0; e2 < NJB; ++e2) {
      const int e = hl + e2 * NHL;
      if (hl >= 0 && e < kB * kB) L.jb[pb][e] = ((jb_ok >> e2) & 1u) ? jb_reg[e2] : 0.f;
    }
    if (hl >= 0 && hl < kB) {
      const bool hok = (jb_ok >> 31) & 1u;
      L.hb[pb][hl] = hok ? h_reg : 0.f;
      L.rsb[pb][hl] = (hok && !I8 && p.rowsum) ? rs_reg : 0.f;
    }
  };
  // Products of K step j (fragments fa) with this tile's spins hx.
  auto mfma_step = [&](const Frag (&fa)[NST][NP], uint32_t hx, f32x16& acc, i32x16& iacc) {
    if (I8) {
      uint32_t d[4];
      spins_i8(hx, emul, exr, d);
      i32x4 bv;
#pragma unroll
      for (int e = 0; e < 4; ++e) bv[e] = (int)d[e];
      const i32x4 a8 = __builtin_bit_cast(i32x4, fa[0][0]);
      iacc = __builtin_amdgcn_mfma_i32_32x32x32_i8(a8, bv, iacc, 0, 0, 0);
    } else {
      const uint32_t nz = spins_nz(hx);
#pragma unroll
      for (int st = 0; st < NST; ++st) {
        uint32_t d[4];
        if (st == 0)
          spins_bf16_r<0>(nz, vmask, vone, d);
        else
          spins_bf16_r<1>(nz, vmask, vone, d);
        u32x4v ud = {d[0], d[1], d[2], d[3]};
        const bf16x8 bv = __builtin_bit_cast(bf16x8, ud);
#pragma unroll
        for (int pc = 0; pc < NP; ++pc)
          acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, fa[st][pc]), bv,
                                                        acc, 0, 0, 0);
      }
    }
  };
  // Fields from the C layout (register v: row 8(v/4) + 4h + v%4 of column
  // lane%32) into [replica][row], four consecutive rows per 16-byte store.
  auto write_fields_col = [&](float* dst, const f32x16& acc, const i32x16& iacc, int qcol) {
#pragma unroll
    for (int g4 = 0; g4 < 4; ++g4) {
      const int i = 8 * g4 + 4 * h;
      float4 x;
      if (I8)
        x = make_float4((float)iacc[4 * g4], (float)iacc[4 * g4 + 1], (float)iacc[4 * g4 + 2],
                        (float)iacc[4 * g4 + 3]);
      else
        x = make_float4(acc[4 * g4], acc[4 * g4 + 1], acc[4 * g4 + 2], acc[4 * g4 + 3]);
      *reinterpret_cast<float4*>(&dst[fld(qcol, i)]) = x;
    }
  };

  // Partial fields of row block cr over this wave's K range, skipping K block
  // ex (its product is added by the proposal wave once the block's spins are
  // final); ring holds the first PF stages of row block cr on entry.
  auto compute_part = [&](int cr, int ex, float* dst) {
    f32x16 acc[2] = {};
    i32x16 iacc[2] = {};
    const int jx = ex - c0;  // the skipped stage, if it is in this wave's range
    constexpr int kUnrollJ = NW <= 8 ? (2 * NW + PF - 1) / PF : 1;
#pragma unroll kUnrollJ
    for (int jb0 = 0; jb0 < 2 * NW; jb0 += PF) {
      if (jb0 >= jn) break;
      // stages left from this chunk on, recomputed per chunk (a scalar
      // compare a slot instead of per-slot lane masks kept across the loop)
      int jrem = __builtin_amdgcn_readfirstlane(jn - jb0);
      asm volatile("" : "+s"(jrem));
#pragma unroll
      for (int u = 0; u < PF; ++u) {
        const int j = jb0 + u;
        // consume slot u, then refill it with stage j + PF (no copy of the
        // fragments: the load is issued after the products that read them);
        // slots past the 2 NW spin halves exist only when PF does not divide
        // 2 NW and are never live
        if (j < 2 * NW && u < jrem && j != jx) {
          mfma_step(ring[u], get_half<NW>(sa[0], j), acc[0], iacc[0]);
          mfma_step(ring[u], get_half<NW>(sa[1], j), acc[1], iacc[1]);
        }
        load_stage(cr, j + PF, ring[u]);
      }
    }
    write_fields_col(dst, acc[0], iacc[0], qtt[0]);
    write_fields_col(dst, acc[1], iacc[1], qtt[1]);
  };
  // Acceptance brackets of row block cr (in-block mask xm) for sweep counter
  // sw, stored by visit index, lane = replica (beta from L.rb): groups of 4
  // rows (one Philox block each) dealt over the helpers on SIMDs 1..MW, so
  // that none lands on the SIMD of the proposal chain.
  auto brackets = [&](int cr, uint32_t xm, uint64_t sw, int pb) {
    float cb, cz;
    bracket_factors(L.rb[q], cb, cz);
    const int pos0 = 32 * cr;
    for (int g4 = w - MW - 2; g4 < kB / 4; g4 += MW) {
      const u32x4 blk = sweep_block(cid, sw, (uint32_t)(pos0 + 4 * g4), p.k0, p.k1);
      const uint32_t wd[4] = {blk.x, blk.y, blk.z, blk.w};
      float lo[4], hi[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        accept_bounds_f(wd[e], cb, cz, lo[e], hi[e]);
        if (pos0 + 4 * g4 + e >= n || !act) lo[e] = hi[e] = INFINITY;  // padding, idle lanes
      }
      // rows 4 g4 + e are visits 4 (g4 ^ xm / 4) + (e ^ xm % 4): one 16-byte
      // store each, its entries exchanged by the uniform low bits of xm
      if (xm & 1u) {
        swap2(lo[0], lo[1]), swap2(lo[2], lo[3]);
        swap2(hi[0], hi[1]), swap2(hi[2], hi[3]);
      }
      if (xm & 2u) {
        swap2(lo[0], lo[2]), swap2(lo[1], lo[3]);
        swap2(hi[0], hi[2]), swap2(hi[1], hi[3]);
      }
      const int v4 = 4 * (g4 ^ (int)(xm >> 2));
      *reinterpret_cast<float4*>(&L.lo[pb][q][v4]) = make_float4(lo[0], lo[1], lo[2], lo[3]);
      *reinterpret_cast<float4*>(&L.hi[pb][q][v4]) = make_float4(hi[0], hi[1], hi[2], hi[3]);
    }
  };
  auto owner_of = [&](int c) { return 1 + c / kb; };  // matrix wave owning K block c
  auto publish_bits = [&](int c, int pb) {  // by the owner of block c
    const int j = c - c0;
#pragma unroll
    for (int tt = 0; tt < 2; ++tt) {
      const uint32_t mine = get_half<NW>(sa[tt], j);
      const uint32_t got = (uint32_t)__shfl_xor((int)mine, 32, 64);  // the other half
      if (lane < 32) L.bb[pb][qtt[tt]] = block_of<I8>(mine, got);
    }
  };

  // Visit schedule of sweep counter swn into parity par (wave 4, one lane
  // per step): the block visited at step t and that block's in-block mask.
  auto fill_sched = [&](int par, uint64_t swn, bool dry) {
    if (lane < nblk) {
      int bo = lane;
      uint32_t mk = 0;
      if (p.order == kBlockRandom && !dry) {
        const FeistelKey fk = block_order_key((uint32_t)nblk, swn, p.k0, p.k1);
        bo = (int)feistel_perm(fk, (uint32_t)lane);
        mk = block_mask(swn, (uint32_t)bo, p.k0, p.k1);
      }
      L.ord[par][lane] = (uint8_t)bo;
      L.msk[par][lane] = (uint8_t)mk;
    }
  };
  // Reads are volatile so that each use reloads the entry (a short LDS read)
  // instead of keeping it live in scalar registers across the block's work.
  auto sched_block = [&](int par, int step) {
    if (!BR) return step;
    return __builtin_amdgcn_readfirstlane((int)*(volatile uint8_t*)&L.ord[par][step]);
  };
  auto sched_mask = [&](int par, int step) {
    if (!BR) return 0u;
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)*(volatile uint8_t*)&L.msk[par][step]);
  };

  // Passes over the blocks.  Every `refresh` sweeps (p.refresh > 0 and
  // S > refresh) a dry pass recomputes the energy: the tracked energy adds
  // each sweep's accepted changes (simulation.py:157-158), and each change
  // carries the f32 rounding of the matrix-product fields, which
  // random-walks over millions of flips (the reference recomputes each field
  // from its coupling row).  A dry pass runs the pipeline over every block
  // with the proposals switched off, so each block's fields are the full
  // products F = J s of the current state (with the couplings the products
  // use); wave 0 sums E = s.(F/2 + h) in f64 and resets the tracked energy
  // to it plus the offset the caller's initial energy had from the dry pass
  // before the first sweep (zero unless the caller passed other energies).
  // Integer couplings are exact either way.  Passes: a dry pass, then
  // `refresh` sweeps and a dry pass, repeated; no dry pass after the last
  // sweep.  pass_info gives a pass's kind and its sweep (for a dry pass the
  // sweep that follows it).
  const int kre = p.refresh;
  const bool rec = kre > 0 && p.S > kre;
  const int npass = rec ? p.S + 1 + (p.S - 1) / kre : p.S;
  auto pass_info = [&](int ps, int& s_out) -> bool {
    if (!rec) {
      s_out = ps;
      return false;
    }
    if (ps == 0) {
      s_out = 0;
      return true;
    }
    const int qq = ps - 1, bq = qq / (kre + 1), rq = qq - bq * (kre + 1);
    if (rq == kre) {
      s_out = (bq + 1) * kre;
      return true;
    }
    s_out = bq * kre + rq;
    return false;
  };

  // ---- prologue: the fields of the first block in full, its brackets and spins
  {
    int s0;
    const bool dry0 = pass_info(0, s0);
    if (BR && w == 4) fill_sched(0, (uint64_t)s0 + p.sweep_offset, dry0);
  }
  __syncthreads();  // wave 0's betas (L.rb), the zeroed L.corr and the schedule are visible
  {
    int s0;
    (void)pass_info(0, s0);
    const uint64_t sw0 = (uint64_t)s0 + p.sweep_offset;
    const int cf = sched_block(0, 0), cs = sched_block(0, 1);
    const uint32_t mf = sched_mask(0, 0), ms = sched_mask(0, 1);
    if (mw) {
      prefetch_frags(cf);
      compute_part(cf, -1, L.part[0][m]);
      prefetch_frags(cs);
      if (w == owner_of(cf)) publish_bits(cf, 0);
    }
    if (hw) {
      prefetch_jb(cf, mf);
      publish_block(0);
      prefetch_jb(cs, ms);
      if (w > MW + 1) brackets(cf, mf, sw0, 0);
    }
  }
  __syncthreads();

  const float cu = (p.ptype == 0) ? 2.f : 1.f;  // |dn / (F + h)|
  // F = fsc * (sum of the partial products + rsb) + h: QUBO with bf16
  // operands used +-1 spins, F = (J.s + J.1) / 2 with the row sums J.1 in
  // rsb (zero otherwise); fsc is a power of two, so the fma rounds like the add
  const bool qsum = !I8 && p.ptype == 1;
  const float fsc = qsum ? 0.5f : 1.f;
  double eoff = 0.0;  // wave 0: caller's initial energy minus its recomputation

  // The correction J[c1, c] . S_c of the next block c1 by the proposal wave,
  // once block c's spins are final (on SIMD 0's matrix core, idle otherwise):
  // bits = block c's spins of replica `lane` in block-row order; the A
  // fragments of row block c1, K block c were loaded into fa at the block's
  // start.  Result to L.corr (read by this wave at the next block's start).
  typedef Frag FragSet[NST][NP];
  auto load_corr_frags = [&](int c1, int c, FragSet& fa) {
    const Frag* fb = reinterpret_cast<const Frag*>(p.frag);
#pragma unroll
    for (int st = 0; st < NST; ++st)
#pragma unroll
      for (int pc = 0; pc < NP; ++pc) {
        const size_t f = I8 ? (size_t)c1 * nblk + c
                            : ((size_t)c1 * (2 * nblk) + 2 * c + st) * NP + pc;
        fa[st][pc] = fb[f * 64 + lane];
      }
  };
  auto corr_products = [&](const FragSet& fa, uint32_t bits, int ql) {  // ql: lane
#pragma unroll
    for (int tl = 0; tl < 2; ++tl) {
      // column lane % 32 of tile tl is replica 32 tl + lane % 32
      const uint32_t bt = (uint32_t)__shfl((int)bits, 32 * tl + (ql & 31), 64);
      f32x16 a = {};
      i32x16 ia = {};
      mfma_step(fa, half_of<I8>(bt, h), a, ia);
      write_fields_col(L.corr, a, ia, 32 * tl + (ql & 31));
    }
  };

  uint32_t t = 0;                               // running block counter
  // The two roles run separate copies of the block loop (so that each keeps
  // only its own state in registers); both pass the same barriers: one after
  // the block's work (B), and at the end of a pass one more (three after a
  // sweep).
  if (w == 0) {
    // schedule entries of the current step, each read one block ahead
    int c_nx = sched_block(0, 0);
    uint32_t xm_nx = sched_mask(0, 0);
    for (int ps = 0; ps < npass; ++ps) {
      int s;
      const bool dry = pass_info(ps, s);
      const uint64_t sw = (uint64_t)s + p.sweep_offset;
      const int sp = ps & 1;
      const float beta = act ? L.rb[q] : 0.f;
      double dsum = 0.0, eacc = 0.0;
      uint32_t nacc = 0;
      for (int tb = 0; tb < nblk; ++tb, ++t) {
        const int pb = t & 1;
        // The lane's LDS offsets are recomputed each block from an opaque
        // copy of its index (a few VALU ops): hoisted out of the loop they
        // would stay live across the chain, where registers are scarce.
        int qo = q;
        asm volatile("" : "+v"(qo));
        // block c of this step, its visit k being block row k ^ xm; c1 the next
        const int c = c_nx;
        const uint32_t xm = xm_nx;
        const int pos0 = 32 * c;
        const bool lastb = tb + 1 == nblk;
        const int c1 = sched_block(lastb ? sp ^ 1 : sp, lastb ? 0 : tb + 1);
        xm_nx = sched_mask(lastb ? sp ^ 1 : sp, lastb ? 0 : tb + 1);
        c_nx = c1;
        FragSet fcorr;
        load_corr_frags(c1, c, fcorr);
        if (dry) {
          // ---- dry pass: f64 energy of the block's rows from its full fields
          const uint32_t bits = L.bb[pb][qo];
          for (int i = 0; i < kB; ++i) {
            if (pos0 + i < n) {
              const int o = fld(qo, i);
              double f = (double)L.part[pb][0][o];
              for (int mm = 1; mm < MW; ++mm) f += (double)L.part[pb][mm][o];
              f += (double)L.corr[o];
              if (qsum) f = 0.5 * (f + (double)L.rsb[pb][i]);
              const double si = ((bits >> i) & 1u) ? 1.0 : (p.ptype == 0 ? -1.0 : 0.0);
              eacc += si * (0.5 * f + (double)L.hb[pb][i]);
            }
          }
          L.bnew[pb][qo] = bits;
          if (c1 == c) L.bb[pb ^ 1][qo] = bits;  // the same block next (the owners skip it)
          corr_products(fcorr, bits, qo);
          __syncthreads();  // B
          if (lastb) {
            if (act) {
              if (ps == 0)
                eoff = (double)E - eacc;
              else
                E = (float)(eacc + eoff);
            }
            __syncthreads();  // the next pass's first brackets
          }
          continue;
        }
        // ---- proposals of block c, one lane per replica
        // brackets of the block's 32 visits, read first (their latency
        // overlaps the field sums and they leave the chain free of LDS waits)
        float hib[kB], lob[kB];
#pragma unroll
        for (int t4 = 0; t4 < kB / 4; ++t4) {
          const float4 a = *reinterpret_cast<const float4*>(&L.hi[pb][qo][4 * t4]);
          const float4 b = *reinterpret_cast<const float4*>(&L.lo[pb][qo][4 * t4]);
          hib[4 * t4] = a.x, hib[4 * t4 + 1] = a.y, hib[4 * t4 + 2] = a.z, hib[4 * t4 + 3] = a.w;
          lob[4 * t4] = b.x, lob[4 * t4 + 1] = b.y, lob[4 * t4 + 2] = b.z, lob[4 * t4 + 3] = b.w;
        }
        float fp[kB];
        {
          const int xg = (int)(xm >> 2);
          // visit group r4 holds block rows 4 (r4 ^ xg) + (e ^ (xm & 3)); the
          // in-group part is a compile-time permutation per uniform branch
          auto load_fields = [&](auto pc) __attribute__((always_inline)) {
            constexpr int P = decltype(pc)::value;
            // kFG groups of 4 rows per stage: every LDS read of a stage is
            // issued before its first sum, so a stage costs one LDS latency
#pragma unroll
            for (int g0 = 0; g0 < kB / 4; g0 += kFG) {
              float4 ld[kFG][MW + 3];
#pragma unroll
              for (int j = 0; j < kFG; ++j) {
                const int gr = (g0 + j) ^ xg;
                const int o = fld(qo, 4 * gr);
#pragma unroll
                for (int mm = 0; mm < MW; ++mm)
                  ld[j][mm] = *reinterpret_cast<const float4*>(&L.part[pb][mm][o]);
                ld[j][MW] = *reinterpret_cast<const float4*>(&L.corr[o]);
                ld[j][MW + 1] = *reinterpret_cast<const float4*>(&L.hb[pb][4 * gr]);
                ld[j][MW + 2] = *reinterpret_cast<const float4*>(&L.rsb[pb][4 * gr]);
              }
              __builtin_amdgcn_sched_barrier(0);
#pragma unroll
              for (int j = 0; j < kFG; ++j) {
                float4 sm = ld[j][0];
#pragma unroll
                for (int mm = 1; mm < MW; ++mm) {
                  sm.x = sm.x + ld[j][mm].x;
                  sm.y = sm.y + ld[j][mm].y;
                  sm.z = sm.z + ld[j][mm].z;
                  sm.w = sm.w + ld[j][mm].w;
                }
                const float4 cr = ld[j][MW], h4 = ld[j][MW + 1], s4 = ld[j][MW + 2];
                const float vs[4] = {sm.x + cr.x + s4.x, sm.y + cr.y + s4.y,
                                     sm.z + cr.z + s4.z, sm.w + cr.w + s4.w};
                const float hv[4] = {h4.x, h4.y, h4.z, h4.w};
#pragma unroll
                for (int e = 0; e < 4; ++e)
                  fp[4 * (g0 + j) + e] = __builtin_fmaf(fsc, vs[e ^ P], hv[e ^ P]);
              }
              __builtin_amdgcn_sched_barrier(0);
            }
          };
          switch (xm & 3u) {
            case 0: load_fields(std::integral_constant<int, 0>{}); break;
            case 1: load_fields(std::integral_constant<int, 1>{}); break;
            case 2: load_fields(std::integral_constant<int, 2>{}); break;
            default: load_fields(std::integral_constant<int, 3>{}); break;
          }
        }
        // spins in visit order (bit k = block row k ^ xm)
        uint32_t mb = xor_permute_bits(L.bb[pb][qo], xm);
        const uint32_t mb0 = mb;
        // accepted flips, step k at bit 31 - k: one add with carry per step
        uint32_t tk = 0;
        // bit k of mb is only changed at step k, so every step reads its
        // spin from the block's initial bits: sign bit = !spin, magnitude cu
        const uint32_t nmb = ~mb0, cub = __float_as_uint(cu);
        uint32_t vsgn = 0x80000000u;  // kept in a register (not a literal operand)
        asm volatile("" : "+v"(vsgn));
        float dpart = 0.f;
        uint64_t amb = 0;  // lanes with a decision inside its bracket
        // Rows of J[blk][blk] (the same for every lane) stream from LDS
        // through a ring JR rows ahead of the chain.  Row k holds only the
        // columns right of the diagonal, shifted to the start of the row, so
        // the ceil((31 - k) / 4) 16-byte loads of a row carry no dead lanes.
        constexpr int JR = 2;
        const float4* jb4 = reinterpret_cast<const float4*>(L.jb[pb]);
        float4 jring[JR][kB / 4];
        static_for<JR>([&](auto dc) __attribute__((always_inline)) {
          constexpr int d = decltype(dc)::value;
          static_for<(kB - 1 - d + 3) / 4>([&](auto tc) __attribute__((always_inline)) {
            constexpr int t4 = decltype(tc)::value;
            jring[d][t4] = jb4[d * (kB / 4) + t4];
          });
        });
        static_for<kB>([&](auto kc) __attribute__((always_inline)) {
          constexpr int k = decltype(kc)::value;
          // dn = 2 s (F + h) (ising) or (2x - 1)(F + h) (qubo), exact in f32
          const uint32_t shk = nmb << (31 - k);
          const float sg = __uint_as_float((shk & vsgn) | (cub & ~vsgn));  // one bit insert
          const float dn = fp[k] * sg;
          const float hik = hib[k], lok = lob[k];
          // idle lanes and padding rows have lo = hi = +inf (and finite dn):
          // never taken, never inside their bracket
          const bool take = dn > hik;
          // the decision's lane mask, used as it is by the two selects below
          const uint64_t tm = __builtin_amdgcn_ballot_w64(take);
          amb |= __builtin_amdgcn_ballot_w64(!(dn < lok)) & ~tm;
          float dS;  // take ? -sg : 0
          asm("v_cndmask_b32_e64 %0, 0, -%1, %2" : "=v"(dS) : "v"(sg), "s"(tm));
          {
            uint64_t co;  // tk = 2 tk + take (the mask as carry-in)
            asm("v_addc_co_u32_e64 %0, %1, %2, %2, %3" : "=v"(tk), "=s"(co) : "v"(tk), "s"(tm));
            (void)co;
          }
          // dS * fp[k] is exactly -dn or 0, so this rounds like dpart - dn
          dpart = __builtin_fmaf(dS, fp[k], dpart);
          static_for<(kB - 1 - k + 3) / 4>([&](auto tc) __attribute__((always_inline)) {
            constexpr int t4 = decltype(tc)::value;
            const float4 j4 = jring[k % JR][t4];
            const float jv[4] = {j4.x, j4.y, j4.z, j4.w};
#pragma unroll
            for (int e = 0; e < 4; ++e) {
              const int r = k + 1 + 4 * t4 + e;
              // dS is 0, +-1 or +-2, so the fused form rounds like mul + add
              if (r < kB) fp[r] = __builtin_fmaf(jv[e], dS, fp[r]);
            }
          });
          if (k + JR < kB)
            static_for<(kB - 1 - (k + JR) + 3) / 4>([&](auto tc) __attribute__((always_inline)) {
              constexpr int t4 = decltype(tc)::value;
              jring[k % JR][t4] = jb4[(k + JR) * (kB / 4) + t4];
            });
        });
        mb = mb0 ^ __builtin_bitreverse32(tk);
        if (__builtin_expect(amb != 0, 0)) {
          // A decision fell inside its bracket: rerun the block, taking the
          // exact test for the decisions inside their brackets (rolled loop,
          // fields staged in LDS).
          float* sf = &L.sfp[qo * 33];
          for (int r = 0; r < kB; ++r) {  // visit order, as fp
            const int br = r ^ (int)xm;
            float v = L.part[pb][0][fld(qo, br)];
            for (int mm = 1; mm < MW; ++mm) v = v + L.part[pb][mm][fld(qo, br)];
            v = v + L.corr[fld(qo, br)];
            sf[r] = __builtin_fmaf(fsc, v + L.rsb[pb][br], L.hb[pb][br]);
          }
          mb = mb0;
          dpart = 0.f;
          for (int k = 0; k < kB; ++k) {
            const bool up = (mb >> k) & 1u;
            const float sg = up ? cu : -cu;
            const float dn = sf[k] * sg;
            // idle lanes and padding rows have lo = hi = +inf: never taken
            bool take = act & (dn > L.hi[pb][qo][k]);
            if (act && !take && !(dn < L.lo[pb][qo][k])) {
              const uint32_t pk = (uint32_t)(pos0 + (k ^ (int)xm));
              const uint32_t wd = pick_word(sweep_block(cid, sw, pk, p.k0, p.k1), pk & 3u);
              take = accept_exact(wd, beta, dn);
            }
            mb ^= (uint32_t)take << k;
            dpart += take ? -dn : 0.f;
            const float dS = take ? -sg : 0.f;
            for (int r = k + 1; r < kB; ++r)
              sf[r] = __builtin_fmaf(L.jb[pb][k * kB + (r - k - 1)], dS, sf[r]);
          }
        }
        {
          const uint32_t mbr = xor_permute_bits(mb, xm);  // back to block-row order
          L.bnew[pb][qo] = mbr;
          // the same block next (at a pass boundary): its spins are these
          // (the owners, whose copy is older, skip it)
          if (c1 == c) L.bb[pb ^ 1][qo] = mbr;
          corr_products(fcorr, mbr, qo);
        }
        dsum += (double)dpart;
        nacc += (uint32_t)__builtin_popcount(mb ^ mb0);
        __syncthreads();  // B
        if (lastb) {
          // ---- end of sweep: energy (simulation.py:157-158), best states, PT
          acc += nacc;
          E = (float)((double)E + dsum);
          L.E[q] = E;
          if (p.best_sem == 0) {  // simulation.py:161-164
            const bool imp = act && (E < bestE);
            if (imp) {
              bestE = E;
              p.best_rep_e[(size_t)read * R + rr] = E;
              p.best_rep_sw[(size_t)read * R + rr] = s;
            }
            L.flag[q] = imp ? 1 : 0;
          }
          if (p.use_pt && act && rr < R - 1)
            L.lu[q] = log_u32(pick_word(pt_block((uint64_t)read, sw, (uint32_t)rr, p.k0, p.k1),
                                        (uint32_t)rr & 3u));
          __syncthreads();
          if (act && rr == 0) {  // read leader
            const int b0 = g * R;
            int src = -1;
            if (p.best_sem == 1) {  // simulation.py:222-225 in replica order
              float cur = gE;
              for (int k = 0; k < R; ++k)
                if (L.E[b0 + k] < cur) {
                  cur = L.E[b0 + k];
                  src = k;
                }
              if (src >= 0) {
                gE = cur;
                gsw = s;
                gimp = 1;
              }
            }
            if (p.use_pt) {  // utils.py:44-58
              for (int k = 0; k < R - 1; ++k) {
                const int a = R - k - 1, b = R - k - 2;
                const float dd = (L.E[b0 + L.bi[b0 + a]] - L.E[b0 + L.bi[b0 + b]]) *
                                 (p.betas[a] - p.betas[b]);
                if (dd >= 0.f || L.lu[b0 + k] < (double)dd) {
                  const int tt = L.bi[b0 + a];
                  L.bi[b0 + a] = L.bi[b0 + b];
                  L.bi[b0 + b] = tt;
                }
              }
            }
            for (int k = 0; k < R; ++k) L.rb[b0 + L.bi[b0 + k]] = p.betas[k];
            if (p.logw && p.log_omegas) {  // ais.py:99-112 over the ascending ladder
              double lw = 0.0, bold = 0.0;
              for (int r = 1; r < R; ++r) {
                const int k = p.bsort[r];
                const float be = p.betas[k];
                const float en = L.E[b0 + L.bi[b0 + k]];
                const float a = -be * en;
                lw += (double)a - (-bold * (double)en);
                bold = (double)be;
              }
              p.log_omegas[(size_t)read * p.S + s] = lw;
            }
            L.src[g] = src;
          }
          __syncthreads();
          __syncthreads();  // the next pass's first brackets
        }
      }
    }
  } else {
    // schedule entries carried from block to block: c (this step), c1 (the
    // next) and its mask; c2 (the one after) is read each block
    int c = sched_block(0, 0), c1 = sched_block(0, 1);
    uint32_t m1 = sched_mask(0, 1);
    for (int ps = 0; ps < npass; ++ps) {
      int s, s_next;
      const bool dry = pass_info(ps, s);
      const bool dry_next = pass_info(ps + 1, s_next);
      const uint64_t sw = (uint64_t)s + p.sweep_offset;
      const uint64_t sw_next = (uint64_t)s_next + p.sweep_offset;
      const int sp = ps & 1;
      for (int tb = 0; tb < nblk; ++tb, ++t) {
        const int pb = t & 1, pn = pb ^ 1;
        // the schedule of the next pass, needed from step nblk - 2 on
        if (BR && w == 4 && tb == 0) fill_sched(sp ^ 1, sw_next, dry_next);
        const bool last = tb + 1 == nblk;
        const int p2 = tb + 2 < nblk ? sp : sp ^ 1, i2 = tb + 2 < nblk ? tb + 2 : tb + 2 - nblk;
        const int c2 = sched_block(p2, i2);
        const uint32_t m2 = sched_mask(p2, i2);
        // ---- meanwhile: block c1's spins, fields except K block c, J / h and brackets
        if (mw) {
          // c1's spins have not changed since its last visit unless c1 == c
          // (a pass boundary), which the proposal wave publishes itself
          if (w == owner_of(c1) && c1 != c) publish_bits(c1, pn);
          compute_part(c1, c, L.part[pn][m]);
          prefetch_frags(c2);
        }
        if (hw) {
          publish_block(pn);
          if (w > MW + 1 && !last && !dry) brackets(c1, m1, sw, pn);
          prefetch_jb(c2, m2);
        }
        __syncthreads();  // B
        // ---- the owners of block c take its new spins
        if (mw && w == owner_of(c)) {
          set_half<NW>(sa[0], c - c0, half_of<I8>(L.bnew[pb][qtt[0]], h));
          set_half<NW>(sa[1], c - c0, half_of<I8>(L.bnew[pb][qtt[1]], h));
        }
        if (last) {
          if (!dry) {
            __syncthreads();
#pragma unroll
            for (int tt = 0; tt < 2; ++tt)
              if (mw && p.best_sem == 0 && actt[tt] && L.flag[qtt[tt]])
                store_range<NW, I8>(sa[tt], c0, kb, nblk, n, p.ptype, p.best_rep + base_t[tt]);
            __syncthreads();
            // owners copy their part of a new global best
#pragma unroll
            for (int tt = 0; tt < 2; ++tt)
              if (mw && p.best_sem == 1 && actt[tt] && L.src[gtt[tt]] == qtt[tt] - gtt[tt] * R)
                store_range<NW, I8>(sa[tt], c0, kb, nblk, n, p.ptype,
                                    p.best_state + (size_t)(read0 + gtt[tt]) * n);
          }
          // the next pass's first block (after the exchanges: new betas)
          if (hw && w > MW + 1 && !dry_next) brackets(c1, m1, sw_next, pn);
          __syncthreads();
        }
        c = c1;
        c1 = c2;
        m1 = m2;
      }
    }
  }

  // ---- outputs
#pragma unroll
  for (int tt = 0; tt < 2; ++tt)
    if (actt[tt]) store_range<NW, I8>(sa[tt], c0, kb, nblk, n, p.ptype, p.spins + base_t[tt]);
  if (w == 0 && act) {
    if (p.accepted) atomicAdd(p.accepted + read, acc);
    p.energies[(size_t)read * R + rr] = E;
    p.beta_idx[(size_t)read * R + rr] = L.bi[q];
    if (rr == 0 && p.best_sem == 1) {
      p.best_energy[read] = gE;
      p.best_sweeps[read] = gsw;
      L.src[g] = gimp;
    }
  }
  if (p.best_sem == 1) {
    // never improved: best_state is still the view of states[0] (simulation.py:204)
    __syncthreads();
#pragma unroll
    for (int tt = 0; tt < 2; ++tt)
      if (actt[tt] && qtt[tt] - gtt[tt] * R == 0 && !L.src[gtt[tt]])
        store_range<NW, I8>(sa[tt], c0, kb, nblk, n, p.ptype,
                            p.best_state + (size_t)(read0 + gtt[tt]) * n);
  }
}

// Final argmin over per-replica bests (simulation.py:174-177), one thread per read.
__global__ void best_argmin_kernel(BParams p) {
  const int read = blockIdx.x * blockDim.x + threadIdx.x;
  if (read >= p.reads) return;
  const int R = p.R;
  int bp = 0;
  for (int k = 1; k < R; ++k)
    if (p.best_rep_e[(size_t)read * R + k] < p.best_rep_e[(size_t)read * R + bp]) bp = k;
  p.best_energy[read] = p.best_rep_e[(size_t)read * R + bp];
  p.best_sweeps[read] = p.best_rep_sw[(size_t)read * R + bp];
  const int8_t* src = p.best_rep + ((size_t)read * R + bp) * p.n;
  for (int i = 0; i < p.n; ++i) p.best_state[(size_t)read * p.n + i] = src[i];
}

template <int NP, bool I8, int KW>
hipError_t launch_kw(const BParams& p, int nw, int groups, hipStream_t st) {
  dim3 grid(groups), block(64 * KW);
  switch (nw) {
    case 4: hipLaunchKernelGGL((blocked_pt_kernel<NP, I8, 4, KW>), grid, block, 0, st, p); break;
    case 8: hipLaunchKernelGGL((blocked_pt_kernel<NP, I8, 8, KW>), grid, block, 0, st, p); break;
    case 16: hipLaunchKernelGGL((blocked_pt_kernel<NP, I8, 16, KW>), grid, block, 0, st, p); break;
    case 32: hipLaunchKernelGGL((blocked_pt_kernel<NP, I8, 32, KW>), grid, block, 0, st, p); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}


template <int NP, bool I8, bool BR>
hipError_t launch_pipe_o(const BParams& p, int nw, int groups, hipStream_t st) {
  dim3 grid(groups), block(128 * (kPipeMW + 1));
  switch (nw) {
    case 4: hipLaunchKernelGGL((blocked_pipe_kernel<NP, I8, 4, kPipeMW, BR>), grid, block, 0, st, p); break;
    case 8: hipLaunchKernelGGL((blocked_pipe_kernel<NP, I8, 8, kPipeMW, BR>), grid, block, 0, st, p); break;
    case 16: hipLaunchKernelGGL((blocked_pipe_kernel<NP, I8, 16, kPipeMW, BR>), grid, block, 0, st, p); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int NP, bool I8>
hipError_t launch_pipe(const BParams& p, int nw, int groups, hipStream_t st) {
  return p.order == kBlockRandom ? launch_pipe_o<NP, I8, true>(p, nw, groups, st)
                                 : launch_pipe_o<NP, I8, false>(p, nw, groups, st);
}

template <int NP, bool I8>
hipError_t launch_np(const BParams& p, int nw, int kw, int groups, hipStream_t st) {
  switch (kw) {
    case 1: return launch_kw<NP, I8, 1>(p, nw, groups, st);
    case 2: return launch_kw<NP, I8, 2>(p, nw, groups, st);
    case 4: return launch_kw<NP, I8, 4>(p, nw, groups, st);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

void blocked_release(BlockedProblem& b) {
  if (b.packed) (void)hipFree(b.packed);
  if (b.scratch) (void)hipFree(b.scratch);
  if (b.rowsum) (void)hipFree(b.rowsum);
  if (b.jtl) (void)hipFree(b.jtl);
  b = BlockedProblem();
}

bool blocked_eligible(int n, int dtype, int R, int reads, int order, bool replay) {
  // the block-random order runs on the pipelined kernel (at least 3 blocks)
  const bool ord_ok = order == kSequential || (order == kBlockRandom && n > 2 * kB);
  return dtype == 0 && n >= 32 && n <= kBlockedMaxN && R >= 1 && R <= 64 && reads >= 1 &&
         ord_ok && !replay;
}

// Smallest exact coupling encoding: 0 = int8, 1..3 = bf16 pieces; 4 = none
// of them (f32 needs at most 3 pieces, so 4 is unreachable for finite input).
int blocked_exact_pieces(const float* J, size_t count) {
  bool i8 = true;
  int np = 1;
  for (size_t k = 0; k < count; ++k) {
    const float v = J[k];
    if (i8 && !(v == (float)(int)v && v >= -128.f && v <= 127.f)) i8 = false;
    float r = v;
    int need = 0;
    while (r != 0.f && need < 4) {
      r = r - bf16_to_f(bf16_rn(r));
      ++need;
    }
    if (need > np) np = need;
  }
  return i8 ? 0 : np;
}

namespace {

// Decisions of one run that depend only on its shape and the encoding.
struct BlockedPlan {
  int nblk, groups, kw, kb, nw, kbp, nwp, kind, np;
  bool pipe, i8, inexact;
};

// Coupling encoding and kernel shape; b.exact_kind must be known.
BlockedPlan plan_blocked(const BlockedProblem& b, int n, int reads, int R, int order, int pieces) {
  BlockedPlan pl;
  pl.nblk = (n + kB - 1) / kB;
  // Waves per replica group: enough groups x waves to give every SIMD of the
  // device work (1024 on MI355X), at most 4, at least one block per wave.
  pl.groups = (reads + (64 / R) - 1) / (64 / R);
  pl.kw = 1;
  while (pl.kw < 4 && 2 * pl.kw <= pl.nblk && pl.groups * pl.kw < 1024) pl.kw *= 2;
  // The pipelined kernel (proposal wave + six matrix waves) takes every
  // problem of at least three blocks; PYSA_AMD_BLOCKED_SPLITK=1 selects the
  // split-K kernel instead (kept for small problems and comparisons).
  static const bool force_splitk = [] {
    const char* e = getenv("PYSA_AMD_BLOCKED_SPLITK");
    return e && e[0] == '1';
  }();
  pl.pipe = pl.nblk >= kPipeMW && (!force_splitk || order == kBlockRandom);
  pl.kbp = (pl.nblk + kPipeMW - 1) / kPipeMW;
  pl.nwp = 4;
  while (2 * pl.nwp < pl.kbp) pl.nwp *= 2;
  pl.kb = (pl.nblk + pl.kw - 1) / pl.kw;
  pl.nw = 4;
  while (2 * pl.nw < pl.kb) pl.nw *= 2;  // words of 2 blocks each: kb <= 2 * nw
  // Coupling encoding: exact where possible; `pieces` forces 2 or 3 bf16 pieces.
  int kind = b.exact_kind;  // 0 int8, else bf16 pieces
  if (kind != 0) {
    if (pieces == 2 || pieces == 3)
      kind = (kind < pieces) ? kind : pieces;  // fewer pieces never help when more are exact
    else
      kind = (kind > 2) ? 2 : kind;  // default: hi+lo split for real couplings
    if (pieces == 3 && b.exact_kind >= 3) kind = 3;
  }
  pl.kind = kind;
  pl.i8 = kind == 0;
  pl.np = pl.i8 ? 1 : kind;
  // The pieces do not hold every coupling exactly: the in-block updates and
  // the QUBO row sums then use the piece sums too (the couplings the products
  // multiply with), so all fields derive from one coupling matrix.
  pl.inexact = !pl.i8 && kind < b.exact_kind;
  return pl;
}

// Whether the problem's device copies match the plan (packed couplings,
// piece sums, row sums).
bool encoding_ready(const BlockedProblem& b, int ptype, const BlockedPlan& pl) {
  if (!b.packed || b.packed_kind != pl.kind) return false;
  if (pl.inexact && (!b.jtl || b.jtl_kind != pl.kind)) return false;
  if (!pl.i8 && ptype == 1 && (!b.rowsum || b.rowsum_kind != pl.kind)) return false;
  return true;
}

// Build the device copies the plan needs (allocates on first use).
int prepare_encoding(BlockedProblem& b, const void* dJ, const void* dJt, int n, int ptype,
                     const BlockedPlan& pl, hipStream_t st, std::string* why) {
  const int nblk = pl.nblk, kind = pl.kind, np = pl.np;
  if (b.packed_kind != kind || !b.packed) {
    if (b.packed) (void)hipFree(b.packed);
    b.packed = nullptr;
    b.packed_kind = -1;
    const size_t bytes =
        pl.i8 ? (size_t)nblk * nblk * 64 * 16 : (size_t)nblk * 2 * nblk * np * 64 * 16;
    if (hipMalloc(&b.packed, bytes) != hipSuccess) {
      b.packed = nullptr;
      *why = "allocating packed couplings";
      return PYSA_AMD_EHIP;
    }
    b.packed_bytes = bytes;
    const long long threads = (long long)(bytes / 16);
    const unsigned nb = (unsigned)((threads + 255) / 256);
    if (pl.i8)
      hipLaunchKernelGGL(pack_i8_kernel, dim3(nb), dim3(256), 0, st, (const float*)dJ, n, nblk,
                         (int8_t*)b.packed);
    else
      hipLaunchKernelGGL(pack_bf16_kernel, dim3(nb), dim3(256), 0, st, (const float*)dJ, n, nblk,
                         np, (uint16_t*)b.packed);
    if (hipGetLastError() != hipSuccess) {
      *why = "packing couplings";
      return PYSA_AMD_EHIP;
    }
    b.packed_kind = kind;
  }
  if (pl.inexact && b.jtl_kind != kind) {
    if (!b.jtl && hipMalloc(&b.jtl, sizeof(float) * (size_t)n * n) != hipSuccess) {
      b.jtl = nullptr;
      *why = "allocating piece-sum couplings";
      return PYSA_AMD_EHIP;
    }
    const long long cnt2 = (long long)n * n;
    hipLaunchKernelGGL(pieces_sum_kernel, dim3((unsigned)((cnt2 + 255) / 256)), dim3(256), 0, st,
                       (const float*)dJt, cnt2, np, b.jtl);
    if (hipGetLastError() != hipSuccess) {
      *why = "piece-sum couplings";
      return PYSA_AMD_EHIP;
    }
    b.jtl_kind = kind;
  }
  if (!pl.i8 && ptype == 1 && b.rowsum_kind != kind) {
    if (!b.rowsum && hipMalloc(&b.rowsum, sizeof(float) * n) != hipSuccess) {
      b.rowsum = nullptr;
      *why = "allocating row sums";
      return PYSA_AMD_EHIP;
    }
    hipLaunchKernelGGL(rowsum_kernel, dim3((n + 255) / 256), dim3(256), 0, st, (const float*)dJ,
                       n, pl.inexact ? np : 0, b.rowsum);
    if (hipGetLastError() != hipSuccess) {
      *why = "row sums";
      return PYSA_AMD_EHIP;
    }
    b.rowsum_kind = kind;
  }
  return 0;
}

// Best-state scratch of a run: reads*R*n spins, then reads*R energies and
// reads*R sweep indices.
size_t scratch_need(int n, int reads, int R) {
  const size_t cnt = (size_t)reads * R;
  return cnt * n + cnt * 4 + cnt * 4;
}

__global__ void zero_u64_kernel(unsigned long long* p, int count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = 0ull;
}

// Enqueue one run on stream st (no allocation, no synchronisation).
hipError_t launch_blocked(const BlockedProblem& b, const BlockedPlan& pl, const void* dJt,
                          const void* dh, int n, int ptype, int reads, int R, int S, int order,
                          int use_pt, int logw, int best_sem, uint64_t seed, uint64_t sweep_offset,
                          const void* betas, const int32_t* bsort, int8_t* states, void* energies,
                          int64_t* beta_idx, int8_t* best_state, void* best_energy,
                          int32_t* best_sweeps, double* log_omegas, uint64_t* accepted,
                          void* scratch_mem, hipStream_t st) {
  const size_t cnt = (size_t)reads * R;
  BParams p;
  p.n = n;
  p.nblk = pl.nblk;
  p.ptype = ptype;
  p.R = R;
  p.reads = reads;
  p.S = S;
  p.rpw = 64 / R;
  p.np = pl.np;
  p.use_pt = use_pt;
  p.logw = logw;
  p.best_sem = best_sem;
  p.kb = pl.pipe ? pl.kbp : pl.kb;
  p.k0 = (uint32_t)seed;
  p.k1 = (uint32_t)(seed >> 32);
  p.sweep_offset = sweep_offset;
  p.frag = b.packed;
  p.Jt = pl.inexact ? b.jtl : (const float*)dJt;
  p.h = (const float*)dh;
  p.order = order;
  // PYSA_AMD_ENERGY_REFRESH=<sweeps> overrides the interval for every
  // problem (0 = never), a switch for tests and comparisons
  const char* ev = getenv("PYSA_AMD_ENERGY_REFRESH");
  p.refresh = ev && ev[0] ? atoi(ev) : (b.exact_sums == 1 ? 0 : 64);
  p.rowsum = (!pl.i8 && ptype == 1) ? b.rowsum : nullptr;
  p.betas = (const float*)betas;
  p.bsort = bsort;
  p.spins = states;
  p.energies = (float*)energies;
  p.beta_idx = beta_idx;
  p.best_rep = (int8_t*)scratch_mem;
  p.best_rep_e = (float*)((int8_t*)scratch_mem + cnt * n);
  p.best_rep_sw = (int*)((int8_t*)scratch_mem + cnt * n + cnt * 4);
  p.best_state = best_state;
  p.best_energy = (float*)best_energy;
  p.best_sweeps = best_sweeps;
  p.log_omegas = logw ? log_omegas : nullptr;
  p.accepted = (unsigned long long*)accepted;
  hipError_t e = hipSuccess;
  if (best_sem == 0) {
    e = hipMemcpyAsync(p.best_rep, states, cnt * n, hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess) return e;
  }
  if (accepted) {  // a fill kernel, not a memset, so that the run can be captured in a graph
    hipLaunchKernelGGL(zero_u64_kernel, dim3((reads + 255) / 256), dim3(256), 0, st,
                       (unsigned long long*)accepted, reads);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  const int np = pl.np;
  if (pl.pipe) {
    if (pl.i8)
      e = launch_pipe<1, true>(p, pl.nwp, pl.groups, st);
    else if (np == 1)
      e = launch_pipe<1, false>(p, pl.nwp, pl.groups, st);
    else if (np == 2)
      e = launch_pipe<2, false>(p, pl.nwp, pl.groups, st);
    else
      e = launch_pipe<3, false>(p, pl.nwp, pl.groups, st);
  } else if (pl.i8)
    e = launch_np<1, true>(p, pl.nw, pl.kw, pl.groups, st);
  else if (np == 1)
    e = launch_np<1, false>(p, pl.nw, pl.kw, pl.groups, st);
  else if (np == 2)
    e = launch_np<2, false>(p, pl.nw, pl.kw, pl.groups, st);
  else
    e = launch_np<3, false>(p, pl.nw, pl.kw, pl.groups, st);
  if (e == hipSuccess && best_sem == 0) {
    hipLaunchKernelGGL(best_argmin_kernel, dim3((reads + 63) / 64), dim3(64), 0, st, p);
    e = hipGetLastError();
  }
  return e;
}

struct EventPair {
  hipEvent_t a = nullptr, b = nullptr;
  hipError_t create() {
    hipError_t e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    return e;
  }
  ~EventPair() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
  }
};

const char* kNotEligible =
    "blocked form needs float32 couplings, 32 <= n <= 2048, R <= 64, sequential or "
    "block-random order (the latter for n > 64) and Philox randomness (no replay)";

}  // namespace

// The allocation-free entry point's scratch on each device (a copy of the
// file-scope array `scratch` per device), looked up once per device with that
// device current.
namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

int noalloc_scratch(void** ptr, size_t* bytes, std::string* why) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    *why = "no current HIP device (or more than 64 devices)";
    return PYSA_AMD_EHIP;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      *why = std::string("looking up the allocation-free scratch: ") + hipGetErrorString(e);
      return PYSA_AMD_EHIP;
    }
    g_scratch_ptr[dev] = p;
  }
  *ptr = g_scratch_ptr[dev];
  *bytes = kNoAllocScratchBytes;
  return 0;
}

// Whether J and h are integers small enough that every field and energy sum
// (at most sum_j |J_ij| + |h_i| in magnitude) is exact in f32: the tracked
// energy then never drifts and the recomputation passes are skipped.
bool integer_sums_exact(const float* J, const float* h, int n) {
  double worst = 0.0;
  for (int i = 0; i < n; ++i) {
    if (!(fabsf(h[i]) < 16777216.f) || h[i] != (float)(int64_t)h[i]) return false;
    double row = fabs((double)h[i]);
    for (int j = 0; j < n; ++j) {
      const float v = J[(size_t)i * n + j];
      if (!(fabsf(v) < 16777216.f) || v != (float)(int64_t)v) return false;
      row += fabs((double)v);
    }
    worst = row > worst ? row : worst;
  }
  return worst < 16777216.0 / 4;  // margin for the QUBO row-sum form and 2 x dn
}

int run_blocked(BlockedProblem& b, const void* dJ, const void* dJt, const void* dh, int n,
                int ptype, int dtype, int reads, int R, int S, int order, int use_pt, int logw,
                int best_sem, int pieces, uint64_t seed, uint64_t sweep_offset, const void* betas,
                const int32_t* bsort, const double* rep_lr, const double* rep_pt, int8_t* states,
                void* energies, int64_t* beta_idx, int8_t* best_state, void* best_energy,
                int32_t* best_sweeps, double* log_omegas, uint64_t* accepted, hipStream_t st,
                double* ms, std::string* why) {
  if (!blocked_eligible(n, dtype, R, reads, order, rep_lr != nullptr || rep_pt != nullptr)) {
    *why = kNotEligible;
    return PYSA_AMD_ENOTSUP;
  }
  if (b.exact_kind < 0) {
    std::vector<float> hJ((size_t)n * n), hh(n);
    if (hipMemcpy(hJ.data(), dJ, hJ.size() * 4, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(hh.data(), dh, hh.size() * 4, hipMemcpyDeviceToHost) != hipSuccess) {
      *why = "reading couplings back";
      return PYSA_AMD_EHIP;
    }
    b.exact_kind = blocked_exact_pieces(hJ.data(), hJ.size());
    b.exact_sums = integer_sums_exact(hJ.data(), hh.data(), n) ? 1 : 0;
  }
  const BlockedPlan pl = plan_blocked(b, n, reads, R, order, pieces);
  int rc = prepare_encoding(b, dJ, dJt, n, ptype, pl, st, why);
  if (rc) return rc;
  const size_t need = scratch_need(n, reads, R);
  if (b.scratch_bytes < need) {
    if (b.scratch) (void)hipFree(b.scratch);
    b.scratch = nullptr;
    b.scratch_bytes = 0;
    if (hipMalloc(&b.scratch, need) != hipSuccess) {
      b.scratch = nullptr;
      *why = "allocating scratch";
      return PYSA_AMD_EHIP;
    }
    b.scratch_bytes = need;
  }
  EventPair ev;
  if (ev.create() != hipSuccess) {
    *why = "creating events";
    return PYSA_AMD_EHIP;
  }
  (void)hipEventRecord(ev.a, st);
  hipError_t le = launch_blocked(b, pl, dJt, dh, n, ptype, reads, R, S, order, use_pt, logw,
                                 best_sem, seed, sweep_offset, betas, bsort, states, energies,
                                 beta_idx, best_state, best_energy, best_sweeps, log_omegas,
                                 accepted, b.scratch, st);
  (void)hipEventRecord(ev.b, st);
  if (le != hipSuccess) {
    *why = std::string("blocked kernel launch: ") + hipGetErrorString(le);
    return PYSA_AMD_EHIP;
  }
  if (hipEventSynchronize(ev.b) != hipSuccess) {
    *why = "blocked kernel execution";
    return PYSA_AMD_EHIP;
  }
  float t = 0.f;
  (void)hipEventElapsedTime(&t, ev.a, ev.b);
  if (ms) *ms = t;
  b.last_kind = pl.kind;
  return 0;
}

int run_blocked_noalloc(const BlockedProblem& b, const void* dJt, const void* dh, int n,
                        int ptype, int dtype, int reads, int R, int S, int order, int use_pt,
                        int logw, int best_sem, int pieces, uint64_t seed, uint64_t sweep_offset,
                        const void* betas, const int32_t* bsort, int8_t* states, void* energies,
                        int64_t* beta_idx, int8_t* best_state, void* best_energy,
                        int32_t* best_sweeps, double* log_omegas, uint64_t* accepted,
                        void* scratch_mem, size_t scratch_bytes, hipStream_t st,
                        std::string* why) {
  if (!blocked_eligible(n, dtype, R, reads, order, false)) {
    *why = kNotEligible;
    return PYSA_AMD_ENOTSUP;
  }
  const size_t need = scratch_need(n, reads, R);
  if (need > scratch_bytes) {
    *why = "reads * replicas * (n + 8) = " + std::to_string(need) +
           " bytes exceeds the allocation-free scratch of " + std::to_string(scratch_bytes) +
           " bytes";
    return PYSA_AMD_EINVAL;
  }
  if (b.exact_kind < 0) {
    *why = "couplings not yet encoded for the MFMA form: run this form once through an "
           "allocating entry point first";
    return PYSA_AMD_ENOTSUP;
  }
  const BlockedPlan pl = plan_blocked(b, n, reads, R, order, pieces);
  if (!encoding_ready(b, ptype, pl)) {
    *why = "couplings encoded for another form: run this form once through an allocating "
           "entry point first";
    return PYSA_AMD_ENOTSUP;
  }
  hipError_t le = launch_blocked(b, pl, dJt, dh, n, ptype, reads, R, S, order, use_pt, logw,
                                 best_sem, seed, sweep_offset, betas, bsort, states, energies,
                                 beta_idx, best_state, best_energy, best_sweeps, log_omegas,
                                 accepted, scratch_mem, st);
  if (le != hipSuccess) {
    *why = std::string("blocked kernel launch: ") + hipGetErrorString(le);
    return PYSA_AMD_EHIP;
  }
  return 0;
}

}  // namespace pysa
