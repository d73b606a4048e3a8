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
// Layout of a wave: 64 replicas = floor(64/R) reads of R replicas; replicas
// q and q+32 form MFMA column tiles 0 and 1.  One launch runs all sweeps;
// parallel tempering, best tracking and log-omega run per read inside the
// wave after every sweep.
#include "blocked.hpp"

#include <math.h>
#include <string.h>

#include <vector>

#include "../../include/pysa_amd.h"
#include "engine.hpp"
#include "rng.hpp"

namespace pysa {
namespace {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(16))) float f32x16;
typedef __attribute__((ext_vector_type(4))) int i32x4;
typedef __attribute__((ext_vector_type(16))) int i32x16;
typedef __attribute__((ext_vector_type(4))) unsigned u32x4v;

constexpr int kB = 32;            // block size (rows of one MFMA tile)
constexpr int kWavesPerGroup = 4; // waves per workgroup
constexpr int kLdsPerWave = 2 * kB * 33 * 4 + 64 * 4 * 4 + 64 * 8;  // transpose + PT scratch

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
__device__ __forceinline__ float rlf(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

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

// Eight spins (bits of x) as bf16 pairs: +1/-1 (ising) or 1/0 (qubo).
__device__ __forceinline__ void spins_bf16(uint32_t x, int ptype, uint32_t (&d)[4]) {
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const uint32_t t = (x >> (2 * e)) & 3u;
    const uint32_t sp = ((t | (t << 15)) & 0x10001u);  // bit0 -> 0, bit1 -> 16
    if (ptype == 0)
      d[e] = 0xBF80BF80u ^ (sp << 15);  // up clears the sign bit
    else
      d[e] = sp * 0x3F80u;
  }
}
// Sixteen spins (bits of x) as int8 +1/-1 (ising) or 1/0 (qubo).
__device__ __forceinline__ void spins_i8(uint32_t x, int ptype, uint32_t (&d)[4]) {
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const uint32_t t = (x >> (4 * e)) & 15u;
    const uint32_t sp = (t * 0x00204081u) & 0x01010101u;  // bit b -> byte b
    d[e] = (ptype == 0) ? (0xFFFFFFFFu - sp * 0xFEu) : sp;
  }
}

// Certified acceptance test: returns det_log(u) < x64 for u = (w + 0.5) 2^-32,
// deciding in single precision when the gap exceeds the error bound of the
// fast logarithm (and of the rounding of x to f32), else in exact f64.
__device__ __forceinline__ bool log_less(uint32_t w, float x32, double x64) {
  const float uf = ((float)w + 0.5f) * 2.3283064365386963e-10f;
  const float L = __builtin_amdgcn_logf(uf) * 0.69314718055994531f;
  const float m = 3.9e-6f * fabsf(L) + 9.6e-7f + 2.4e-7f * fabsf(x32);
  if (L < x32 - m) return true;
  if (L > x32 + m) return false;
  return log_u32(w) < x64;
}

// Spin bits of one MFMA column tile as seen by one lane: word c/2 holds the
// 16-bit half of blocks 2(c/2) and 2(c/2)+1.  A vector type keeps the words
// in VGPRs under wave-uniform dynamic indexing.
template <int NW>
struct Bits {
  typedef uint32_t V __attribute__((ext_vector_type(NW)));
  V w;
};

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
  int n, nblk, ptype, R, reads, S, rpw, np, use_pt, logw, best_sem, refresh;
  uint32_t k0, k1;
  uint64_t sweep_offset;
  const void* frag;
  const float* Jt;  // f32 J^T, row j = coefficients J[:][j] (in-block updates)
  const float* h;
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

// Write the spins of replica `q` of this wave (held in tile q/32, lanes
// q%32 and q%32+32) to dst[0..n): called by both lanes of the pair.
template <int NW, bool I8>
__device__ __forceinline__ void store_state(const Bits<NW>& s, int nblk, int n, int ptype, int8_t* dst) {
  const int h = (threadIdx.x & 63) >> 5;
  const int8_t dn = ptype == 0 ? int8_t(-1) : int8_t(0);
  for (int c = 0; c < nblk; ++c) {
    const uint32_t hv = get_half<NW>(s, c);
#pragma unroll
    for (int b = 0; b < 16; ++b) {
      const int rel = I8 ? (16 * h + b) : ((b < 8) ? (8 * h + b) : (16 + 8 * h + (b - 8)));
      const int pos = 32 * c + rel;
      if (pos < n) dst[pos] = ((hv >> b) & 1) ? int8_t(1) : dn;
    }
  }
}

template <int NP, bool I8, int NW>
__global__ __launch_bounds__(64 * kWavesPerGroup) void blocked_pt_kernel(BParams p) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  unsigned char* my = lds_raw + (size_t)wave * kLdsPerWave;
  float* s_tr = reinterpret_cast<float*>(my);                       // 2 tiles x 32 x 33
  float* s_E = reinterpret_cast<float*>(my + 2 * kB * 33 * 4);       // 64
  int* s_bi = reinterpret_cast<int*>(s_E + 64);                      // 64: slot -> replica per read
  float* s_rb = reinterpret_cast<float*>(s_bi + 64);                 // 64: beta of each replica
  int* s_flag = reinterpret_cast<int*>(s_rb + 64);                   // 64
  double* s_lu = reinterpret_cast<double*>(s_flag + 64);             // 64: PT log-uniforms

  const int R = p.R, n = p.n, nblk = p.nblk;
  const int gw = blockIdx.x * kWavesPerGroup + wave;  // global wave
  const int read0 = gw * p.rpw;
  // In-block role: lane = replica slot q of this wave.
  const int q = lane;
  const int g = q / R, rr = q - g * R;  // read within the wave, replica within the read
  const int read = read0 + g;
  const bool act = (g < p.rpw) && (read < p.reads);
  const int h = lane >> 5;
  // MFMA-operand role: lane holds tile-0 replica lane%32 and tile-1 replica 32+lane%32.
  const int qa = lane & 31, qb = 32 + (lane & 31);
  const int ga = qa / R, gb = qb / R;
  const bool acta = (ga < p.rpw) && (read0 + ga < p.reads);
  const bool actb = (gb < p.rpw) && (read0 + gb < p.reads);
  const size_t base_a = ((size_t)(read0 + ga) * R + (qa - ga * R)) * n;
  const size_t base_b = ((size_t)(read0 + gb) * R + (qb - gb * R)) * n;

  // Spin bits in operand layout.
  Bits<NW> sa, sb;
  sa.w = 0;
  sb.w = 0;
  for (int c = 0; c < nblk; ++c) {
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
    set_half<NW>(sa, c, va);
    set_half<NW>(sb, c, vb);
  }
  float E = act ? p.energies[(size_t)read * R + rr] : 0.f;
  // Slot map and betas (sa.py:325, simulation.py:151-152).
  if (act) {
    s_bi[q] = rr;  // beta_idx = arange(R) for each read
    s_rb[q] = p.betas[rr];
  }
  float bestE = E;
  if (p.best_sem == 0 && act) {
    p.best_rep_e[(size_t)read * R + rr] = E;
    p.best_rep_sw[(size_t)read * R + rr] = 0;
  }
  float gE = 0.f;
  int gsw = 0, gimp = 0;
  if (act && rr == 0) gE = E;  // simulation.py:205 (read leader)
  const uint64_t cid = (uint64_t)read * R + rr;
  unsigned long long acc = 0;
  __syncthreads();

  for (int s = 0; s < p.S; ++s) {
    const uint64_t sw = (uint64_t)s + p.sweep_offset;
    const float beta = act ? s_rb[q] : 0.f;
    const double bd = (double)beta;
    double dsum = 0.0;
    for (int c = 0; c < nblk; ++c) {
      // ---- 1. block fields on the matrix cores
      f32x16 acc0 = {}, acc1 = {};
      i32x16 iacc0 = {}, iacc1 = {};
      for (int cp = 0; cp < nblk; ++cp) {
        const uint32_t ha = get_half<NW>(sa, cp), hb = get_half<NW>(sb, cp);
        if (I8) {
          const i32x4* fr = reinterpret_cast<const i32x4*>(p.frag) + ((size_t)c * nblk + cp) * 64 + lane;
          const i32x4 a = *fr;
          uint32_t da[4], db[4];
          spins_i8(ha, p.ptype, da);
          spins_i8(hb, p.ptype, db);
          i32x4 ba, bb;
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            ba[e] = (int)da[e];
            bb[e] = (int)db[e];
          }
          iacc0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, ba, iacc0, 0, 0, 0);
          iacc1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, bb, iacc1, 0, 0, 0);
        } else {
#pragma unroll
          for (int st = 0; st < 2; ++st) {
            const bf16x8* fr = reinterpret_cast<const bf16x8*>(p.frag) +
                               (((size_t)c * (2 * nblk) + 2 * cp + st) * NP) * 64 + lane;
            bf16x8 a[NP];
#pragma unroll
            for (int pc = 0; pc < NP; ++pc) a[pc] = fr[pc * 64];
            uint32_t da[4], db[4];
            spins_bf16((ha >> (8 * st)) & 0xFFu, p.ptype, da);
            spins_bf16((hb >> (8 * st)) & 0xFFu, p.ptype, db);
            u32x4v ua = {da[0], da[1], da[2], da[3]}, ub = {db[0], db[1], db[2], db[3]};
            const bf16x8 ba = __builtin_bit_cast(bf16x8, ua), bb = __builtin_bit_cast(bf16x8, ub);
#pragma unroll
            for (int pc = 0; pc < NP; ++pc) {
              acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[pc], ba, acc0, 0, 0, 0);
              acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[pc], bb, acc1, 0, 0, 0);
            }
          }
        }
      }
      // ---- 2. transpose through LDS: lane q receives the 32 fields of replica q
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int i = 8 * (v >> 2) + 4 * h + (v & 3);
        const float f0 = I8 ? (float)iacc0[v] : acc0[v];
        const float f1 = I8 ? (float)iacc1[v] : acc1[v];
        s_tr[(lane & 31) * 33 + i] = f0;
        s_tr[(32 + (lane & 31)) * 33 + i] = f1;
      }
      __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0)
      __builtin_amdgcn_wave_barrier();
      float f[kB];
#pragma unroll
      for (int i = 0; i < kB; ++i) f[i] = s_tr[q * 33 + i];
      // Block spins of replica q: halves live in lanes q%32 (h=0) and q%32+32 (h=1)
      // of tile q/32.
      const uint32_t mine = (lane < 32) ? get_half<NW>(sa, c) : get_half<NW>(sb, c);
      const uint32_t give = (lane < 32) ? get_half<NW>(sb, c) : get_half<NW>(sa, c);
      const uint32_t got = (uint32_t)__shfl_xor((int)give, 32, 64);
      uint32_t bits = (lane < 32) ? block_of<I8>(mine, got) : block_of<I8>(got, mine);
      // ---- 3. the block's proposals (sequential order).  Flipping spin j
      // adds J[k][j] dS to the field of spin k: row j of J^T.
      const int pos0 = 32 * c;
      const float* Jb = p.Jt + (size_t)pos0 * n + pos0;
      uint32_t wrd[4] = {0, 0, 0, 0};
#pragma unroll
      for (int i = 0; i < kB; ++i) {
        const int pos = pos0 + i;
        if (pos >= n) break;
        if ((i & 3) == 0) {
          const u32x4 blk = sweep_block(cid, sw, (uint32_t)pos, p.k0, p.k1);
          wrd[0] = blk.x;
          wrd[1] = blk.y;
          wrd[2] = blk.z;
          wrd[3] = blk.w;
        }
        const bool up = (bits >> i) & 1u;
        const float fh = f[i] + p.h[pos];
        // dn = 2 s (F + h) (ising) or (2x - 1)(F + h) (qubo), exact in f32
        const float dn32 = (p.ptype == 0) ? (up ? 2.f : -2.f) * fh : (up ? fh : -fh);
        bool take = dn32 >= 0.f;
        if (!take) take = log_less(wrd[i & 3], beta * dn32, bd * (double)dn32);
        take = take && act;
        bits ^= (uint32_t)take << i;
        dsum += take ? -(double)dn32 : 0.0;
        acc += take ? 1u : 0u;
        const float dS = take ? ((p.ptype == 0) ? (up ? -2.f : 2.f) : (up ? -1.f : 1.f)) : 0.f;
        const float* jr = Jb + (size_t)i * n;
#pragma unroll
        for (int k = i + 1; k < kB; ++k)
          if (pos0 + k < n) f[k] += jr[k] * dS;
      }
      // ---- 4. new block spins back into operand layout
      const uint32_t h0 = half_of<I8>(bits, 0), h1 = half_of<I8>(bits, 1);
      // lane l keeps half h=l/32 of tile-0 replica l%32 and tile-1 replica 32+l%32
      const uint32_t other = (uint32_t)__shfl_xor((int)((lane < 32) ? h1 : h0), 32, 64);
      if (lane < 32) {
        set_half<NW>(sa, c, h0);     // tile 0, replica lane, half 0
        set_half<NW>(sb, c, other);  // tile 1, replica lane+32, half 0 (lane+32's h0)
      } else {
        set_half<NW>(sa, c, other);  // tile 0, replica lane-32, half 1 (lane-32's h1)
        set_half<NW>(sb, c, h1);     // tile 1, replica lane, half 1
      }
    }
    // ---- end of sweep: energy (simulation.py:157-158), best, PT, log-omega
    E = (float)((double)E + dsum);
    if (p.best_sem == 0) {  // simulation.py:161-164
      const bool imp = act && (E < bestE);
      if (imp) {
        bestE = E;
        p.best_rep_e[(size_t)read * R + rr] = E;
        p.best_rep_sw[(size_t)read * R + rr] = s;
      }
      const bool impa = __shfl((int)imp, qa, 64) != 0, impb = __shfl((int)imp, qb, 64) != 0;
      if (impa) store_state<NW, I8>(sa, nblk, n, p.ptype, p.best_rep + base_a);
      if (impb) store_state<NW, I8>(sb, nblk, n, p.ptype, p.best_rep + base_b);
    }
    s_E[q] = E;
    // PT log-uniforms, one lane per pair of this read (drawn ahead of the walk).
    if (p.use_pt && act && rr < R - 1)
      s_lu[q] = log_u32(pick_word(pt_block((uint64_t)read, sw, (uint32_t)rr, p.k0, p.k1),
                                  (uint32_t)rr & 3u));
    __builtin_amdgcn_s_waitcnt(0xc07f);
    __builtin_amdgcn_wave_barrier();
    int src = -1;
    if (act && rr == 0) {  // read leader
      const int b0 = g * R;
      if (p.best_sem == 1) {  // simulation.py:222-225 in replica order
        float cur = gE;
        for (int k = 0; k < R; ++k)
          if (s_E[b0 + k] < cur) {
            cur = s_E[b0 + k];
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
          const float dd = (s_E[b0 + s_bi[b0 + a]] - s_E[b0 + s_bi[b0 + b]]) *
                           (p.betas[a] - p.betas[b]);
          if (dd >= 0.f || s_lu[b0 + k] < (double)dd) {
            const int t = s_bi[b0 + a];
            s_bi[b0 + a] = s_bi[b0 + b];
            s_bi[b0 + b] = t;
          }
        }
      }
      for (int k = 0; k < R; ++k) s_rb[b0 + s_bi[b0 + k]] = p.betas[k];
      if (p.logw && p.log_omegas) {  // ais.py:99-112 over the ascending ladder
        double lw = 0.0, bold = 0.0;
        for (int r = 1; r < R; ++r) {
          const int k = p.bsort[r];
          const float be = p.betas[k];
          const float en = s_E[b0 + s_bi[b0 + k]];
          const float a = -be * en;
          lw += (double)a - (-bold * (double)en);
          bold = (double)be;
        }
        p.log_omegas[(size_t)read * p.S + s] = lw;
      }
      s_flag[g] = src;
    }
    __builtin_amdgcn_s_waitcnt(0xc07f);
    __builtin_amdgcn_wave_barrier();
    if (p.best_sem == 1) {
      // lanes holding replica src of each read copy it to best_state
      const int fa = (acta ? s_flag[ga] : -1), fb = (actb ? s_flag[gb] : -1);
      if (fa >= 0 && fa == qa - ga * R)
        store_state<NW, I8>(sa, nblk, n, p.ptype, p.best_state + (size_t)(read0 + ga) * n);
      if (fb >= 0 && fb == qb - gb * R)
        store_state<NW, I8>(sb, nblk, n, p.ptype, p.best_state + (size_t)(read0 + gb) * n);
    }
    __builtin_amdgcn_wave_barrier();
  }

  // ---- outputs
  if (acta) store_state<NW, I8>(sa, nblk, n, p.ptype, p.spins + base_a);
  if (actb) store_state<NW, I8>(sb, nblk, n, p.ptype, p.spins + base_b);
  if (act) {
    p.energies[(size_t)read * R + rr] = E;
    p.beta_idx[(size_t)read * R + rr] = s_bi[q];
    if (p.accepted) atomicAdd(p.accepted + read, acc);
    if (rr == 0 && p.best_sem == 1) {
      p.best_energy[read] = gE;
      p.best_sweeps[read] = gsw;
    }
  }
  if (p.best_sem == 1) {
    // never improved: best_state is still the view of states[0] (simulation.py:204)
    __builtin_amdgcn_wave_barrier();
    if (act && rr == 0) s_flag[g] = gimp;
    __builtin_amdgcn_s_waitcnt(0xc07f);
    __builtin_amdgcn_wave_barrier();
    if (acta && qa - ga * R == 0 && !s_flag[ga])
      store_state<NW, I8>(sa, nblk, n, p.ptype, p.best_state + (size_t)(read0 + ga) * n);
    if (actb && qb - gb * R == 0 && !s_flag[gb])
      store_state<NW, I8>(sb, nblk, n, p.ptype, p.best_state + (size_t)(read0 + gb) * n);
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

template <int NP, bool I8>
hipError_t launch_np(const BParams& p, int nw, int groups, hipStream_t st) {
  const size_t lds = (size_t)kLdsPerWave * kWavesPerGroup;
  dim3 grid(groups), block(64 * kWavesPerGroup);
  switch (nw) {
    case 4: hipLaunchKernelGGL((blocked_pt_kernel<NP, I8, 4>), grid, block, lds, st, p); break;
    case 8: hipLaunchKernelGGL((blocked_pt_kernel<NP, I8, 8>), grid, block, lds, st, p); break;
    case 16: hipLaunchKernelGGL((blocked_pt_kernel<NP, I8, 16>), grid, block, lds, st, p); break;
    case 32: hipLaunchKernelGGL((blocked_pt_kernel<NP, I8, 32>), grid, block, lds, st, p); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace

void blocked_release(BlockedProblem& b) {
  if (b.packed) (void)hipFree(b.packed);
  if (b.scratch) (void)hipFree(b.scratch);
  b = BlockedProblem();
}

bool blocked_eligible(int n, int dtype, int R, int reads, int order, bool replay) {
  return dtype == 0 && n >= 32 && n <= kBlockedMaxN && R >= 1 && R <= 64 && reads >= 1 &&
         order == kSequential && !replay;
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

int run_blocked(BlockedProblem& b, const void* dJ, const void* dJt, const void* dh, int n, int ptype, int dtype,
                int reads, int R, int S, int order, int use_pt, int logw, int best_sem, int pieces,
                uint64_t seed, uint64_t sweep_offset, const void* betas, const int32_t* bsort,
                const double* rep_lr, const double* rep_pt, int8_t* states, void* energies,
                int64_t* beta_idx, int8_t* best_state, void* best_energy, int32_t* best_sweeps,
                double* log_omegas, uint64_t* accepted, hipStream_t st, double* ms,
                std::string* why) {
  if (!blocked_eligible(n, dtype, R, reads, order, rep_lr != nullptr || rep_pt != nullptr)) {
    *why = "blocked form needs float32 couplings, 32 <= n <= 2048, R <= 64, sequential order "
           "and Philox randomness (no replay)";
    return PYSA_AMD_ENOTSUP;
  }
  const int nblk = (n + kB - 1) / kB;
  int nw = 4;
  while (2 * nw < nblk) nw *= 2;  // words of 2 blocks each: nblk <= 2 * nw
  // Coupling encoding: exact where possible; `pieces` forces 2 or 3 bf16 pieces.
  if (b.exact_kind < 0) {
    std::vector<float> hJ((size_t)n * n);
    if (hipMemcpy(hJ.data(), dJ, hJ.size() * 4, hipMemcpyDeviceToHost) != hipSuccess) {
      *why = "reading couplings back";
      return PYSA_AMD_EHIP;
    }
    b.exact_kind = blocked_exact_pieces(hJ.data(), hJ.size());
  }
  int kind = b.exact_kind;  // 0 int8, else bf16 pieces
  if (kind != 0) {
    if (pieces == 2 || pieces == 3)
      kind = (kind < pieces) ? kind : pieces;  // fewer pieces never help when more are exact
    else
      kind = (kind > 2) ? 2 : kind;  // default: hi+lo split for real couplings
    if (pieces == 3 && b.exact_kind >= 3) kind = 3;
  }
  const bool i8 = kind == 0;
  const int np = i8 ? 1 : kind;
  if (b.packed_kind != kind || !b.packed) {
    if (b.packed) (void)hipFree(b.packed);
    b.packed = nullptr;
    const size_t bytes = i8 ? (size_t)nblk * nblk * 64 * 16 : (size_t)nblk * 2 * nblk * np * 64 * 16;
    if (hipMalloc(&b.packed, bytes) != hipSuccess) {
      *why = "allocating packed couplings";
      return PYSA_AMD_EHIP;
    }
    b.packed_bytes = bytes;
    const long long threads = (long long)(bytes / 16);
    const unsigned nb = (unsigned)((threads + 255) / 256);
    if (i8)
      hipLaunchKernelGGL(pack_i8_kernel, dim3(nb), dim3(256), 0, st, (const float*)dJ, n, nblk,
                         (int8_t*)b.packed);
    else
      hipLaunchKernelGGL(pack_bf16_kernel, dim3(nb), dim3(256), 0, st, (const float*)dJ, n, nblk, np,
                         (uint16_t*)b.packed);
    if (hipGetLastError() != hipSuccess) {
      *why = "packing couplings";
      return PYSA_AMD_EHIP;
    }
    b.packed_kind = kind;
  }
  const size_t cnt = (size_t)reads * R;
  const size_t need = cnt * n + cnt * 4 + cnt * 4;
  if (b.scratch_bytes < need) {
    if (b.scratch) (void)hipFree(b.scratch);
    b.scratch = nullptr;
    if (hipMalloc(&b.scratch, need) != hipSuccess) {
      *why = "allocating scratch";
      return PYSA_AMD_EHIP;
    }
    b.scratch_bytes = need;
  }
  BParams p;
  p.n = n;
  p.nblk = nblk;
  p.ptype = ptype;
  p.R = R;
  p.reads = reads;
  p.S = S;
  p.rpw = 64 / R;
  p.np = np;
  p.use_pt = use_pt;
  p.logw = logw;
  p.best_sem = best_sem;
  p.refresh = 0;
  p.k0 = (uint32_t)seed;
  p.k1 = (uint32_t)(seed >> 32);
  p.sweep_offset = sweep_offset;
  p.frag = b.packed;
  p.Jt = (const float*)dJt;
  p.h = (const float*)dh;
  p.betas = (const float*)betas;
  p.bsort = bsort;
  p.spins = states;
  p.energies = (float*)energies;
  p.beta_idx = beta_idx;
  p.best_rep = (int8_t*)b.scratch;
  p.best_rep_e = (float*)((int8_t*)b.scratch + cnt * n);
  p.best_rep_sw = (int*)((int8_t*)b.scratch + cnt * n + cnt * 4);
  p.best_state = best_state;
  p.best_energy = (float*)best_energy;
  p.best_sweeps = best_sweeps;
  p.log_omegas = logw ? log_omegas : nullptr;
  p.accepted = (unsigned long long*)accepted;
  if (best_sem == 0 &&
      hipMemcpyAsync(p.best_rep, states, cnt * n, hipMemcpyDeviceToDevice, st) != hipSuccess) {
    *why = "initialising best states";
    return PYSA_AMD_EHIP;
  }
  if (accepted && hipMemsetAsync(accepted, 0, sizeof(uint64_t) * reads, st) != hipSuccess) {
    *why = "clearing counters";
    return PYSA_AMD_EHIP;
  }
  const int waves = (reads + p.rpw - 1) / p.rpw;
  const int groups = (waves + kWavesPerGroup - 1) / kWavesPerGroup;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    *why = "creating events";
    return PYSA_AMD_EHIP;
  }
  (void)hipEventRecord(e0, st);
  hipError_t le;
  if (i8)
    le = launch_np<1, true>(p, nw, groups, st);
  else if (np == 1)
    le = launch_np<1, false>(p, nw, groups, st);
  else if (np == 2)
    le = launch_np<2, false>(p, nw, groups, st);
  else
    le = launch_np<3, false>(p, nw, groups, st);
  if (le == hipSuccess && best_sem == 0) {
    hipLaunchKernelGGL(best_argmin_kernel, dim3((reads + 63) / 64), dim3(64), 0, st, p);
    le = hipGetLastError();
  }
  (void)hipEventRecord(e1, st);
  if (le != hipSuccess) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *why = std::string("blocked kernel launch: ") + hipGetErrorString(le);
    return PYSA_AMD_EHIP;
  }
  if (hipEventSynchronize(e1) != hipSuccess) {
    *why = "blocked kernel execution";
    return PYSA_AMD_EHIP;
  }
  float t = 0.f;
  (void)hipEventElapsedTime(&t, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (ms) *ms = t;
  b.last_kind = kind;
  return 0;
}

}  // namespace pysa
