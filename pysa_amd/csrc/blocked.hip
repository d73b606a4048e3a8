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
#include "blocked_pipe.hpp"
#include "engine.hpp"
#include "rng.hpp"

// Best-state scratch of the allocation-free entry point
// (pysa_amd_metropolis_device_noalloc): one copy per device, held until the
// process ends.  Host code never uses this name as an address: it passes the
// device address that noalloc_scratch() looks up.
static __device__ __align__(256) uint8_t scratch[pysa::kNoAllocScratchBytes];

namespace pysa {
using namespace blk;
namespace {


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
  int nblk, groups, kw, kb, nw, kbp, nwp, kind, np, nt;
  bool pipe, i8, inexact;
};

// Compute units of the current device (cached per device).
int device_cus() {
  static std::mutex mu;
  static int cus[kMaxDevices] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return 256;
  std::lock_guard<std::mutex> lk(mu);
  if (!cus[dev]) {
    int v = 0;
    cus[dev] = (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) ==
                    hipSuccess && v > 0) ? v : 256;
  }
  return cus[dev];
}

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
  // The problem's launch shape (pysa_amd_problem_set_shape) picks the kernel
  // where the size allows both; the split-K kernel has no block-random order.
  const bool want_splitk = force_splitk || b.shape_kernel == 2;
  pl.pipe = order == kBlockRandom || (pl.nblk >= kPipeMW && (b.shape_kernel == 1 || !want_splitk));
  // Tiles of 32 replicas per workgroup of the pipelined kernel: two, unless
  // two-tile groups leave compute units idle and a read fits one tile (one
  // workgroup per CU either way: LDS and registers hold one).
  const int groups2 = (reads + (64 / R) - 1) / (64 / R);
  if (R > 32)
    pl.nt = 2;
  else if (b.shape_tiles == 1 || b.shape_tiles == 2)
    pl.nt = b.shape_tiles;
  else
    pl.nt = groups2 < device_cus() ? 1 : 2;
  if (pl.pipe) pl.groups = (reads + (32 * pl.nt / R) - 1) / (32 * pl.nt / R);
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
  BParams p{};
  p.n = n;
  p.nblk = pl.nblk;
  p.ptype = ptype;
  p.R = R;
  p.reads = reads;
  p.S = S;
  p.rpw = pl.pipe ? 32 * pl.nt / R : 64 / R;
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
  // Dry passes every 256 sweeps (the C3 batch's tracked energies stay within
  // 3e-6 of f64 over 5000 sweeps; every 64 sweeps cost 1.3 % more time for
  // 2e-6).  PYSA_AMD_ENERGY_REFRESH=<sweeps> overrides the interval for every
  // problem (0 = never), a switch for tests and comparisons.
  const char* ev = getenv("PYSA_AMD_ENERGY_REFRESH");
  p.refresh = ev && ev[0] ? atoi(ev) : (b.exact_sums == 1 ? 0 : 256);
  p.hzero = b.h_zero == 1 ? 1 : 0;
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
      e = launch_pipe_i8(p, pl.nwp, pl.nt, pl.groups, st);
    else if (np == 1)
      e = launch_pipe_bf16_1(p, pl.nwp, pl.nt, pl.groups, st);
    else if (np == 2)
      e = launch_pipe_bf16_2(p, pl.nwp, pl.nt, pl.groups, st);
    else
      e = launch_pipe_bf16_3(p, pl.nwp, pl.nt, pl.groups, st);
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
    b.h_zero = 1;
    for (int i = 0; i < n; ++i)
      if (hh[i] != 0.f) b.h_zero = 0;
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

int run_blocked_ais(BlockedProblem& b, const void* dJ, const void* dJt, const void* dh, int n,
                    int ptype, int chains, int nt, const float* temps, int order, int ordering,
                    int pieces, uint64_t seed, int8_t* states, double* energy, double* log_w,
                    hipStream_t st, double* ms, std::string* why) {
  if (!blocked_eligible(n, 0, 1, chains, order, false) || n <= 2 * kB) {
    *why = "MFMA AIS needs float32 couplings, 65 <= n <= 2048, sequential or block-random order, "
           "the uniform reference and Philox randomness";
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
    b.h_zero = 1;
    for (int i = 0; i < n; ++i)
      if (hh[i] != 0.f) b.h_zero = 0;
  }
  // one chain per read (R = 1: the chain index is the oracle's), two tiles
  // of 32 chains per workgroup on the pipelined kernel
  BlockedPlan pl = plan_blocked(b, n, chains, 1, order, pieces);
  pl.pipe = true;
  pl.nt = 2;
  pl.groups = (chains + 63) / 64;
  int rc = prepare_encoding(b, dJ, dJt, n, ptype, pl, st, why);
  if (rc) return rc;
  BParams p{};
  p.n = n;
  p.nblk = pl.nblk;
  p.ptype = ptype;
  p.R = 1;
  p.reads = chains;
  p.S = nt;
  p.rpw = 64;
  p.np = pl.np;
  p.use_pt = 0;
  p.logw = 0;
  p.best_sem = 0;
  p.order = order;
  // dry passes recompute the energy every 64 steps unless every sum is exact
  const char* ev = getenv("PYSA_AMD_ENERGY_REFRESH");
  p.refresh = ev && ev[0] ? atoi(ev) : (b.exact_sums == 1 ? 0 : 64);
  p.hzero = b.h_zero == 1 ? 1 : 0;
  p.kb = pl.kbp;
  p.k0 = (uint32_t)seed;
  p.k1 = (uint32_t)(seed >> 32);
  p.sweep_offset = 0;
  p.frag = b.packed;
  p.Jt = pl.inexact ? b.jtl : (const float*)dJt;
  p.h = (const float*)dh;
  p.rowsum = (!pl.i8 && ptype == 1) ? b.rowsum : nullptr;
  p.spins = states;
  p.temps = temps;
  p.ordering = ordering;
  p.ais_e = energy;
  p.ais_w = log_w;
  EventPair ev2;
  if (ev2.create() != hipSuccess) {
    *why = "creating events";
    return PYSA_AMD_EHIP;
  }
  (void)hipEventRecord(ev2.a, st);
  hipError_t le;
  if (pl.i8)
    le = launch_pipe_ais_i8(p, pl.nwp, pl.groups, st);
  else if (pl.np == 1)
    le = launch_pipe_ais_bf16_1(p, pl.nwp, pl.groups, st);
  else if (pl.np == 2)
    le = launch_pipe_ais_bf16_2(p, pl.nwp, pl.groups, st);
  else
    le = launch_pipe_ais_bf16_3(p, pl.nwp, pl.groups, st);
  (void)hipEventRecord(ev2.b, st);
  if (le != hipSuccess) {
    *why = std::string("MFMA AIS launch: ") + hipGetErrorString(le);
    (void)hipStreamSynchronize(st);
    return PYSA_AMD_EHIP;
  }
  if (hipEventSynchronize(ev2.b) != hipSuccess) {
    *why = "MFMA AIS execution";
    return PYSA_AMD_EHIP;
  }
  float t = 0.f;
  (void)hipEventElapsedTime(&t, ev2.a, ev2.b);
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
