// Batched form: many independent dense instances of different sizes, all
// reads of all of them in one persistent launch.
//
// PySA has no batching across instances; a batch of problems is a Python
// loop over Solver objects, each running its reads one after another
// (pysa/sa.py:425-427 inside a caller's loop).  Per instance the chain is the
// one of pysa/simulation.py:123-236 (sweeps :41-76, update_spin
// pysa/ising.py:40-63 / pysa/qubo.py:40-63, exchange pysa/utils.py:26-58).
//
// Couplings are J = scale * S with S a symmetric zero-diagonal int8 matrix
// (+-1 signs for Sherrington-Kirkpatrick instances).  Local fields are kept
// as exact integers c = S s, so the field of spin i is scale*c_i + h_i,
// rounded once in float32 -- the correctly rounded value of the reference's
// float32 row dot J_i . s (which rounds at every partial sum in an order set
// by BLAS) -- and the fields never drift.
//
// Layout.  One workgroup per (instance, read), one wave per replica (R <= 16
// waves); a replica's integer fields stay in VGPRs for the whole run, with
// element i in lane (i>>2)&63, slot ((i>>8)<<2)|(i&3), so an accepted flip
// adds one int8 coupling row, 4 bytes per lane per 256 elements, to every
// field.  Workgroups are issued largest instance first so the tail of the
// launch holds the short ones.
#include <stdint.h>

#include "engine.hpp"
#include "rng.hpp"

namespace pysa {
namespace {

__device__ __forceinline__ int rl(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }
__device__ __forceinline__ uint32_t rl(uint32_t v, int lane) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, lane);
}
__device__ __forceinline__ double rl(double v, int lane) {
  const uint64_t b = (uint64_t)__double_as_longlong(v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, lane);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), lane);
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

__device__ __forceinline__ int elem_of(int lane, int g) {
  return ((g >> 2) << 8) + (lane << 2) + (g & 3);
}

// Integer fields and spin bits of one replica (G <= 32 slots per lane).
template <int G>
struct IChain {
  int f[G];
  uint32_t bits;  // bit g: element elem_of(lane, g) is up (+1 / 1)

  __device__ __forceinline__ int pick(int g) const {
    int v[G];
#pragma unroll
    for (int k = 0; k < G; ++k) v[k] = f[k];
#pragma unroll
    for (int half = G / 2; half >= 1; half >>= 1) {
      const bool hi = (g & half) != 0;
#pragma unroll
      for (int k = 0; k < half; ++k) v[k] = hi ? v[k + half] : v[k];
    }
    return v[0];
  }
  // f += d * row (row: int8, zero beyond n, 16-byte aligned rows)
  __device__ __forceinline__ void add_row(const int8_t* __restrict__ row, int n, int d) {
    const int lane = threadIdx.x & 63;
    int x[G / 4];
#pragma unroll
    for (int g4 = 0; g4 < G / 4; ++g4) {
      const int i0 = (g4 << 8) + (lane << 2);
      x[g4] = (i0 < n) ? *reinterpret_cast<const int*>(row + i0) : 0;
    }
#pragma unroll
    for (int g4 = 0; g4 < G / 4; ++g4)
#pragma unroll
      for (int q = 0; q < 4; ++q) f[4 * g4 + q] += d * ((x[g4] << (24 - 8 * q)) >> 24);
  }
};

template <int G>
__global__ __launch_bounds__(G >= 32 ? 512 : 1024) void batched_pt_kernel(BatchParams p) {
  const int item = p.order_map ? p.order_map[blockIdx.x] : (int)blockIdx.x;
  const int inst = item / p.reads, rd = item - inst * p.reads;
  const int tid = threadIdx.x, lane = tid & 63, r = tid >> 6;
  const int R = p.R, n = p.inst_n[inst], ld = p.ld[inst];
  const float sc = p.scale[inst];
  const int8_t* __restrict__ S = p.signs + p.j_off[inst];
  const int64_t so = p.s_off[inst];                      // states of this instance
  const int64_t gread = (int64_t)inst * p.reads + rd;     // global read index
  int8_t* spins = p.states + so + (int64_t)rd * R * n;
  int8_t* brep = p.best_rep + so + (int64_t)rd * R * n;
  int8_t* bstate = p.best_state + so / R + (int64_t)rd * n;
  const float* betas = p.betas + (int64_t)inst * R;

  __shared__ float s_h[kBatchedMaxN];
  __shared__ float s_E[16], s_beta[16], s_bs[16], s_bE[16];
  __shared__ long long s_bi[16];
  __shared__ double s_lu[16];
  __shared__ int s_bsw[16], s_src, s_gsw, s_imp;
  __shared__ float s_gE;
  for (int i = tid; i < n; i += blockDim.x) s_h[i] = p.h ? p.h[p.h_off[inst] + i] : 0.f;
  if (tid < R) {
    s_E[tid] = p.energies[gread * R + tid];
    s_bE[tid] = s_E[tid];
    s_bsw[tid] = 0;
    s_bi[tid] = tid;  // beta_idx = arange(R) (sa.py:325)
    s_beta[tid] = betas[tid];
  }
  if (tid == 0) {
    s_gE = p.energies[gread * R];  // simulation.py:205
    s_gsw = 0;
    s_imp = 0;
  }

  // Load the spins and build the exact integer fields c = S s.
  IChain<G> ch;
  ch.bits = 0;
#pragma unroll
  for (int g = 0; g < G; ++g) {
    ch.f[g] = 0;
    const int i = elem_of(lane, g);
    if (i < n && spins[(int64_t)r * n + i] == 1) ch.bits |= 1u << g;
  }
  for (int j = 0; j < n; ++j) {
    const int sj = spins[(int64_t)r * n + j];
    if (sj != 0) ch.add_row(S + (int64_t)j * ld, n, sj);
  }
  if (p.best_sem == 0)
    for (int i = lane; i < n; i += 64) brep[(int64_t)r * n + i] = spins[(int64_t)r * n + i];
  __syncthreads();

  const uint64_t cid = (uint64_t)gread * R + r;
  unsigned long long acc = 0;
  for (int s = 0; s < p.S; ++s) {
    const uint64_t sw = (uint64_t)s + p.sweep_offset;
    if (tid < R) s_bs[s_bi[tid]] = s_beta[tid];  // simulation.py:151-152
    if (p.use_pt && tid < R - 1)
      s_lu[tid] = log_u32(pick_word(pt_block((uint64_t)gread, sw, (uint32_t)tid, p.k0, p.k1),
                                    (uint32_t)tid & 3u));
    __syncthreads();
    // ---- one sweep of replica r (simulation.py:41-76)
    const float beta = s_bs[r];
    const double bd = (double)beta;
    FeistelKey fk;
    const bool brand = p.order == kBlockRandom;
    const int nv = brand ? kOrderBlock * ((n + kOrderBlock - 1) / kOrderBlock) : n;
    if (brand)
      fk = block_order_key((uint32_t)((n + kOrderBlock - 1) / kOrderBlock), sw, p.k0, p.k1);
    else if (p.order != kSequential)
      fk = feistel_make((uint32_t)n, perm_keys(p.order == kRandomShared ? 0ull : cid, sw,
                                               p.order == kRandomShared ? kStreamShared
                                                                        : kStreamPerm,
                                               p.k0, p.k1));
    double sum = 0.0;
    for (int w = 0; w < nv; w += 64) {
      const int vi = w + lane;
      int pos_l = -1;
      double lr_l = 0.0;
      if (vi < nv) {
        if (brand)
          pos_l = block_random_pos(fk, sw, (uint32_t)vi, n, p.k0, p.k1);
        else
          pos_l = (p.order == kSequential) ? vi : (int)feistel_perm(fk, (uint32_t)vi);
        if (pos_l >= 0)
          lr_l = log_u32(pick_word(sweep_block(cid, sw, (uint32_t)pos_l, p.k0, p.k1),
                                   (uint32_t)pos_l & 3u));
      }
      const float hl = (pos_l >= 0) ? s_h[pos_l] : 0.f;
      const int cnt = min(64, nv - w);
      for (int t = 0; t < cnt; ++t) {
        const int pos = rl(pos_l, t);
        if (pos < 0) continue;  // padding slot (wave-uniform)
        const double lr = rl(lr_l, t);
        const int pl = (pos >> 2) & 63, ps = ((pos >> 8) << 2) | (pos & 3);
        const bool up = (rl(ch.bits, pl) >> ps) & 1u;
        const int c = rl(ch.pick(ps), pl);
        const float fh = (float)c * sc + __int_as_float(rl(__float_as_int(hl), t));
        // dn = 2 s (F + h) (ising.py:48-49) or (2x - 1)(F + h) (qubo.py:48-49)
        const double dn = (p.ptype == 0 ? (up ? 2.0 : -2.0) : (up ? 1.0 : -1.0)) * (double)fh;
        if (dn >= 0.0 || lr < bd * dn) {  // ising.py:52
          sum += -dn;
          ++acc;
          if (lane == pl) ch.bits ^= 1u << ps;
          ch.add_row(S + (int64_t)pos * ld, n, p.ptype == 0 ? (up ? -2 : 2) : (up ? -1 : 1));
        }
      }
    }
    const float en = (float)((double)s_E[r] + sum);  // simulation.py:157-158
    if (lane == 0) s_E[r] = en;
    if (p.best_sem == 0 && en < s_bE[r]) {  // simulation.py:161-164
#pragma unroll
      for (int g = 0; g < G; ++g) {
        const int i = elem_of(lane, g);
        if (i < n)
          brep[(int64_t)r * n + i] = ((ch.bits >> g) & 1u) ? 1 : (p.ptype == 0 ? -1 : 0);
      }
      if (lane == 0) {
        s_bE[r] = en;
        s_bsw[r] = s;
      }
    }
    __syncthreads();
    if (tid == 0) {
      int src = -1;
      if (p.best_sem == 1) {  // simulation.py:222-225 in replica order
        float cur = s_gE;
        for (int k = 0; k < R; ++k)
          if (s_E[k] < cur) {
            cur = s_E[k];
            src = k;
          }
        if (src >= 0) {
          s_gE = cur;
          s_gsw = s;
          s_imp = 1;
        }
      }
      s_src = src;
      if (p.use_pt) {  // utils.py:44-58
        for (int k = 0; k < R - 1; ++k) {
          const int a = R - k - 1, b = R - k - 2;
          const float dd = (s_E[s_bi[a]] - s_E[s_bi[b]]) * (s_beta[a] - s_beta[b]);
          if (dd >= 0.f || s_lu[k] < (double)dd) {
            const long long t = s_bi[a];
            s_bi[a] = s_bi[b];
            s_bi[b] = t;
          }
        }
      }
    }
    __syncthreads();
    if (p.best_sem == 1 && s_src == r) {  // the improving replica copies itself
#pragma unroll
      for (int g = 0; g < G; ++g) {
        const int i = elem_of(lane, g);
        if (i < n) bstate[i] = ((ch.bits >> g) & 1u) ? 1 : (p.ptype == 0 ? -1 : 0);
      }
    }
  }

  // Outputs.
#pragma unroll
  for (int g = 0; g < G; ++g) {
    const int i = elem_of(lane, g);
    if (i < n) spins[(int64_t)r * n + i] = ((ch.bits >> g) & 1u) ? 1 : (p.ptype == 0 ? -1 : 0);
  }
  if (p.accepted && lane == 0) atomicAdd(p.accepted + gread, acc);
  __syncthreads();
  if (tid < R) {
    p.beta_idx[gread * R + tid] = s_bi[tid];
    p.energies[gread * R + tid] = s_E[tid];
  }
  if (tid == 0) {
    if (p.best_sem == 0) {  // simulation.py:174-177, first minimum wins
      int bp = 0;
      for (int k = 1; k < R; ++k)
        if (s_bE[k] < s_bE[bp]) bp = k;
      s_src = bp;
      p.best_energy[gread] = s_bE[bp];
      p.best_sweeps[gread] = s_bsw[bp];
    } else {
      s_src = s_imp ? -1 : 0;  // not improved: best_state is the view of states[0] (:204)
      p.best_energy[gread] = s_gE;
      p.best_sweeps[gread] = s_gsw;
    }
  }
  __syncthreads();
  if (s_src >= 0) {
    const int8_t* src = (p.best_sem == 0 ? brep : spins) + (int64_t)s_src * n;
    for (int i = tid; i < n; i += blockDim.x) bstate[i] = src[i];
  }
}

}  // namespace

int batched_regs_for(int max_n) {
  int g = 4;
  while (64 * g < max_n) g *= 2;
  return g;
}

hipError_t launch_batched(const BatchParams& p, int items, int max_n, hipStream_t st) {
  if (items < 1) return hipSuccess;
  if (max_n < 1 || max_n > kBatchedMaxN || p.R < 1 || p.R > 16) return hipErrorInvalidValue;
  if (batched_regs_for(max_n) >= 32 && p.R > 8) return hipErrorInvalidValue;
  const dim3 grid((unsigned)items), block((unsigned)(64 * p.R));
  switch (batched_regs_for(max_n)) {
    case 4: hipLaunchKernelGGL(batched_pt_kernel<4>, grid, block, 0, st, p); break;
    case 8: hipLaunchKernelGGL(batched_pt_kernel<8>, grid, block, 0, st, p); break;
    case 16: hipLaunchKernelGGL(batched_pt_kernel<16>, grid, block, 0, st, p); break;
    case 32: hipLaunchKernelGGL(batched_pt_kernel<32>, grid, block, 0, st, p); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace pysa
