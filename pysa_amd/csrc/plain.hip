// Plain form of the engine: the drop-in for PySA's Metropolis sweep.
//
// Reference path (what these kernels replace):
//   pysa/simulation.py:123-236  simulation_parallel / simulation_sequential
//   pysa/simulation.py:41-76    random_sweep / sequential_sweep
//   pysa/ising.py:40-63, pysa/qubo.py:40-63  update_spin
//   pysa/utils.py:26-58         pt
//   pysa/ais.py:79-112          get_log_omega
//   pysa/simulation.py:266-317  ais_simulation_parallel / ais_sample
//
// Layout: one workgroup per read; each wave owns one replica at a time.  The
// replica's local fields F = J s live in VGPRs (lane l holds F[64m + l] in
// register m) and its spins as one bit per register slot.  A proposal reads
// F[pos] with a readlane, so dE is O(1); an accepted flip streams one
// coupling row (coalesced, 64 lanes x 4 bytes per step) and updates every
// field.  The sweep loop, best tracking, parallel tempering and log-omega
// run inside one persistent launch; replicas of a read meet at a workgroup
// barrier once per sweep.
#include "engine.hpp"
#include "rng.hpp"

namespace pysa {
namespace {

__device__ __forceinline__ int rl(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }
__device__ __forceinline__ float rl(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
__device__ __forceinline__ uint64_t rl(uint64_t v, int lane) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, lane);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), lane);
  return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ double rl(double v, int lane) {
  return __longlong_as_double((long long)rl((uint64_t)__double_as_longlong(v), lane));
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ int wave_sum(int v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// One chain held by one wave.
template <typename T, int CH>
struct Chain {
  T f[CH];        // f[m] at lane l = F[64m + l]
  uint64_t bits;  // bit m at lane l: spin 64m + l is up (+1 ising / 1 qubo)

  __device__ void load(const T* F, const int8_t* s, int n) {
    const int lane = threadIdx.x & 63;
    bits = 0;
#pragma unroll
    for (int m = 0; m < CH; ++m) {
      const int i = 64 * m + lane;
      f[m] = (i < n) ? F[i] : T(0);
      if (i < n && s[i] == 1) bits |= (1ull << m);
    }
  }
  __device__ void store(T* F, int8_t* s, int n, int ptype) const {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int m = 0; m < CH; ++m) {
      const int i = 64 * m + lane;
      if (i < n) {
        F[i] = f[m];
        s[i] = ((bits >> m) & 1) ? int8_t(1) : (ptype == 0 ? int8_t(-1) : int8_t(0));
      }
    }
  }
  __device__ void store_spins(int8_t* s, int n, int ptype) const {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int m = 0; m < CH; ++m) {
      const int i = 64 * m + lane;
      if (i < n) s[i] = ((bits >> m) & 1) ? int8_t(1) : (ptype == 0 ? int8_t(-1) : int8_t(0));
    }
  }

  // F = J s accumulated in f64 and rounded once; returns E = s.(F/2 + h) in
  // f64 (the periodic recomputation that bounds drift).
  __device__ double refresh(const T* __restrict__ Jt, const T* __restrict__ h, int n, int ptype) {
    const int lane = threadIdx.x & 63;
    double e = 0.0;
#pragma unroll
    for (int m = 0; m < CH; ++m) {
      const int i = 64 * m + lane;
      double a = 0.0;
      for (int jc = 0; 64 * jc < n; ++jc) {
        const int jl = 64 * jc + lane;
        const int sv_l = (jl < n) ? (((bits >> jc) & 1) ? 1 : (ptype == 0 ? -1 : 0)) : 0;
        const int cnt = min(64, n - 64 * jc);
        for (int t = 0; t < cnt; ++t) {
          const int sj = rl(sv_l, t);
          if (sj == 0) continue;
          if (i < n) a += (double)Jt[(size_t)(64 * jc + t) * n + i] * (double)sj;
        }
      }
      f[m] = (T)a;
      if (i < n) {
        const double si = ((bits >> m) & 1) ? 1.0 : (ptype == 0 ? -1.0 : 0.0);
        e += si * (0.5 * a + (double)h[i]);
      }
    }
    return wave_sum(e);
  }

  __device__ int count_up(int n) const {
    const int lane = threadIdx.x & 63;
    int c = 0;
#pragma unroll
    for (int m = 0; m < CH; ++m)
      if (64 * m + lane < n) c += (int)((bits >> m) & 1);
    return wave_sum(c);
  }

  // One Metropolis sweep at inverse temperature beta
  // (simulation.py:60-76 / 41-57 with ising.py:40-63 / qubo.py:40-63).
  // AIS adds the reference-distribution term of update_spin_ref
  // (ising.py:74-83).  Returns the f64 sum of accepted energy changes.
  template <bool AIS>
  __device__ double sweep(int n, int ptype, const T* __restrict__ Jt, const T* __restrict__ h,
                          T beta, uint64_t chain, uint64_t sweep_ctr, int order, uint32_t k0,
                          uint32_t k1, const double* __restrict__ rep_lr,
                          const int* __restrict__ rep_perm, int ref, double lratio,
                          unsigned long long& acc, int& ones_delta) {
    const int lane = threadIdx.x & 63;
    FeistelKey fk;
    if (order != kSequential && rep_perm == nullptr)
      fk = feistel_make((uint32_t)n, perm_keys(order == kRandomShared ? 0ull : chain, sweep_ctr,
                                               order == kRandomShared ? kStreamShared : kStreamPerm,
                                               k0, k1));
    const double bd = (double)beta;
    double sum = 0.0;
    for (int c = 0; c < CH; ++c) {
      if (64 * c >= n) break;
      const int i = 64 * c + lane;
      int pos_l = 0;
      double lr_l = 0.0;
      if (i < n) {
        if (rep_perm)
          pos_l = rep_perm[i];
        else
          pos_l = (order == kSequential) ? i : (int)feistel_perm(fk, (uint32_t)i);
        if (rep_lr)
          lr_l = rep_lr[pos_l];
        else
          lr_l = log_u32(pick_word(sweep_block(chain, sweep_ctr, (uint32_t)pos_l, k0, k1),
                                   (uint32_t)pos_l & 3u));
      }
      const int cnt = min(64, n - 64 * c);
      for (int t = 0; t < cnt; ++t) {
        const int pos = rl(pos_l, t);
        const double lr = rl(lr_l, t);
        const int m = pos >> 6, ln = pos & 63;
        const T fv = rl(f[m], ln);
        const bool up = (rl(bits, ln) >> m) & 1;
        const T fh = fv + h[pos];
        const double dn = (ptype == 0) ? 2.0 * (up ? 1.0 : -1.0) * (double)fh
                                       : (2.0 * (up ? 1.0 : 0.0) - 1.0) * (double)fh;
        bool accept;
        if (AIS) {
          const double refd = (ref == 1) ? (up ? lratio : -lratio) : 0.0;
          const double full = bd * dn + (1.0 - bd) * refd;
          accept = full >= 0 || lr < full;
        } else {
          accept = dn >= 0 || lr < bd * dn;
        }
        if (accept) {
          sum += -dn;
          ++acc;
          ones_delta += up ? -1 : 1;
          const T dS = (ptype == 0) ? (up ? T(-2) : T(2)) : (up ? T(-1) : T(1));
          if (lane == ln) bits ^= (1ull << m);
          const T* row = Jt + (size_t)pos * n;
#pragma unroll
          for (int mm = 0; mm < CH; ++mm) {
            const int idx = 64 * mm + lane;
            if (idx < n) f[mm] += row[idx] * dS;
          }
        }
      }
    }
    return sum;
  }
};

constexpr int kMaxR = 1024;

// Waves per workgroup for a chain of CH field registers of type T: keeps the
// fields plus working registers inside the VGPR budget at that occupancy.
template <typename T, int CH>
constexpr int waves_for() {
  return (CH * (int)(sizeof(T) / 4) <= 16) ? 16 : (CH * (int)(sizeof(T) / 4) <= 32) ? 8 : 4;
}

template <typename T, int CH>
__global__ __launch_bounds__((64 * waves_for<T, CH>())) void plain_pt_kernel(PTParams<T> p) {
  const int read = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, W = blockDim.x >> 6;
  const int R = p.R, n = p.n;
  __shared__ T s_bs[kMaxR];
  __shared__ T s_beta[kMaxR];
  __shared__ T s_E[kMaxR];
  __shared__ long long s_bi[kMaxR];
  __shared__ double s_ptlu[kMaxR];
  __shared__ int s_src;
  __shared__ T s_gE;
  __shared__ int s_gsw, s_improved;

  for (int k = tid; k < R; k += blockDim.x) {
    s_bi[k] = p.beta_idx[(size_t)read * R + k];
    s_E[k] = p.energies[(size_t)read * R + k];
    s_beta[k] = p.betas[k];
  }
  if (tid == 0) {
    s_gE = p.energies[(size_t)read * R];  // simulation.py:205
    s_gsw = 0;
    s_improved = 0;
  }
  // Local fields of every replica, F = J s in f64 rounded once.
  for (int r = wave; r < R; r += W) {
    Chain<T, CH> ch;
    const size_t off = ((size_t)read * R + r) * n;
    ch.load(p.fields + off, p.spins + off, n);
    ch.refresh(p.Jt, p.h, n, p.ptype);
    ch.store(p.fields + off, p.spins + off, n, p.ptype);
  }
  __syncthreads();

  unsigned long long acc = 0;
  for (int s = 0; s < p.S; ++s) {
    const uint64_t sw = (uint64_t)s + p.sweep_offset;
    for (int k = tid; k < R; k += blockDim.x) s_bs[s_bi[k]] = s_beta[k];  // simulation.py:151-152
    // PT log-uniforms, one per pair, drawn in parallel ahead of the serial walk.
    if (p.use_pt)
      for (int k = tid; k < R - 1; k += blockDim.x) {
        double lu;
        if (p.rep_pt)
          lu = p.rep_pt[((size_t)read * p.S + s) * (R - 1) + k];
        else
          lu = log_u32(pick_word(pt_block((uint64_t)read, sw, (uint32_t)k, p.k0, p.k1), (uint32_t)k & 3u));
        s_ptlu[k] = lu;
      }
    __syncthreads();
    for (int r = wave; r < R; r += W) {  // simulation.py:154-164
      Chain<T, CH> ch;
      const size_t off = ((size_t)read * R + r) * n;
      ch.load(p.fields + off, p.spins + off, n);
      const size_t rb = (((size_t)read * p.S + s) * R + r) * (size_t)n;
      int od = 0;
      const double d = ch.template sweep<false>(
          n, p.ptype, p.Jt, p.h, s_bs[r], (uint64_t)read * R + r, sw, p.order, p.k0, p.k1,
          p.rep_lr ? p.rep_lr + rb : nullptr, p.rep_perm ? p.rep_perm + rb : nullptr, 0, 0.0, acc,
          od);
      ch.store(p.fields + off, p.spins + off, n, p.ptype);
      const T en = (T)((double)s_E[r] + d);  // simulation.py:157-158
      if (lane == 0) s_E[r] = en;
      if (p.best_sem == 0) {
        const size_t bi = (size_t)read * R + r;
        if (en < p.best_rep_e[bi]) {  // simulation.py:161-164
          ch.store_spins(p.best_rep + bi * n, n, p.ptype);
          if (lane == 0) {
            p.best_rep_e[bi] = en;
            p.best_rep_sw[bi] = s;
          }
        }
      }
    }
    __syncthreads();
    if (tid == 0) {
      if (p.best_sem == 1) {  // simulation.py:222-225 in replica order
        int src = -1;
        T cur = s_gE;
        for (int k = 0; k < R; ++k)
          if (s_E[k] < cur) {
            cur = s_E[k];
            src = k;
          }
        if (src >= 0) {
          s_gE = cur;
          s_gsw = s;
          s_improved = 1;
        }
        s_src = src;
      }
      if (p.use_pt) {  // utils.py:44-58
        for (int k = 0; k < R - 1; ++k) {
          const int a = R - k - 1, b = R - k - 2;
          const T dd = (T)(s_E[s_bi[a]] - s_E[s_bi[b]]) * (T)(s_beta[a] - s_beta[b]);
          if (dd >= 0 || s_ptlu[k] < (double)dd) {
            const long long t = s_bi[a];
            s_bi[a] = s_bi[b];
            s_bi[b] = t;
          }
        }
      }
      if (p.logw && p.log_omegas) {  // ais.py:99-112 over the ascending ladder
        double lw = 0.0, bold = 0.0;
        for (int r = 1; r < R; ++r) {
          const int q = p.bsort[r];
          const T be = s_beta[q];
          const T en = s_E[s_bi[q]];
          const T a = -be * en;
          lw += (double)a - (-bold * (double)en);
          bold = (double)be;
        }
        p.log_omegas[(size_t)read * p.S + s] = lw;
      }
    }
    __syncthreads();
    if (p.best_sem == 1 && s_src >= 0) {
      const int8_t* src = p.spins + ((size_t)read * R + s_src) * n;
      for (int i = tid; i < n; i += blockDim.x) p.best_state[(size_t)read * n + i] = src[i];
    }
  }

  // Outputs.
  for (int k = tid; k < R; k += blockDim.x) {
    p.beta_idx[(size_t)read * R + k] = s_bi[k];
    p.energies[(size_t)read * R + k] = s_E[k];
  }
  if (lane == 0 && p.accepted) atomicAdd(p.accepted + read, acc);
  __syncthreads();
  if (p.best_sem == 0) {  // simulation.py:174-177, first minimum wins
    if (tid == 0) {
      int bp = 0;
      for (int k = 1; k < R; ++k)
        if (p.best_rep_e[(size_t)read * R + k] < p.best_rep_e[(size_t)read * R + bp]) bp = k;
      s_src = bp;
      p.best_energy[read] = p.best_rep_e[(size_t)read * R + bp];
      p.best_sweeps[read] = p.best_rep_sw[(size_t)read * R + bp];
    }
    __syncthreads();
    const int8_t* src = p.best_rep + ((size_t)read * R + s_src) * n;
    for (int i = tid; i < n; i += blockDim.x) p.best_state[(size_t)read * n + i] = src[i];
  } else {
    if (tid == 0) {
      p.best_energy[read] = s_gE;
      p.best_sweeps[read] = s_gsw;
    }
    if (!s_improved) {  // best_state is still the view of states[0] (simulation.py:204)
      const int8_t* src = p.spins + (size_t)read * R * n;
      for (int i = tid; i < n; i += blockDim.x) p.best_state[(size_t)read * n + i] = src[i];
    }
  }
}

// AIS: one wave per chain, no barriers (simulation.py:266-317).
template <typename T, int CH>
__global__ __launch_bounds__(256) void plain_ais_kernel(AISParams<T> p) {
  const int lane = threadIdx.x & 63;
  const int chain = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (chain >= p.chains) return;
  const int n = p.n;
  Chain<T, CH> ch;
  const size_t off = (size_t)chain * n;
  ch.load(p.fields + off, p.spins + off, n);
  double E = ch.refresh(p.Jt, p.h, n, p.ptype);
  int ones = ch.count_up(n);
  const double lratio = p.lp - p.l1p;
  double lw = 0.0, bold = 0.0;
  unsigned long long acc = 0;
  for (int t = 0; t < p.nt; ++t) {
    const T beta = p.temps[t];
    const double bd = (double)beta;
    if (p.ordering == 1) {
      const double refE = (p.ref == 1) ? -((double)(n - ones) * p.lp + (double)ones * p.l1p) : 0.0;
      lw += (-bd * E - (1.0 - bd) * refE) - (-bold * E - (1.0 - bold) * refE);
    }
    const size_t rb = ((size_t)chain * p.nt + t) * (size_t)n;
    int od = 0;
    const double d = ch.template sweep<true>(n, p.ptype, p.Jt, p.h, beta, (uint64_t)chain,
                                             (uint64_t)t, p.order, p.k0, p.k1,
                                             p.rep_lr ? p.rep_lr + rb : nullptr,
                                             p.rep_perm ? p.rep_perm + rb : nullptr, p.ref, lratio,
                                             acc, od);
    E += d;
    ones += od;
    if (p.ordering == 0) {  // simulation.py:309-314, after the transition
      const double refE = (p.ref == 1) ? -((double)(n - ones) * p.lp + (double)ones * p.l1p) : 0.0;
      lw += (-bd * E - (1.0 - bd) * refE) - (-bold * E - (1.0 - bold) * refE);
    }
    bold = bd;
  }
  ch.store(p.fields + off, p.spins + off, n, p.ptype);
  if (lane == 0) {
    p.energy[chain] = E;
    p.logw[chain] = lw;
  }
}

// E = s.(J s / 2 + h) in f64, one workgroup per state (ising.py:32-37).
// Thread i accumulates F_i = sum_j J[i][j] s_j from rows of J^T (coalesced
// across threads); the state is read through the scalar cache.
template <typename T>
__global__ __launch_bounds__(256) void energy_f64_kernel(const T* __restrict__ Jt,
                                                         const T* __restrict__ h,
                                                         const int8_t* __restrict__ spins,
                                                         double* out, int n) {
  __shared__ double red[4];
  const int8_t* st = spins + (size_t)blockIdx.x * n;
  double e = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    double f = 0.0;
    for (int j = 0; j < n; ++j) f += (double)Jt[(size_t)j * n + i] * (double)st[j];
    e += (double)st[i] * (0.5 * f + (double)h[i]);
  }
  e = wave_sum(e);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = e;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
    out[blockIdx.x] = t;
  }
}

int chunks_for(int n) {
  int ch = 1;
  while (64 * ch < n) ch *= 2;
  return ch;
}

}  // namespace

template <typename T>
hipError_t launch_plain_pt(const PTParams<T>& p, hipStream_t st) {
  if (p.n < 1 || p.n > kPlainMaxN || p.R < 1 || p.R > kMaxR || p.reads < 1) return hipErrorInvalidValue;
  const int ch = chunks_for(p.n);
  dim3 grid(p.reads);
#define PYSA_LAUNCH(CHV)                                                                   \
  case CHV:                                                                                \
    hipLaunchKernelGGL((plain_pt_kernel<T, CHV>), grid, dim3(64 * min(p.R, waves_for<T, CHV>())), \
                       0, st, p);                                                          \
    break;
  switch (ch) {
    PYSA_LAUNCH(1)
    PYSA_LAUNCH(2)
    PYSA_LAUNCH(4)
    PYSA_LAUNCH(8)
    PYSA_LAUNCH(16)
    PYSA_LAUNCH(32)
    PYSA_LAUNCH(64)
    default:
      return hipErrorInvalidValue;
  }
#undef PYSA_LAUNCH
  return hipGetLastError();
}

template <typename T>
hipError_t launch_plain_ais(const AISParams<T>& p, hipStream_t st) {
  if (p.n < 1 || p.n > kPlainMaxN || p.chains < 1) return hipErrorInvalidValue;
  const int ch = chunks_for(p.n);
  const int W = 4;
  dim3 grid((p.chains + W - 1) / W), block(64 * W);
#define PYSA_LAUNCH(CHV) \
  case CHV:              \
    hipLaunchKernelGGL((plain_ais_kernel<T, CHV>), grid, block, 0, st, p); break;
  switch (ch) {
    PYSA_LAUNCH(1)
    PYSA_LAUNCH(2)
    PYSA_LAUNCH(4)
    PYSA_LAUNCH(8)
    PYSA_LAUNCH(16)
    PYSA_LAUNCH(32)
    PYSA_LAUNCH(64)
    default:
      return hipErrorInvalidValue;
  }
#undef PYSA_LAUNCH
  return hipGetLastError();
}

template <typename T>
hipError_t launch_energy_f64(const T* Jt, const T* h, const int8_t* spins, double* out, int n,
                             long long count, hipStream_t st) {
  if (count <= 0) return hipSuccess;
  if (count > 0x7fffffffLL) return hipErrorInvalidValue;
  hipLaunchKernelGGL((energy_f64_kernel<T>), dim3((unsigned)count), dim3(256), 0, st, Jt, h,
                     spins, out, n);
  return hipGetLastError();
}

template hipError_t launch_plain_pt<float>(const PTParams<float>&, hipStream_t);
template hipError_t launch_plain_pt<double>(const PTParams<double>&, hipStream_t);
template hipError_t launch_plain_ais<float>(const AISParams<float>&, hipStream_t);
template hipError_t launch_plain_ais<double>(const AISParams<double>&, hipStream_t);
template hipError_t launch_energy_f64<float>(const float*, const float*, const int8_t*, double*, int,
                                             long long, hipStream_t);
template hipError_t launch_energy_f64<double>(const double*, const double*, const int8_t*, double*,
                                              int, long long, hipStream_t);
}  // namespace pysa
