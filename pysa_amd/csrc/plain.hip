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
// Layout.  One workgroup per read (PT) or per group of chains (AIS); each
// wave carries one chain at a time.  A chain's local fields F = J s live in VGPRs:
// element i sits in lane (i>>2)&63, register slot ((i>>8)<<2)|(i&3), so four
// consecutive slots of a lane hold four consecutive elements and a coupling
// row moves as 16-byte vectors.  Spins are one bit per slot.  A proposal reads
// F[pos] with a readlane, so dE is O(1); an accepted flip adds one coupling row
// (times dS in {+-1, +-2}) to every field.
//
// The sweep loop, best tracking, parallel tempering and log-omega run inside
// one persistent launch; the replicas of a read meet at a workgroup barrier
// once per sweep.
#include "engine.hpp"
#include "rng.hpp"

namespace pysa {
namespace {

__device__ __forceinline__ int rl(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }
__device__ __forceinline__ float rl(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
__device__ __forceinline__ uint32_t rl(uint32_t v, int lane) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, lane);
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

// Element <-> (lane, slot) map of the field layout.
__device__ __forceinline__ int elem_of(int lane, int g) {
  return ((g >> 2) << 8) + (lane << 2) + (g & 3);
}
__device__ __forceinline__ int lane_of(int i) { return (i >> 2) & 63; }
__device__ __forceinline__ int slot_of(int i) { return ((i >> 8) << 2) | (i & 3); }

__device__ __forceinline__ uint32_t shfl_bits(uint32_t w, int src) {
  return (uint32_t)__shfl((int)w, src, 64);
}
__device__ __forceinline__ uint64_t shfl_bits(uint64_t w, int src) {
  const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)w, src, 64);
  const uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(w >> 32), src, 64);
  return ((uint64_t)hi << 32) | lo;
}

template <typename T>
struct alignas(4 * sizeof(T)) T4 {
  T v[4];
};

// One chain: G field registers per lane (G a multiple of 4, n <= 64*G).
template <int G>
struct BitsFor {
  using type = uint64_t;
};
template <>
struct BitsFor<4> {
  using type = uint32_t;
};
template <>
struct BitsFor<8> {
  using type = uint32_t;
};
template <>
struct BitsFor<16> {
  using type = uint32_t;
};
template <>
struct BitsFor<32> {
  using type = uint32_t;
};

template <typename T, int G>
struct Chain {
  using B = typename BitsFor<G>::type;
  T f[G];
  B bits;  // bit g: the spin of element elem_of(lane, g) is up (+1 / 1)

  __device__ void load(const T* F, const int8_t* s, int n) {
    const int lane = threadIdx.x & 63;
    bits = 0;
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int i = elem_of(lane, g);
      f[g] = (i < n) ? F[i] : T(0);
      if (i < n && s[i] == 1) bits |= (B(1) << g);
    }
  }
  __device__ void store(T* F, int8_t* s, int n, int ptype) const {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int i = elem_of(lane, g);
      if (i < n) {
        F[i] = f[g];
        s[i] = ((bits >> g) & 1) ? int8_t(1) : (ptype == 0 ? int8_t(-1) : int8_t(0));
      }
    }
  }
  __device__ void store_spins(int8_t* s, int n, int ptype) const {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int i = elem_of(lane, g);
      if (i < n) s[i] = ((bits >> g) & 1) ? int8_t(1) : (ptype == 0 ? int8_t(-1) : int8_t(0));
    }
  }
  __device__ bool up_at(int pos) const { return (rl(bits, lane_of(pos)) >> slot_of(pos)) & 1; }

  // f[g] for a wave-uniform slot g, as a select tree on g's bits: register
  // indices stay compile-time constants, so the fields never leave VGPRs.
  __device__ __forceinline__ T pick(int g) const {
    T v[G];
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

  // F = J s accumulated in f64 and rounded once to T; returns
  // E = s.(F/2 + h) in f64 (the periodic recomputation that bounds drift).
  __device__ double refresh(const T* __restrict__ Jt, const T* h, int n, int ptype) {
    const int lane = threadIdx.x & 63;
    double e = 0.0;
#pragma unroll
    for (int g4 = 0; g4 < G / 4; ++g4) {
      double a[4] = {0.0, 0.0, 0.0, 0.0};
      const int i0 = (g4 << 8) + (lane << 2);
      for (int j = 0; j < n; ++j) {
        const bool upj = up_at(j);
        if (ptype == 1 && !upj) continue;
        const double sj = upj ? 1.0 : -1.0;
        const T* row = Jt + (size_t)j * n;
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if (i0 + q < n) a[q] += (double)row[i0 + q] * sj;
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int g = 4 * g4 + q;
        f[g] = (T)a[q];
        if (i0 + q < n) {
          const double si = ((bits >> g) & 1) ? 1.0 : (ptype == 0 ? -1.0 : 0.0);
          e += si * (0.5 * a[q] + (double)h[i0 + q]);
        }
      }
    }
    return wave_sum(e);
  }

  // The same from the nonzeros of a symmetric J (sparse form): row i's
  // entries in ascending column order, so the f64 sums equal the dense
  // ones term for term (a zero coupling adds exactly nothing).
  __device__ double refresh_ell(const int* __restrict__ cols, const T* __restrict__ vals, int D,
                                const T* h, int n, int ptype) {
    const int lane = threadIdx.x & 63;
    double e = 0.0;
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int i = elem_of(lane, g);
      double a = 0.0;
      for (int k = 0; k < D; ++k) {
        const int j = (i < n) ? cols[(size_t)i * D + k] : -1;
        // every lane takes part in the exchange (a disabled source lane reads as zero)
        const B w = shfl_bits(bits, lane_of(j < 0 ? 0 : j));
        if (j < 0) continue;
        const bool upj = (w >> slot_of(j)) & 1;
        if (ptype == 1 && !upj) continue;
        a += (double)vals[(size_t)i * D + k] * (upj ? 1.0 : -1.0);
      }
      f[g] = (T)a;
      if (i < n) {
        const double si = ((bits >> g) & 1) ? 1.0 : (ptype == 0 ? -1.0 : 0.0);
        e += si * (0.5 * a + (double)h[i]);
      }
    }
    return wave_sum(e);
  }

  // Add m to f[g] for a wave-uniform slot g (a scalar branch tree on g's
  // bits, so the register index stays a compile-time constant).
  template <int LO, int HI>
  __device__ __forceinline__ void add_at(int g, T m) {
    if constexpr (HI - LO == 1) {
      f[LO] += m;
    } else {
      constexpr int MID = (LO + HI) / 2;
      if (g < MID)
        add_at<LO, MID>(g, m);
      else
        add_at<MID, HI>(g, m);
    }
  }
  // Add dS times row p of a symmetric sparse J (its ELL row: d <= 64
  // entries, column -1 past the row's nonzeros): lane k loads entry k (one
  // load latency for the whole row), then each nonzero's owner lane updates
  // its register; the other lanes add zero, as the dense row would.
  __device__ __forceinline__ void add_ell(const int* __restrict__ cols,
                                          const T* __restrict__ vals, int d, T dS) {
    const int lane = threadIdx.x & 63;
    const int kl = lane < d ? lane : 0;
    const int cl = cols[kl];
    const T vl = vals[kl];
    for (int k = 0; k < d; ++k) {
      const int j = rl(cl, k);
      if (j < 0) break;
      add_at<0, G>(slot_of(j), lane == lane_of(j) ? rl(vl, k) * dS : T(0));
    }
  }

  __device__ int count_up(int n) const {
    const int lane = threadIdx.x & 63;
    int c = 0;
#pragma unroll
    for (int g = 0; g < G; ++g)
      if (elem_of(lane, g) < n) c += (int)((bits >> g) & 1);
    return wave_sum(c);
  }

  // Add dS times a staged row (LDS, 16-byte aligned, zero-padded to 64*G).
  __device__ __forceinline__ void add_row(const T* row, T dS) {
    const int lane = threadIdx.x & 63;
    const T4<T>* r4 = reinterpret_cast<const T4<T>*>(row);
#pragma unroll
    for (int g4 = 0; g4 < G / 4; ++g4) {
      const T4<T> x = r4[(g4 << 6) + lane];
#pragma unroll
      for (int q = 0; q < 4; ++q) f[4 * g4 + q] += x.v[q] * dS;
    }
  }
  // Add dS times a coupling row padded with zeros to ldp elements (a
  // multiple of 256, 16-byte aligned): a 256-element chunk is loaded whole
  // or, beyond ldp, not at all (a wave-uniform test, no per-lane masks).
  __device__ __forceinline__ void add_row_padded(const T* __restrict__ row, int ldp, T dS) {
    const int lane = threadIdx.x & 63;
    const T4<T>* r4 = reinterpret_cast<const T4<T>*>(row);
    T4<T> x[G / 4];
    if (ldp >= 64 * G) {
#pragma unroll
      for (int g4 = 0; g4 < G / 4; ++g4) x[g4] = r4[(g4 << 6) + lane];
    } else {
#pragma unroll
      for (int g4 = 0; g4 < G / 4; ++g4)
        x[g4] = ((g4 << 8) < ldp) ? r4[(g4 << 6) + lane] : T4<T>{{T(0), T(0), T(0), T(0)}};
    }
#pragma unroll
    for (int g4 = 0; g4 < G / 4; ++g4)
#pragma unroll
      for (int q = 0; q < 4; ++q) f[4 * g4 + q] += x[g4].v[q] * dS;
  }
};

// Decision for one proposal (ising.py:48-52 / qubo.py:48-52; with AIS the
// reference-distribution term of update_spin_ref, ising.py:75-83).
template <typename T, bool AIS>
__device__ __forceinline__ bool decide(int ptype, bool up, T fh, double lr, double bd, int ref,
                                       double lratio, double& dn) {
  dn = (ptype == 0) ? 2.0 * (up ? 1.0 : -1.0) * (double)fh
                    : (2.0 * (up ? 1.0 : 0.0) - 1.0) * (double)fh;
  if (AIS) {
    const double refd = (ref == 1) ? (up ? lratio : -lratio) : 0.0;
    const double full = bd * dn + (1.0 - bd) * refd;
    return full >= 0 || lr < full;
  }
  return dn >= 0 || lr < bd * dn;
}

// Log-uniform of (chain, position) from replay or Philox.
__device__ __forceinline__ double log_uniform(const double* rep, uint64_t chain, uint64_t sw,
                                              int pos, uint32_t k0, uint32_t k1) {
  if (rep) return rep[pos];
  return log_u32(pick_word(sweep_block(chain, sw, (uint32_t)pos, k0, k1), (uint32_t)pos & 3u));
}

constexpr int kMaxR = 1024;

// Waves per workgroup for G field registers of type T, keeping fields and
// working registers inside the VGPR budget.
template <typename T, int G>
constexpr int kernel_waves() {
  return G * (int)(sizeof(T) / 4) <= 16 ? 16 : G * (int)(sizeof(T) / 4) <= 32 ? 8 : 4;
}


struct SweepCtx {
  int n, ptype, order, ref;
  uint32_t k0, k1;
  uint64_t sw;
  double lratio;
  // sparse form: ELL rows of J (ell_d > 0), else the dense padded rows
  const int* ell_cols = nullptr;
  const void* ell_vals = nullptr;
  int ell_d = 0;
};

// One sweep of one chain (simulation.py:41-76): visit order from the
// sequential order, a Feistel permutation (per chain, or shared by all chains
// of the sweep) or a replayed order; an accepted flip loads its coupling row
// from global memory (L2 / Infinity Cache resident) as 16-byte vectors.
// SP: the sparse form (an accepted flip updates the nonzeros of its ELL row),
// a separate instantiation so that the dense kernels keep their registers.
template <typename T, int G, bool AIS, bool SP = false>
__device__ double sweep_direct(Chain<T, G>& ch, bool active, const T* hs, const SweepCtx& c,
                               const T* __restrict__ Jp, int ldp, T beta, uint64_t cid,
                               const double* rep_lr, const int* rep_perm,
                               unsigned long long& acc, int& od) {
  const int n = c.n, lane = threadIdx.x & 63;
  od = 0;
  if (!active) return 0.0;
  // Visit slots: n, or 32 per block in the block-random order (padding
  // slots of the last block visit nothing).
  const bool brand = c.order == kBlockRandom && rep_perm == nullptr;
  const int nv = brand ? kOrderBlock * ((n + kOrderBlock - 1) / kOrderBlock) : n;
  FeistelKey fk;
  if (brand)
    fk = block_order_key((uint32_t)((n + kOrderBlock - 1) / kOrderBlock), c.sw, c.k0, c.k1);
  else if (c.order != kSequential && rep_perm == nullptr)
    fk = feistel_make((uint32_t)n,
                      perm_keys(c.order == kRandomShared ? 0ull : cid, c.sw,
                                c.order == kRandomShared ? kStreamShared : kStreamPerm, c.k0, c.k1));
  double sum = 0.0;
  for (int w = 0; w < nv; w += 64) {
    const int vi = w + lane;
    int pos_l = -1;
    double lr_l = 0.0;
    if (vi < nv) {
      if (rep_perm)
        pos_l = rep_perm[vi];
      else if (brand)
        pos_l = block_random_pos(fk, c.sw, (uint32_t)vi, n, c.k0, c.k1);
      else
        pos_l = (c.order == kSequential) ? vi : (int)feistel_perm(fk, (uint32_t)vi);
      if (pos_l >= 0) lr_l = log_uniform(rep_lr, cid, c.sw, pos_l, c.k0, c.k1);
    }
    const T h_l = (pos_l >= 0) ? hs[pos_l] : T(0);
    const int cnt = min(64, nv - w);
    for (int t = 0; t < cnt; ++t) {
      const int pos = rl(pos_l, t);
      if (pos < 0) continue;  // padding slot (wave-uniform)
      const double lr = rl(lr_l, t);
      const int pl = lane_of(pos), ps = slot_of(pos);
      const bool up = (rl(ch.bits, pl) >> ps) & 1;
      const T fh = rl(ch.pick(ps), pl) + rl(h_l, t);
      double dn;
      if (decide<T, AIS>(c.ptype, up, fh, lr, (double)beta, c.ref, c.lratio, dn)) {
        sum += -dn;
        ++acc;
        od += up ? -1 : 1;
        if (lane == pl) ch.bits ^= (typename Chain<T, G>::B(1) << ps);
        const T dS = (c.ptype == 0) ? (up ? T(-2) : T(2)) : (up ? T(-1) : T(1));
        if constexpr (SP)
          ch.add_ell(c.ell_cols + (size_t)pos * c.ell_d,
                     static_cast<const T*>(c.ell_vals) + (size_t)pos * c.ell_d, c.ell_d, dS);
        else
          ch.add_row_padded(Jp + (size_t)pos * ldp, ldp, dS);
      }
    }
  }
  return sum;
}

// Waves per workgroup of the PT kernel: the sparse form's update and refresh
// need more registers per lane, so it runs half as many waves.
template <typename T, int G, bool SP>
constexpr int pt_waves() {
  return SP ? (kernel_waves<T, G>() > 4 ? kernel_waves<T, G>() / 2 : 4) : kernel_waves<T, G>();
}

template <typename T, int G, bool SP>
__global__ __launch_bounds__((64 * pt_waves<T, G, SP>())) void plain_pt_kernel(PTParams<T> p) {
  const int read = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, W = blockDim.x >> 6;
  const int R = p.R, n = p.n;
  __shared__ T s_h[kPlainMaxN];
  __shared__ T s_bs[kMaxR];
  __shared__ T s_beta[kMaxR];
  __shared__ T s_E[kMaxR];
  __shared__ long long s_bi[kMaxR];
  __shared__ double s_ptlu[kMaxR];
  __shared__ double s_eoff[kMaxR];  // caller's initial energy minus its f64 recomputation
  __shared__ int s_src;
  __shared__ T s_gE;
  __shared__ int s_gsw, s_improved;

  for (int i = tid; i < n; i += blockDim.x) s_h[i] = p.h[i];
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
  __syncthreads();
  SweepCtx c;
  c.n = n;
  c.ptype = p.ptype;
  c.order = p.order;
  c.ref = 0;
  c.k0 = p.k0;
  c.k1 = p.k1;
  c.lratio = 0.0;
  c.ell_cols = p.ell_cols;
  c.ell_vals = p.ell_vals;
  c.ell_d = p.ell_d;
  unsigned long long acc = 0;
  for (int s = 0; s < p.S; ++s) {
    c.sw = (uint64_t)s + p.sweep_offset;
    // Local fields F = J s recomputed in f64 (rounded once) at the start and
    // every `refresh` sweeps, which bounds the drift of float couplings.  The
    // reference recomputes every field from its coupling row
    // (ising.py:48), so its energy changes carry no drift; here the tracked
    // energy is reset with the fields to the f64 energy plus the offset the
    // caller's initial energy had from it (zero unless the caller passed
    // other energies), so the drift of incrementally updated fields does not
    // accumulate.  Integer couplings are exact either way: the reset then
    // leaves the value unchanged.
    if (s == 0 || (p.refresh > 0 && s % p.refresh == 0)) {
      for (int r = wave; r < R; r += W) {
        Chain<T, G> ch;
        const size_t off = ((size_t)read * R + r) * n;
        ch.load(p.fields + off, p.spins + off, n);
        double ex;
        if constexpr (SP)
          ex = ch.refresh_ell(p.ell_cols, p.ell_vals, p.ell_d, s_h, n, p.ptype);
        else
          ex = ch.refresh(p.Jt, s_h, n, p.ptype);
        ch.store(p.fields + off, p.spins + off, n, p.ptype);
        if (lane == 0) {
          if (s == 0)
            s_eoff[r] = (double)s_E[r] - ex;
          else
            s_E[r] = (T)(ex + s_eoff[r]);
        }
      }
    }
    for (int k = tid; k < R; k += blockDim.x) s_bs[s_bi[k]] = s_beta[k];  // simulation.py:151-152
    // PT log-uniforms, one per pair, drawn in parallel ahead of the serial walk.
    if (p.use_pt)
      for (int k = tid; k < R - 1; k += blockDim.x) {
        double lu;
        if (p.rep_pt)
          lu = p.rep_pt[((size_t)read * p.S + s) * (R - 1) + k];
        else
          lu = log_u32(pick_word(pt_block((uint64_t)read, c.sw, (uint32_t)k, p.k0, p.k1),
                                 (uint32_t)k & 3u));
        s_ptlu[k] = lu;
      }
    __syncthreads();
    for (int r = wave; r < R; r += W) {  // replicas of this read (simulation.py:154-164)
      Chain<T, G> ch;
      const size_t off = ((size_t)read * R + r) * n;
      ch.load(p.fields + off, p.spins + off, n);
      const size_t rb = (((size_t)read * p.S + s) * R + r) * (size_t)n;
      int od;
      const double d = sweep_direct<T, G, false, SP>(ch, true, s_h, c, p.Jp, p.ldp, s_bs[r],
                                                 (uint64_t)read * R + r,
                                                 p.rep_lr ? p.rep_lr + rb : nullptr,
                                                 p.rep_perm ? p.rep_perm + rb : nullptr, acc, od);
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

// AIS: each wave carries one chain through the whole beta schedule
// (simulation.py:266-317); chains are independent, so there are no barriers
// after h is staged.
template <typename T, int G>
__global__ __launch_bounds__((64 * kernel_waves<T, G>())) void plain_ais_kernel(AISParams<T> p) {
  __shared__ T s_h[kPlainMaxN];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, W = blockDim.x >> 6;
  const int n = p.n;
  for (int i = tid; i < n; i += blockDim.x) s_h[i] = p.h[i];
  __syncthreads();
  const int chain = blockIdx.x * W + wave;
  if (chain >= p.chains) return;
  Chain<T, G> ch;
  ch.load(p.fields + (size_t)chain * n, p.spins + (size_t)chain * n, n);
  double E = ch.refresh(p.Jt, s_h, n, p.ptype);
  int ones = ch.count_up(n);
  SweepCtx c;
  c.n = n;
  c.ptype = p.ptype;
  c.order = p.order;
  c.ref = p.ref;
  c.k0 = p.k0;
  c.k1 = p.k1;
  c.lratio = p.lp - p.l1p;
  double bold = 0.0, lw = 0.0;
  unsigned long long acc = 0;
  for (int t = 0; t < p.nt; ++t) {
    c.sw = (uint64_t)t;
    const T beta = p.temps[t];
    const double bd = (double)beta;
    // The reference recomputes the energy after every step
    // (simulation.py:309); here fields and energy are recomputed in f64
    // every `refresh` steps and updated by the accepted changes in between,
    // which bounds the drift of float couplings.
    if (t > 0 && p.refresh > 0 && t % p.refresh == 0) E = ch.refresh(p.Jt, s_h, n, p.ptype);
    if (p.ordering == 1) {  // Neal's ordering: weight before the transition
      const double refE = (p.ref == 1) ? -((double)(n - ones) * p.lp + (double)ones * p.l1p) : 0.0;
      lw += (-bd * E - (1.0 - bd) * refE) - (-bold * E - (1.0 - bold) * refE);
    }
    const size_t rb = ((size_t)chain * p.nt + t) * (size_t)n;
    int od;
    const double d = sweep_direct<T, G, true>(ch, true, s_h, c, p.Jp, p.ldp, beta, (uint64_t)chain,
                                              p.rep_lr ? p.rep_lr + rb : nullptr,
                                              p.rep_perm ? p.rep_perm + rb : nullptr, acc, od);
    E += d;
    ones += od;
    if (p.ordering == 0) {  // simulation.py:309-314, after the transition
      const double refE = (p.ref == 1) ? -((double)(n - ones) * p.lp + (double)ones * p.l1p) : 0.0;
      lw += (-bd * E - (1.0 - bd) * refE) - (-bold * E - (1.0 - bold) * refE);
    }
    bold = bd;
  }
  ch.store(p.fields + (size_t)chain * n, p.spins + (size_t)chain * n, n, p.ptype);
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

// Field registers per lane for n spins: a power of two from 4 to 64 with
// 64 * G >= n.
int regs_for(int n) {
  int g = 4;
  while (64 * g < n) g *= 2;
  return g;
}

template <typename T, int G>
hipError_t launch_pt_shape(const PTParams<T>& p, hipStream_t st) {
  if (p.ell_d > 0) {
    const int W = min(p.R, pt_waves<T, G, true>());
    hipLaunchKernelGGL((plain_pt_kernel<T, G, true>), dim3(p.reads), dim3(64 * W), 0, st, p);
  } else {
    const int W = min(p.R, pt_waves<T, G, false>());
    hipLaunchKernelGGL((plain_pt_kernel<T, G, false>), dim3(p.reads), dim3(64 * W), 0, st, p);
  }
  return hipGetLastError();
}

template <typename T, int G>
hipError_t launch_ais_shape(const AISParams<T>& p, hipStream_t st) {
  const int W = kernel_waves<T, G>();
  hipLaunchKernelGGL((plain_ais_kernel<T, G>), dim3((p.chains + W - 1) / W), dim3(64 * W), 0, st, p);
  return hipGetLastError();
}

}  // namespace

template <typename T>
hipError_t launch_plain_pt(const PTParams<T>& p, hipStream_t st) {
  if (p.n < 1 || p.n > kPlainMaxN || p.R < 1 || p.R > kMaxR || p.reads < 1)
    return hipErrorInvalidValue;
  switch (regs_for(p.n)) {
    case 4: return launch_pt_shape<T, 4>(p, st);
    case 8: return launch_pt_shape<T, 8>(p, st);
    case 16: return launch_pt_shape<T, 16>(p, st);
    case 32: return launch_pt_shape<T, 32>(p, st);
    case 64: return launch_pt_shape<T, 64>(p, st);
    default: return hipErrorInvalidValue;
  }
}

template <typename T>
hipError_t launch_plain_ais(const AISParams<T>& p, hipStream_t st) {
  if (p.n < 1 || p.n > kPlainMaxN || p.chains < 1) return hipErrorInvalidValue;
  switch (regs_for(p.n)) {
    case 4: return launch_ais_shape<T, 4>(p, st);
    case 8: return launch_ais_shape<T, 8>(p, st);
    case 16: return launch_ais_shape<T, 16>(p, st);
    case 32: return launch_ais_shape<T, 32>(p, st);
    case 64: return launch_ais_shape<T, 64>(p, st);
    default: return hipErrorInvalidValue;
  }
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
