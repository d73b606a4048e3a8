// ORACLE -- test infrastructure only.
//
// A CPU restatement of PySA's dense Metropolis / parallel-tempering path and
// of its annealed-importance-sampling sibling.  It is the checker the HIP
// kernels are compared against (and, built with -O3 -fopenmp, the CPU
// baseline that bench.py times).  Nothing in the product links, imports or
// calls this file; only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg do.
//
// It follows the reference line by line, with the O(N) coupling-row dot per
// proposal that the reference uses:
//   energy            pysa/ising.py:32-37, pysa/qubo.py:32-37
//   update_spin       pysa/ising.py:40-63, pysa/qubo.py:40-63
//   sweeps            pysa/simulation.py:41-76 (log_r indexed by position)
//   simulation        pysa/simulation.py:123-181 (per-replica best) and
//                     pysa/simulation.py:184-236 (global best, aliasing quirk)
//   pt                pysa/utils.py:26-58 (pairs from the hot end, lazy draw)
//   get_log_omega     pysa/ais.py:79-112
//   ais_sample        pysa/simulation.py:293-317, update_spin_ref
//                     pysa/ising.py:66-92, pysa/qubo.py:66-92,
//                     ref probabilities pysa/ais.py:137-226
//
// Randomness comes either from host-supplied arrays (replay: log-uniforms
// indexed [read][sweep][replica][position], permutations indexed
// [read][sweep][replica][visit], PT log-uniforms indexed [read][sweep][pair])
// or from this file's own Philox4x32-10 + deterministic-log + Feistel
// implementation, written independently of the product's and pinned by the
// Random123 known-answer vectors in tests/.
//
// Parity status per function (what pins it) is listed in the module docstring
// of oracle/__init__.py.
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ---------------------------------------------------------------- RNG -----
struct W4 {
  uint32_t v[4];
};

W4 philox10(W4 c, uint32_t k0, uint32_t k1) {
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c.v[0];
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c.v[2];
    W4 n;
    n.v[0] = (uint32_t)(p1 >> 32) ^ c.v[1] ^ k0;
    n.v[1] = (uint32_t)p1;
    n.v[2] = (uint32_t)(p0 >> 32) ^ c.v[3] ^ k1;
    n.v[3] = (uint32_t)p0;
    c = n;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

double u01(uint32_t x) { return ((double)x + 0.5) * 2.3283064365386962890625e-10; }

// Natural log from +,-,*,/,fma in a fixed order (see the product's rng
// header for the method; this is a separate implementation of the same
// specification).
double dlog(double x) {
  uint64_t b;
  memcpy(&b, &x, 8);
  int e = (int)((b >> 52) & 0x7ff) - 1023;
  uint64_t mb = (b & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull;
  double m;
  memcpy(&m, &mb, 8);
  if (m > 1.4142135623730951) {
    m *= 0.5;
    e += 1;
  }
  const double s = (m - 1.0) / (m + 1.0);
  const double z = s * s;
  static const double c[12] = {1.0 / 25.0, 1.0 / 23.0, 1.0 / 21.0, 1.0 / 19.0,
                               1.0 / 17.0, 1.0 / 15.0, 1.0 / 13.0, 1.0 / 11.0,
                               1.0 / 9.0,  1.0 / 7.0,  1.0 / 5.0,  1.0 / 3.0};
  double p = c[0];
  for (int i = 1; i < 12; ++i) p = fma(p, z, c[i]);
  const double s2 = s + s;
  const double lnm = fma(s2 * z, p, s2);
  const double de = (double)e;
  return fma(de, 0.6931471787393093, fma(de, 1.8206359985041462e-09, lnm));
}

uint32_t mix(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85EBCA6Bu;
  h ^= h >> 13;
  h *= 0xC2B2AE35u;
  h ^= h >> 16;
  return h;
}

struct Perm {
  uint32_t k[4], hb, n;
  Perm(uint32_t n_, W4 keys) : n(n_) {
    for (int i = 0; i < 4; ++i) k[i] = keys.v[i];
    uint32_t bits = 0;
    while (bits < 32 && (1u << bits) < n) ++bits;
    hb = (bits + 1) / 2;
  }
  uint32_t once(uint32_t x) const {
    const uint32_t mask = (hb >= 32) ? 0xFFFFFFFFu : ((1u << hb) - 1u);
    uint32_t L = x >> hb, R = x & mask;
    for (int i = 0; i < 4; ++i) {
      uint32_t t = R;
      R = L ^ (mix(R ^ k[i]) & mask);
      L = t;
    }
    return (L << hb) | R;
  }
  uint32_t at(uint32_t i) const {
    if (n <= 1) return 0;
    uint32_t y = once(i);
    while (y >= n) y = once(y);
    return y;
  }
};

// chain bits 24-39 and sweep bits 32-47 share the last counter word
uint32_t chain_sweep_hi(uint64_t chain, uint64_t sweep) {
  return (uint32_t)((chain >> 24) & 0xFFFFu) | ((uint32_t)((sweep >> 32) & 0xFFFFu) << 16);
}
W4 ctr_sweep(uint64_t chain, uint64_t sweep, uint32_t pos) {
  W4 c = {{pos >> 2, (uint32_t)sweep, ((uint32_t)chain & 0xFFFFFFu), chain_sweep_hi(chain, sweep)}};
  return c;
}
W4 ctr_pt(uint64_t read, uint64_t sweep, uint32_t pair) {
  W4 c = {{pair >> 2, (uint32_t)sweep, (1u << 24) | (uint32_t)((sweep >> 32) & 0xFFFFFFu),
           (uint32_t)read}};
  return c;
}
W4 ctr_perm(uint64_t chain, uint64_t sweep, uint32_t purpose) {
  W4 c = {{0u, (uint32_t)sweep, ((uint32_t)chain & 0xFFFFFFu) | (purpose << 24),
           chain_sweep_hi(chain, sweep)}};
  return c;
}

// Block-random visit order (order 3), shared by every chain of a run: the
// 32-spin blocks in the order of a Feistel permutation of [0, nblk) keyed by
// the sweep (stream purpose 4), and inside block b the slots k = 0..31 visit
// position 32 b + (k ^ m_b), m_b the low 5 bits of word b % 4 of the Philox
// block (b / 4, sweep, 5 << 24, 0); positions >= n are skipped.
struct BlockOrder {
  int n, nblk;
  uint64_t sweep;
  uint32_t k0, k1;
  Perm blocks;
  BlockOrder(int n_, uint64_t sw, uint32_t k0_, uint32_t k1_)
      : n(n_), nblk((n_ + 31) / 32), sweep(sw), k0(k0_), k1(k1_),
        blocks((uint32_t)((n_ + 31) / 32), philox10(ctr_perm(0, sw, 4u), k0_, k1_)) {}
  // Every position of the sweep in visit order.
  std::vector<int> positions() const {
    std::vector<int> out;
    out.reserve(n);
    for (int t = 0; t < nblk; ++t) {
      const uint32_t b = blocks.at((uint32_t)t);
      W4 c = {{b >> 2, (uint32_t)sweep, 5u << 24, (uint32_t)((sweep >> 32) & 0xFFFFu)}};
      const uint32_t m = philox10(c, k0, k1).v[b & 3] & 31u;
      for (uint32_t k = 0; k < 32; ++k) {
        const int pos = (int)(32 * b + (k ^ m));
        if (pos < n) out.push_back(pos);
      }
    }
    return out;
  }
};

// ------------------------------------------------------------ problem -----
// Coupling-row dot in the problem's float type, fixed order: 32 interleaved
// partial sums (four independent 8-wide vector accumulators) combined
// pairwise (pysa/ising.py:48 `couplings[pos].dot(state)`).
template <typename T>
T row_dot(const T* row, const T* s, int n) {
  T acc[32];
  for (int q = 0; q < 32; ++q) acc[q] = 0;
  int j = 0;
  for (; j + 32 <= n; j += 32)
    for (int q = 0; q < 32; ++q) acc[q] += row[j + q] * s[j + q];
  for (int q = 0; j < n; ++j, ++q) acc[q] += row[j] * s[j];
  for (int w = 16; w >= 1; w >>= 1)
    for (int q = 0; q < w; ++q) acc[q] = acc[q] + acc[q + w];
  return acc[0];
}

struct RunCfg {
  int n, ptype, reads, R, S, order, use_pt, logw, best_sem;
  uint64_t seed, sweep_offset;
  const double* rep_lr;
  const int32_t* rep_perm;
  const double* rep_pt;
  int nthreads;
  // Batched instances: couplings scale * signs (int8, leading dimension ld)
  // with the field taken as the f32 value scale * (signs_i . s) + h_i, and
  // read indices offset by read_base in the random streams.
  const int8_t* signs = nullptr;
  int ld = 0;
  float scale = 1.f;
  uint64_t read_base = 0;
};

// Negated energy change of flipping `pos` (ising.py:48-49, qubo.py:48-49):
// f64 product of the f64 spin factor and the float-type sum row.s + h.
template <typename T>
inline double neg_delta(int ptype, T spin, T field_plus_h) {
  if (ptype == 0) return 2.0 * (double)spin * (double)field_plus_h;
  return (2.0 * (double)spin - 1.0) * (double)field_plus_h;
}

template <typename T>
inline T flipped(int ptype, T spin) {
  if (ptype == 0) return -spin;
  return spin != (T)0 ? (T)0 : (T)1;
}

// One sweep of one replica (simulation.py:60-76 / :41-57).  Returns the f64
// sum of the energy changes; counts acceptances.
template <typename T>
double sweep_one(const RunCfg& c, const T* J, const T* h, T* s, T beta, int read, int k, int sweep,
                 uint64_t* acc) {
  const int n = c.n;
  const uint32_t k0 = (uint32_t)c.seed, k1 = (uint32_t)(c.seed >> 32);
  const uint64_t chain = ((uint64_t)read + c.read_base) * (uint64_t)c.R + (uint64_t)k;
  const uint64_t sw = (uint64_t)sweep + c.sweep_offset;
  const size_t base = (((size_t)read * c.S + sweep) * c.R + k) * (size_t)n;
  Perm perm(n, philox10(ctr_perm(c.order == 2 ? 0 : chain, sw, c.order == 2 ? 3u : 2u), k0, k1));
  std::vector<int> border;
  if (c.order == 3 && !c.rep_perm) border = BlockOrder(n, sw, k0, k1).positions();
  double de = 0.0;
  W4 blk = {{0, 0, 0, 0}};
  uint32_t blk_id = 0xFFFFFFFFu;
  for (int i = 0; i < n; ++i) {
    int pos;
    if (c.rep_perm)
      pos = c.rep_perm[base + i];
    else if (c.order == 0)
      pos = i;
    else if (c.order == 3)
      pos = border[i];
    else
      pos = (int)perm.at((uint32_t)i);
    double lr;
    if (c.rep_lr) {
      lr = c.rep_lr[base + pos];
    } else {
      if ((uint32_t)pos >> 2 != blk_id) {
        blk_id = (uint32_t)pos >> 2;
        blk = philox10(ctr_sweep(chain, sw, (uint32_t)pos), k0, k1);
      }
      lr = dlog(u01(blk.v[pos & 3]));
    }
    T fh;
    if (c.signs) {
      long long dot = 0;  // exact integer count signs_pos . s
      const int8_t* row = c.signs + (size_t)pos * c.ld;
      for (int j = 0; j < n; ++j) dot += (long long)row[j] * (long long)s[j];
      const float t1 = (float)dot * c.scale;
      fh = (T)(t1 + (float)h[pos]);
    } else {
      fh = (T)(row_dot<T>(J + (size_t)pos * n, s, n) + h[pos]);
    }
    const double dn = neg_delta<T>(c.ptype, s[pos], fh);
    if (dn >= 0 || lr < (double)beta * dn) {
      s[pos] = flipped<T>(c.ptype, s[pos]);
      de += -dn;
      ++*acc;
    }
  }
  return de;
}

template <typename T>
int metropolis(const RunCfg& c, const T* J, const T* h, const T* betas, const int32_t* bsort,
               int8_t* states, T* energies, int64_t* beta_idx_out, int8_t* best_state,
               T* best_energy, int32_t* best_sweeps, double* log_omegas, uint64_t* accepted) {
  const int n = c.n, R = c.R, S = c.S;
  const uint32_t k0 = (uint32_t)c.seed, k1 = (uint32_t)(c.seed >> 32);
#ifdef _OPENMP
  if (c.nthreads > 0) omp_set_num_threads(c.nthreads);
#endif
  for (int read = 0; read < c.reads; ++read) {  // sa.py:425-427, serial reads
    std::vector<T> st((size_t)R * n), bst;
    std::vector<T> E(R), bE(R), bs(R);
    std::vector<int32_t> bsw(R, 0);
    std::vector<int64_t> bi(R);
    std::vector<uint64_t> acc(R, 0);
    for (size_t q = 0; q < (size_t)R * n; ++q) st[q] = (T)states[(size_t)read * R * n + q];
    for (int k = 0; k < R; ++k) {
      E[k] = energies[(size_t)read * R + k];
      bi[k] = k;  // sa.py:325
    }
    // simulation.py:143-145 (parallel) / :204-206 (sequential)
    bE = E;
    if (c.best_sem == 0) bst = st;
    T gE = E[0];
    int gsw = 0;
    bool aliased = true;  // best_state is a view of states[0] until improved
    std::vector<T> gst;
    for (int s = 0; s < S; ++s) {
      for (int k = 0; k < R; ++k) bs[bi[k]] = betas[k];  // simulation.py:151-152
#pragma omp parallel for schedule(static) if (c.nthreads != 1)
      for (int k = 0; k < R; ++k) {  // simulation.py:154 prange
        const double d = sweep_one<T>(c, J, h, &st[(size_t)k * n], bs[k], read, k, s, &acc[k]);
        E[k] = (T)((double)E[k] + d);  // simulation.py:157-158
        if (c.best_sem == 0 && E[k] < bE[k]) {  // simulation.py:161-164
          bE[k] = E[k];
          memcpy(&bst[(size_t)k * n], &st[(size_t)k * n], sizeof(T) * n);
          bsw[k] = s;
        }
      }
      if (c.best_sem == 1) {  // simulation.py:222-225, in replica order
        for (int k = 0; k < R; ++k)
          if (E[k] < gE) {
            gE = E[k];
            gst.assign(st.begin() + (size_t)k * n, st.begin() + (size_t)(k + 1) * n);
            gsw = s;
            aliased = false;
          }
      }
      if (c.use_pt) {  // utils.py:44-58
        for (int k = 0; k < R - 1; ++k) {
          const int a = R - k - 1, b = R - k - 2;
          const T dd = (T)(E[bi[a]] - E[bi[b]]) * (T)(betas[a] - betas[b]);
          bool sw = dd >= 0;
          if (!sw) {
            double lu;
            if (c.rep_pt)
              lu = c.rep_pt[((size_t)read * S + s) * (R - 1) + k];
            else {
              W4 w = philox10(ctr_pt((uint64_t)read + c.read_base, (uint64_t)s + c.sweep_offset,
                                     (uint32_t)k), k0, k1);
              lu = dlog(u01(w.v[k & 3]));
            }
            sw = lu < (double)dd;
          }
          if (sw) {
            int64_t t = bi[a];
            bi[a] = bi[b];
            bi[b] = t;
          }
        }
      }
      if (c.logw && log_omegas) {  // ais.py:85-112 with the ladder sorted by beta
        double lw = 0.0, bold = 0.0;
        for (int r = 1; r < R; ++r) {
          const T be = betas[bsort[r]];
          const T en = E[bi[bsort[r]]];
          const T a = -be * en;
          const double bterm = -bold * (double)en;
          lw += (double)a - bterm;
          bold = (double)be;
        }
        log_omegas[(size_t)read * S + s] = lw;
      }
    }
    // outputs
    for (size_t q = 0; q < (size_t)R * n; ++q) states[(size_t)read * R * n + q] = (int8_t)st[q];
    for (int k = 0; k < R; ++k) {
      energies[(size_t)read * R + k] = E[k];
      beta_idx_out[(size_t)read * R + k] = bi[k];
    }
    if (c.best_sem == 0) {  // simulation.py:174-177, first minimum wins
      int bp = 0;
      for (int k = 1; k < R; ++k)
        if (bE[k] < bE[bp]) bp = k;
      for (int q = 0; q < n; ++q) best_state[(size_t)read * n + q] = (int8_t)bst[(size_t)bp * n + q];
      best_energy[read] = bE[bp];
      best_sweeps[read] = bsw[bp];
    } else {
      const T* src = aliased ? &st[0] : &gst[0];
      for (int q = 0; q < n; ++q) best_state[(size_t)read * n + q] = (int8_t)src[q];
      best_energy[read] = gE;
      best_sweeps[read] = gsw;
    }
    if (accepted) {
      uint64_t a = 0;
      for (int k = 0; k < R; ++k) a += acc[k];
      accepted[read] = a;
    }
  }
  return 0;
}

// AIS (sa.py:506-553, simulation.py:293-317).  ordering 0 = the reference's
// (weight taken after the sweep at beta_t), 1 = Neal's (before the sweep).
template <typename T>
int ais(int n, int ptype, const T* J, const T* h, const T* temps, int nt, int chains, int ref,
        double p, int order, int ordering, uint64_t seed, const double* rep_lr,
        const int32_t* rep_perm, int8_t* states, double* logw, int nthreads) {
  const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
  const double lp = (ref == 1) ? log(p) : 0.0, l1p = (ref == 1) ? log(1.0 - p) : 0.0;
#pragma omp parallel for schedule(dynamic, 1) if (nthreads != 1)
  for (int ch = 0; ch < chains; ++ch) {
    std::vector<T> s(n);
    for (int q = 0; q < n; ++q) s[q] = (T)states[(size_t)ch * n + q];
    // energy in f64 (ising.py:32-37) and the count of "1" bits (ais.py:205-216)
    auto energy = [&]() {
      double e = 0.0;
      for (int i = 0; i < n; ++i) {
        double r = 0.0;
        for (int j = 0; j < n; ++j) r += (double)J[(size_t)i * n + j] * (double)s[j];
        e += (double)s[i] * (0.5 * r + (double)h[i]);
      }
      return e;
    };
    double E = energy();
    int ones = 0;
    for (int q = 0; q < n; ++q) ones += (s[q] == (T)1);
    double lw = 0.0, bold = 0.0;
    for (int t = 0; t < nt; ++t) {
      const T beta = temps[t];
      const double refE = (ref == 1) ? -((double)(n - ones) * lp + (double)ones * l1p) : 0.0;
      if (ordering == 1) {
        lw += (-(double)beta * E - (1.0 - (double)beta) * refE) -
              (-bold * E - (1.0 - bold) * refE);
      }
      // sweep_ref at beta (simulation.py:79-120, ising.py:66-92)
      const uint64_t chain = (uint64_t)ch;
      const size_t base = ((size_t)ch * nt + t) * (size_t)n;
      Perm perm(n, philox10(ctr_perm(order == 2 ? 0 : chain, (uint64_t)t, order == 2 ? 3u : 2u), k0, k1));
      std::vector<int> border;
      if (order == 3 && !rep_perm) border = BlockOrder(n, (uint64_t)t, k0, k1).positions();
      W4 blk = {{0, 0, 0, 0}};
      uint32_t blk_id = 0xFFFFFFFFu;
      for (int i = 0; i < n; ++i) {
        int pos = rep_perm ? rep_perm[base + i]
                           : (order == 0 ? i : order == 3 ? border[i] : (int)perm.at((uint32_t)i));
        double lr;
        if (rep_lr)
          lr = rep_lr[base + pos];
        else {
          if ((uint32_t)pos >> 2 != blk_id) {
            blk_id = (uint32_t)pos >> 2;
            blk = philox10(ctr_sweep(chain, (uint64_t)t, (uint32_t)pos), k0, k1);
          }
          lr = dlog(u01(blk.v[pos & 3]));
        }
        const T fh = (T)(row_dot<T>(J + (size_t)pos * n, &s[0], n) + h[pos]);
        const double dn = neg_delta<T>(ptype, s[pos], fh);
        double refd = 0.0;
        if (ref == 1) refd = (s[pos] == (T)1) ? (lp - l1p) : (l1p - lp);
        const double full = (double)beta * dn + (1.0 - (double)beta) * refd;
        if (full >= 0 || lr < full) {
          ones += (s[pos] == (T)1) ? -1 : 1;
          s[pos] = flipped<T>(ptype, s[pos]);
        }
      }
      E = energy();  // recomputed after every step (simulation.py:309)
      if (ordering == 0) {
        const double refE2 = (ref == 1) ? -((double)(n - ones) * lp + (double)ones * l1p) : 0.0;
        lw += (-(double)beta * E - (1.0 - (double)beta) * refE2) -
              (-bold * E - (1.0 - bold) * refE2);
      }
      bold = (double)beta;
    }
    for (int q = 0; q < n; ++q) states[(size_t)ch * n + q] = (int8_t)s[q];
    logw[ch] = lw;
  }
  return 0;
}

}  // namespace

extern "C" {

// Philox4x32-10 block for tests (known-answer vectors).
void oracle_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
  W4 c = {{ctr[0], ctr[1], ctr[2], ctr[3]}};
  W4 r = philox10(c, key[0], key[1]);
  for (int i = 0; i < 4; ++i) out[i] = r.v[i];
}

void oracle_det_log(const double* x, double* y, int64_t n) {
  for (int64_t i = 0; i < n; ++i) y[i] = dlog(x[i]);
}

void oracle_perm(uint32_t n, const uint32_t* keys, uint32_t* out) {
  W4 k = {{keys[0], keys[1], keys[2], keys[3]}};
  Perm p(n, k);
  for (uint32_t i = 0; i < n; ++i) out[i] = p.at(i);
}

// The uniform/permutation stream of one (chain, sweep), for tests.
void oracle_sweep_stream(uint64_t seed, uint64_t chain, uint64_t sweep, int n, int order,
                         double* log_r, int32_t* perm) {
  const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  Perm p(n, philox10(ctr_perm(order == 2 ? 0 : chain, sweep, order == 2 ? 3u : 2u), k0, k1));
  for (int pos = 0; pos < n; ++pos) {
    W4 b = philox10(ctr_sweep(chain, sweep, (uint32_t)pos), k0, k1);
    log_r[pos] = dlog(u01(b.v[pos & 3]));
  }
  if (order == 3) {
    const std::vector<int> b = BlockOrder(n, sweep, k0, k1).positions();
    for (int i = 0; i < n; ++i) perm[i] = b[i];
    return;
  }
  for (int i = 0; i < n; ++i) perm[i] = order == 0 ? i : (int32_t)p.at((uint32_t)i);
}

// E = s.(J s / 2 + h) in f64 for a batch of states (ising.py:37).
void oracle_energy_f64(int n, const double* J, const double* h, const int8_t* states, int64_t count,
                       double* out) {
  for (int64_t c = 0; c < count; ++c) {
    const int8_t* s = states + c * n;
    double E = 0.0;
    for (int i = 0; i < n; ++i) {
      double r = 0.0;
      for (int j = 0; j < n; ++j) r += J[(size_t)i * n + j] * (double)s[j];
      E += (double)s[i] * (0.5 * r + h[i]);
    }
    out[c] = E;
  }
}

int oracle_metropolis(int n, int ptype, int dtype, const void* J, const void* h, int reads, int R,
                      int S, int order, int use_pt, int logw, int best_sem, const void* betas,
                      const int32_t* beta_sort, uint64_t seed, uint64_t sweep_offset,
                      const double* rep_lr, const int32_t* rep_perm, const double* rep_pt,
                      int8_t* states, void* energies, int64_t* beta_idx, int8_t* best_state,
                      void* best_energy, int32_t* best_sweeps, double* log_omegas,
                      uint64_t* accepted, int nthreads) {
  RunCfg c = {n, ptype, reads, R, S, order, use_pt, logw, best_sem, seed, sweep_offset,
              rep_lr, rep_perm, rep_pt, nthreads};
  if (dtype == 0)
    return metropolis<float>(c, (const float*)J, (const float*)h, (const float*)betas, beta_sort,
                             states, (float*)energies, beta_idx, best_state, (float*)best_energy,
                             best_sweeps, log_omegas, accepted);
  return metropolis<double>(c, (const double*)J, (const double*)h, (const double*)betas, beta_sort,
                            states, (double*)energies, beta_idx, best_state, (double*)best_energy,
                            best_sweeps, log_omegas, accepted);
}

// Batched instances (the engine's pysa_amd_batched_device): instance k runs
// the Metropolis / PT chain above with couplings scale[k] * signs_k and read
// indices k*reads + read in the random streams.
int oracle_batched(int count, const int32_t* ns, const int32_t* ld, const int64_t* j_off,
                   const int8_t* signs, const float* scale, const float* h, const int64_t* h_off,
                   int ptype, int reads, int R, int S, int order, int use_pt, int best_sem,
                   const float* betas, uint64_t seed, uint64_t sweep_offset, int8_t* states,
                   float* energies, int64_t* beta_idx, int8_t* best_state, float* best_energy,
                   int32_t* best_sweeps, uint64_t* accepted, int nthreads) {
  int64_t so = 0;
  std::vector<int32_t> bsort(R);
  for (int k = 0; k < R; ++k) bsort[k] = k;
  for (int k = 0; k < count; ++k) {
    const int n = ns[k];
    RunCfg c = {n, ptype, reads, R, S, order, use_pt, 0, best_sem, seed, sweep_offset,
                nullptr, nullptr, nullptr, nthreads};
    c.signs = signs + j_off[k];
    c.ld = ld[k];
    c.scale = scale[k];
    c.read_base = (uint64_t)k * reads;
    std::vector<float> hk(n, 0.f);
    if (h)
      for (int i = 0; i < n; ++i) hk[i] = h[h_off[k] + i];
    const size_t g = (size_t)k * reads;
    metropolis<float>(c, nullptr, hk.data(), betas + (size_t)k * R, bsort.data(), states + so,
                      energies + g * R, beta_idx + g * R, best_state + so / R, best_energy + g,
                      best_sweeps + g, nullptr, accepted ? accepted + g : nullptr);
    so += (int64_t)reads * R * n;
  }
  return 0;
}

int oracle_ais(int n, int ptype, int dtype, const void* J, const void* h, const void* temps, int nt,
               int chains, int ref, double p, int order, int ordering, uint64_t seed,
               const double* rep_lr, const int32_t* rep_perm, int8_t* states, double* logw,
               int nthreads) {
  if (dtype == 0)
    return ais<float>(n, ptype, (const float*)J, (const float*)h, (const float*)temps, nt, chains,
                      ref, p, order, ordering, seed, rep_lr, rep_perm, states, logw, nthreads);
  return ais<double>(n, ptype, (const double*)J, (const double*)h, (const double*)temps, nt, chains,
                     ref, p, order, ordering, seed, rep_lr, rep_perm, states, logw, nthreads);
}

}  // extern "C"
