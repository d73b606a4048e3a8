// Shared declarations of the engine's kernels and launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

namespace pysa {

// Visit orders: sequential; a fresh Feistel permutation per chain and sweep;
// one permutation per sweep shared by every chain; the block-random order
// (rng.hpp) shared by every chain, the random order the MFMA form runs.
enum Order : int { kSequential = 0, kRandom = 1, kRandomShared = 2, kBlockRandom = 3 };

// Device-side parameters of one Metropolis / parallel-tempering run over
// `reads` independent reads of R replicas of one problem (plain form).
template <typename T>
struct PTParams {
  int n, ptype, R, S, reads;
  int order, use_pt, logw, best_sem;
  int refresh;  // recompute local fields in f64 every `refresh` sweeps (0 = only at start)
  uint32_t k0, k1;
  uint64_t sweep_offset;
  const T* Jt;       // n*n, row p holds the coefficients J[i][p] (column p of J)
  const T* Jp;       // the rows of Jt padded with zeros to ldp (a multiple of 256)
  int ldp;
  // Sparse form (null otherwise): the nonzeros of row p of the symmetric J
  // in ascending column order, padded to ell_d entries per row with column -1
  const int* ell_cols;  // n*ell_d
  const T* ell_vals;    // n*ell_d
  int ell_d;
  const T* h;        // n
  const T* betas;    // R, slot order (slot 0 = coldest)
  const int* bsort;  // R, slots sorted by ascending beta (log-omega)
  const double* rep_lr;   // reads*S*R*n or null
  const int* rep_perm;    // reads*S*R*n or null
  const double* rep_pt;   // reads*S*(R-1) or null
  int8_t* spins;          // reads*R*n (+1/-1 ising, 1/0 qubo)
  T* fields;              // reads*R*n, F = J s (without h)
  T* energies;            // reads*R
  int64_t* beta_idx;      // reads*R
  int8_t* best_rep;       // reads*R*n (per-replica best states) or reads*n
  T* best_rep_e;          // reads*R or reads
  int* best_rep_sw;       // reads*R or reads
  int8_t* best_state;     // reads*n
  T* best_energy;         // reads
  int* best_sweeps;       // reads
  double* log_omegas;     // reads*S or null
  unsigned long long* accepted;  // reads or null
};

// Annealed importance sampling over `chains` independent chains (plain form).
template <typename T>
struct AISParams {
  int n, ptype, chains, nt, order, ref, ordering;
  int refresh;  // recompute fields and energy in f64 every `refresh` steps (0 = never)
  uint32_t k0, k1;
  double lp, l1p;  // log p, log(1-p) of the Bernoulli reference
  const T* Jt;
  const T* Jp;            // rows padded to ldp, as in PTParams
  int ldp;
  const T* h;
  const T* temps;         // nt
  const double* rep_lr;   // chains*nt*n or null
  const int* rep_perm;    // chains*nt*n or null
  int8_t* spins;          // chains*n
  T* fields;              // chains*n
  double* energy;         // chains, f64 running energy
  double* logw;           // chains
};

template <typename T>
hipError_t launch_energy_f64(const T* Jt, const T* h, const int8_t* spins, double* out, int n,
                             long long count, hipStream_t st);
template <typename T>
hipError_t launch_plain_pt(const PTParams<T>& p, hipStream_t st);
template <typename T>
hipError_t launch_plain_ais(const AISParams<T>& p, hipStream_t st);

// Largest problem size the plain kernels hold in registers.
constexpr int kPlainMaxN = 4096;
// float64 chains above n = 2048 (64 register slots of doubles per lane) are
// refused: that instantiation took a memory fault in testing whose cause is
// not found, so its path stays off.
constexpr int kPlainMaxN64 = 2048;

// Many instances of different sizes in one launch (batched form).  Instance
// k: n[k] spins, couplings scale[k] * S_k with S_k an int8 matrix (row-major,
// leading dimension ld[k], a multiple of 16, at signs + j_off[k]), fields h +
// h_off[k] (h may be null), ladder betas[k*R ..], states at s_off[k] =
// sum_{j<k} reads*R*n[j]; best states at s_off[k]/R.
struct BatchParams {
  int count, reads, R, S, order, use_pt, best_sem, ptype;
  uint32_t k0, k1;
  uint64_t sweep_offset;
  const int32_t* inst_n;
  const int32_t* ld;
  const int64_t* j_off;
  const int8_t* signs;
  const float* scale;
  const float* h;
  const int64_t* h_off;
  const float* betas;
  const int64_t* s_off;
  const int32_t* order_map;  // workgroup -> (instance * reads + read), or null
  int8_t* states;
  float* energies;           // count*reads*R
  int64_t* beta_idx;         // count*reads*R
  int8_t* best_state;
  float* best_energy;        // count*reads
  int32_t* best_sweeps;      // count*reads
  unsigned long long* accepted;  // count*reads or null
  int8_t* best_rep;          // scratch, same layout as states
};
constexpr int kBatchedMaxN = 2048;
hipError_t launch_batched(const BatchParams& p, int items, int max_n, hipStream_t st);

// ---- batched MFMA form (blocked_pipe_bat.hip) ----
// The batched MFMA form runs two tiles of up to 32 chains per workgroup,
// each tile some reads of one instance; the two instances have the same
// number of 32-spin blocks, so they share the block schedule.  States of
// global read g, replica r of a tile start at spin_adj + (g R + r) n
// (spin_adj = the instance's state offset minus the global index of its
// first chain times n), its best state at best_adj + g n; the offsets are
// never negative once a chain's own term is added.  An empty tile has
// read0 == read_end.
struct BatTile {
  int n, ld, inst, read0, read_end;  // spins, row stride of S, instance, global reads [read0, read_end)
  float scale;
  int64_t frag;      // byte offset of the instance's int8 MFMA fragments
  int64_t sgn;       // offset of its int8 coupling matrix S
  int64_t h;         // offset of its local fields, -1 when h = 0
  int64_t spin_adj;  // see above
  int64_t best_adj;
};
struct BatWG {
  int nblk, kb, pad0, pad1;  // 32-spin blocks, K blocks per matrix wave
  BatTile t[2];
};
// One instance of the fragment packing.
struct BatPack {
  int n, nblk, ld, pad;
  int64_t sgn, frag;
};
// Smallest instance of the batched MFMA form (three blocks, one per matrix wave).
constexpr int kBatchedMfmaMinN = 65;
// Enqueues the instances `inst` (indices into the batch, largest first) of a
// batch on the pipelined MFMA kernel: packs their couplings, builds the
// workgroup table and launches.  Host copies of the instance arrays are
// passed in (scale, h_off: null when h is).  Owns the device buffers until
// destroyed, which must follow the stream's completion.  best_rep is the
// batch's state-layout scratch holding the initial states (best_sem 0).
struct BatchedMfmaRun {
  void* frag = nullptr;
  void* tab = nullptr;
  void* pack = nullptr;
  void* best_es = nullptr;
  std::vector<BatWG> wg;
  std::vector<BatPack> pk;
  ~BatchedMfmaRun();
  hipError_t enqueue(const BatchParams& bp, const std::vector<int>& inst, const int32_t* hn,
                     const int32_t* hld, const int64_t* hj, const int64_t* hs,
                     const int64_t* hh_off, const float* hscale, hipStream_t st);
};

}  // namespace pysa
