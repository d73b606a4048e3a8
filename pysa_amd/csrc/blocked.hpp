// MFMA-blocked form of the engine (declarations).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

namespace pysa {

// Largest problem the blocked form holds (spin bits of 64 replicas in VGPRs).
constexpr int kBlockedMaxN = 2048;

// Scratch size of the allocation-free entry point (best states of
// reads * R chains of n spins plus 8 bytes per chain), and devices it serves.
constexpr size_t kNoAllocScratchBytes = size_t(32) << 20;
constexpr int kMaxDevices = 64;

// Packed couplings and scratch of the blocked form, owned by a problem.
struct BlockedProblem {
  void* packed = nullptr;  // MFMA A-operand fragments of J
  size_t packed_bytes = 0;
  int packed_kind = -1;  // 0 int8, 1..3 bf16 pieces
  int exact_kind = -1;   // smallest exact encoding of this J
  int exact_sums = -1;   // 1: J and h integers with every field sum exact in f32
  int h_zero = -1;       // 1: every local field is zero
  int last_kind = -1;    // encoding used by the last run
  void* scratch = nullptr;
  size_t scratch_bytes = 0;
  float* rowsum = nullptr;  // J . 1 per row (QUBO with bf16 operands)
  int rowsum_kind = -1;     // encoding the row sums were taken for
  float* jtl = nullptr;     // J^T as its bf16 piece sums (inexact encodings only)
  int jtl_kind = -1;
  // Launch shape (results do not depend on it): kernel 0 = by size,
  // 1 = pipelined, 2 = split-K; tiles of 32 replicas per workgroup of the
  // pipelined kernel, 0 = by occupancy, 1 or 2.
  int shape_kernel = 0;
  int shape_tiles = 0;
};

void blocked_release(BlockedProblem& b);

// Whether the blocked form can run this request (sequential order, or the
// block-random order for n > 64).
bool blocked_eligible(int n, int dtype, int R, int reads, int order, bool replay);

// pieces: 0 = exact where possible else hi+lo bf16; 2 or 3 = that many bf16
// pieces (3 is exact for every f32).
int run_blocked(BlockedProblem& b, const void* dJ, const void* dJt, const void* dh, int n,
                int ptype, int dtype, int reads, int R, int S, int order, int use_pt, int logw,
                int best_sem, int pieces, uint64_t seed, uint64_t sweep_offset, const void* betas,
                const int32_t* bsort, const double* rep_lr, const double* rep_pt, int8_t* states,
                void* energies, int64_t* beta_idx, int8_t* best_state, void* best_energy,
                int32_t* best_sweeps, double* log_omegas, uint64_t* accepted, hipStream_t st,
                double* ms, std::string* why);

// The same run with no allocation, no event and no synchronisation: the
// couplings must already be encoded for this form (by an earlier run_blocked
// on the problem), the best-state scratch is `scratch_mem` (scratch_bytes
// long), and the work is enqueued on st.
int run_blocked_noalloc(const BlockedProblem& b, const void* dJt, const void* dh, int n,
                        int ptype, int dtype, int reads, int R, int S, int order, int use_pt,
                        int logw, int best_sem, int pieces, uint64_t seed, uint64_t sweep_offset,
                        const void* betas, const int32_t* bsort, int8_t* states, void* energies,
                        int64_t* beta_idx, int8_t* best_state, void* best_energy,
                        int32_t* best_sweeps, double* log_omegas, uint64_t* accepted,
                        void* scratch_mem, size_t scratch_bytes, hipStream_t st,
                        std::string* why);

// Device address and size of the allocation-free scratch on the current
// device (looked up on the first call per device).
int noalloc_scratch(void** ptr, size_t* bytes, std::string* why);

// AIS with the uniform reference on the pipelined MFMA kernel (float32,
// 65 <= n <= 2048, sequential or block-random order), all arrays in device
// memory: `chains` chains of n spins, the schedule temps[nt], energy in (the
// f64 energies of the initial states) and out (tracked), log_w out.
int run_blocked_ais(BlockedProblem& b, const void* dJ, const void* dJt, const void* dh, int n,
                    int ptype, int chains, int nt, const float* temps, int order, int ordering,
                    int pieces, uint64_t seed, int8_t* states, double* energy, double* log_w,
                    hipStream_t st, double* ms, std::string* why);

}  // namespace pysa
