// MFMA-blocked form of the engine (declarations).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

namespace pysa {

// Packed couplings and scratch of the blocked form, owned by a problem.
struct BlockedProblem {
  void* packed = nullptr;  // MFMA-ready coupling tiles
  size_t packed_bytes = 0;
  int packed_b = 0, packed_split = 0;
  void* scratch = nullptr;
  size_t scratch_bytes = 0;
};

void blocked_release(BlockedProblem& b);

// Whether the blocked form can run this request (auto form selection).
bool blocked_eligible(int n, int dtype, int R, int reads, int order, bool replay);

int run_blocked(BlockedProblem& b, const void* dJ, const void* dh, int n, int ptype, int dtype,
                int reads, int R, int S, int order, int use_pt, int logw, int best_sem, int block_b,
                uint64_t seed, uint64_t sweep_offset, const void* betas, const int32_t* bsort,
                const double* rep_lr, const double* rep_pt, int8_t* states, void* energies,
                int64_t* beta_idx, int8_t* best_state, void* best_energy, int32_t* best_sweeps,
                double* log_omegas, uint64_t* accepted, hipStream_t st, double* ms,
                std::string* why);

}  // namespace pysa
