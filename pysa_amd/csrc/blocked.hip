// MFMA-blocked form of the engine.
#include "blocked.hpp"

#include "../../include/pysa_amd.h"

namespace pysa {

void blocked_release(BlockedProblem& b) {
  if (b.packed) (void)hipFree(b.packed);
  if (b.scratch) (void)hipFree(b.scratch);
  b = BlockedProblem();
}

bool blocked_eligible(int, int, int, int, int, bool) { return false; }

int run_blocked(BlockedProblem&, const void*, const void*, int, int, int, int, int, int, int, int,
                int, int, int, uint64_t, uint64_t, const void*, const int32_t*, const double*,
                const double*, int8_t*, void*, int64_t*, int8_t*, void*, int32_t*, double*,
                uint64_t*, hipStream_t, double*, std::string* why) {
  *why = "blocked form not built yet";
  return PYSA_AMD_ENOTSUP;
}

}  // namespace pysa
