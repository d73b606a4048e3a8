// Pipelined MFMA kernel, 2 bf16 pieces.
#include "blocked_pipe.hpp"

namespace pysa {
namespace blk {

hipError_t launch_pipe_bf16_2(const BParams& p, int nw, int nt, int groups, hipStream_t st) {
  return launch_pipe_enc<2, false>(p, nw, nt, groups, st);
}

hipError_t launch_pipe_ais_bf16_2(const BParams& p, int nw, int groups, hipStream_t st) {
  return launch_pipe_ais_enc<2, false>(p, nw, groups, st);
}

}  // namespace blk
}  // namespace pysa
