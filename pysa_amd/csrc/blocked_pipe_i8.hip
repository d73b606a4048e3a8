// Pipelined MFMA kernel, int8 couplings.
#include "blocked_pipe.hpp"

namespace pysa {
namespace blk {

hipError_t launch_pipe_i8(const BParams& p, int nw, int nt, int groups, hipStream_t st) {
  return launch_pipe_enc<1, true>(p, nw, nt, groups, st);
}

hipError_t launch_pipe_ais_i8(const BParams& p, int nw, int groups, hipStream_t st) {
  return launch_pipe_ais_enc<1, true>(p, nw, groups, st);
}

}  // namespace blk
}  // namespace pysa
