// Local solves of systems above 128 equations (workspace-resident).
#include "stencil_gen.hpp"

namespace rbffd {
void launch_gen_mem(const SolveArgs &a, hipStream_t s) { launch_gen_mem_impl(a, s); }
}  // namespace rbffd
