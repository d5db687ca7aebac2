// Generic local-solve kernels for systems padded to 96 equations (one and two
// wavefronts per stencil).
#include "stencil_gen2.hpp"

RBFFD_DEFINE_GEN(96)

namespace rbffd {
void launch_gen2_96(const SolveArgs &a, int R, hipStream_t s) { launch_gen2_bucket<96>(a, R, s); }
}  // namespace rbffd
