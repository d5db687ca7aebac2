// Generic local-solve kernels for systems padded to 128 equations (one and two
// wavefronts per stencil).
#include "stencil_gen2.hpp"

RBFFD_DEFINE_GEN(128)

namespace rbffd {
void launch_gen2_128(const SolveArgs &a, int R, hipStream_t s) { launch_gen2_bucket<128>(a, R, s); }
}  // namespace rbffd
