// Generic local-solve kernels for systems padded to 64 equations.
#include "stencil_gen.hpp"

RBFFD_DEFINE_GEN(64)
