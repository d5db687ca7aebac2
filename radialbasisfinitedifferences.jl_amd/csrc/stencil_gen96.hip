// Generic local-solve kernels for systems padded to 96 equations.
#include "stencil_gen.hpp"

RBFFD_DEFINE_GEN(96)
