// Generic local-solve kernels for systems padded to 48 equations.
#include "stencil_gen.hpp"

RBFFD_DEFINE_GEN(48)
