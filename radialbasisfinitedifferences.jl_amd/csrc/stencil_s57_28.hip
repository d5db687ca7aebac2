// Local-solve kernels for stencils of 57 nodes with 28 monomials.
#include "stencil_impl.hpp"

RBFFD_DEFINE_SHAPE(57, 28)
