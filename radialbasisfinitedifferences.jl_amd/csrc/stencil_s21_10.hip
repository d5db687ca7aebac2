// Local-solve kernels for stencils of 21 nodes with 10 monomials.
#include "stencil_impl.hpp"

RBFFD_DEFINE_SHAPE(21, 10)
