// Local-solve kernels for stencils of 20 nodes with 10 monomials.
#include "stencil_impl.hpp"

RBFFD_DEFINE_SHAPE(20, 10)
