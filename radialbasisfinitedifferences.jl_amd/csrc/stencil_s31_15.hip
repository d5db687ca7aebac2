// Local-solve kernels for stencils of 31 nodes with 15 monomials.
#include "stencil_impl.hpp"

RBFFD_DEFINE_SHAPE(31, 15)
