// Local-solve kernels for stencils of 13 nodes with 6 monomials.
#include "stencil_impl.hpp"

RBFFD_DEFINE_SHAPE(13, 6)
