// Local-solve kernels for stencils of 42 nodes with 21 monomials.
#include "stencil_impl.hpp"

RBFFD_DEFINE_SHAPE(42, 21)
