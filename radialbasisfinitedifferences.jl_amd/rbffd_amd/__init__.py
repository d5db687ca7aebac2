"""rbffd_amd -- Python mirror of the operator API of RadialBasisFiniteDifferences.jl
on librbffd, a gfx950 (MI355X) HIP engine behind a C ABI (include/rbffd.h)."""
from ._lib import NonFiniteError, RbffdError, SingularException, UndefRefError, load
from .device import Context, DeviceArray
from .mesh import processmesh
from .operators import (DeviceOperator, Engine, SpmvPlan, boundary_groups, default_context, generate_operator,
                        hyperviscosity_operator, operators, solve)

__all__ = ["Context", "DeviceArray", "DeviceOperator", "Engine", "SpmvPlan", "boundary_groups", "generate_operator", "hyperviscosity_operator", "operators", "solve", "processmesh",
           "default_context", "load", "SingularException", "UndefRefError", "NonFiniteError", "RbffdError"]
