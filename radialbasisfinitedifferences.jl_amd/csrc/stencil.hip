// stencil.hip -- dispatch of K2 (the local solves) to the per-shape kernels
// compiled in stencil_s<n>_<P>.hip from stencil_impl.hpp.
#include <string>

#include "internal.hpp"

namespace rbffd {

bool solve_supported(int n, int P, int precision)
{
#define RBFFD_CASE(nn, pp) \
    if (n == nn && P == pp) return precision == 0 || nn + pp <= 64;
    RBFFD_SHAPES(RBFFD_CASE)
#undef RBFFD_CASE
    return false;
}

void launch_solve(const SolveArgs &a, int precision, hipStream_t s)
{
    if (a.ntasks <= 0) return;
#define RBFFD_CASE(nn, pp) \
    if (a.n == nn && a.P == pp) { launch_solve_##nn##_##pp(a, precision, s); return; }
    RBFFD_SHAPES(RBFFD_CASE)
#undef RBFFD_CASE
    throw Error{2, "no compiled kernel for stencil size n=" + std::to_string(a.n) + " with " +
                       std::to_string(a.P) + " monomials"};
}

void solve_groups(int n, int P, int precision, int *groups)
{
    groups[0] = groups[1] = groups[2] = 0;
#define RBFFD_CASE(nn, pp) \
    if (n == nn && P == pp) { groups_##nn##_##pp(precision, groups); return; }
    RBFFD_SHAPES(RBFFD_CASE)
#undef RBFFD_CASE
}

}  // namespace rbffd
