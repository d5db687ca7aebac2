// stencil.hip -- dispatch of K2 (the local solves): the specialised per-shape
// kernels compiled in stencil_s<n>_<P>.hip from stencil_impl.hpp, and the
// generic kernel (stencil_gen<NSB>.hip, stencil_genmem.hip) for every other
// shape and for output rows that share a stencil.  Also the grouping of Y
// rows by their stencil.
#include <cstdlib>
#include <string>

#include "internal.hpp"

namespace rbffd {

// RBFFD_SPECIALISED=0 sends every shape to the generic kernels and
// RBFFD_GEN_TWO_WAVES=0 keeps the 96- and 128-equation sizes in one
// wavefront per stencil (both for comparisons).
static bool env_off(const char *name)
{
    const char *e = std::getenv(name);
    return e && e[0] == '0';
}

namespace {
constexpr int kGenMaxN = 128;    // stencil nodes: the neighbour search's limit
constexpr int kGenMaxSys = 512;  // equations of the workspace-resident kernel
constexpr int kGenMaxRows = 64;  // output rows per task (one status bit each)

int gen_bucket(int nsys)
{
    const int sizes[] = {16, 32, 48, 64, 96, 128};
    for (int b : sizes)
        if (nsys <= b) return b;
    return 0;  // workspace-resident kernel
}
}  // namespace

bool solve_supported(int n, int P) { return n >= 1 && n <= kGenMaxN && P >= 1 && n + P <= kGenMaxSys; }

bool solve_tuned(int n, int P, int precision)
{
    if (env_off("RBFFD_SPECIALISED")) return false;
#define RBFFD_CASE(nn, pp) \
    if (n == nn && P == pp) return precision == 0 || nn + pp <= 64;
    RBFFD_SHAPES(RBFFD_CASE)
#undef RBFFD_CASE
    return false;
}

int gen_rows_per_task(int n, int P, int nops, int maxrows, double mean_rows)
{
    const int b = gen_bucket(n + P);
    if (b == 0 || nops < 1 || maxrows <= 1) return 1;
    int rb = 6;
#define RBFFD_RB(nsb) \
    if (b == nsb) rb = gen_rb(nsb);
    RBFFD_GEN_BUCKETS(RBFFD_RB)
#undef RBFFD_RB
    if ((b == 96 || b == 128) && !env_off("RBFFD_GEN_TWO_WAVES")) rb = kGen2MaxRhs;
    // every X node in one task where that fits the narrow layout; else the
    // narrow layout if an X node has about that many rows on average (the
    // rest split into a few tasks); else the widest block
    const int narrow = gen_narrow_lanes(b) ? kGenNarrowRhs / nops : 0;
    int ky;
    if (narrow >= maxrows) ky = maxrows;
    else if (narrow >= 2 && narrow >= (int)(mean_rows + 0.999)) ky = narrow;
    else ky = rb / nops < maxrows ? rb / nops : maxrows;
    return ky < 1 ? 1 : (ky > kGenMaxRows ? kGenMaxRows : ky);
}

int gen_lanes_used(int n, int P, int R)
{
    const int b = gen_bucket(n + P);
    if ((b == 96 || b == 128) && R <= kGen2MaxRhs && !env_off("RBFFD_GEN_TWO_WAVES")) return 128;
    return b ? gen_lanes_for(b, R) : 64;
}

long gen_mem_ld(int n, int P)
{
    const int nsys = n + P;
    return gen_bucket(nsys) ? 0 : 64L * ((nsys + 63) / 64);
}

void launch_solve(const SolveArgs &a, int precision, hipStream_t s)
{
    if (a.ntasks <= 0) return;
    if (!a.grp_x && solve_tuned(a.n, a.P, precision)) {
#define RBFFD_CASE(nn, pp) \
    if (a.n == nn && a.P == pp) { launch_solve_##nn##_##pp(a, precision, s); return; }
        RBFFD_SHAPES(RBFFD_CASE)
#undef RBFFD_CASE
    }
    if (a.grp_x && a.grp_tuned && solve_tuned(a.n, a.P, 0)) {
        if (a.gen_cols > kLuGrpMaxRhs) throw Error{3, "grouped specialised solve: bad launch arguments"};
#define RBFFD_CASE(nn, pp) \
    if (a.n == nn && a.P == pp) { launch_solve_grp_##nn##_##pp(a, s); return; }
        RBFFD_SHAPES(RBFFD_CASE)
#undef RBFFD_CASE
    }
    if (!solve_supported(a.n, a.P))
        throw Error{2, "no kernel for stencil size n=" + std::to_string(a.n) + " with " + std::to_string(a.P) +
                           " monomials (n must be in 1..128 and n + P at most 512)"};
    const int b = gen_bucket(a.n + a.P);
    if (b == 0) {
        if (a.grp_x || !a.scratch || a.mem_ld < a.n + a.P || a.gen_cols > kMaxRhs)
            throw Error{3, "workspace-resident solve: bad launch arguments"};
        launch_gen_mem(a, s);
        return;
    }
    if ((b == 96 || b == 128) && a.gen_cols <= kGen2MaxRhs && !env_off("RBFFD_GEN_TWO_WAVES")) {
        if (b == 96) launch_gen2_96(a, a.gen_cols, s);
        else launch_gen2_128(a, a.gen_cols, s);
        return;
    }
#define RBFFD_GEN(nsb) \
    if (b == nsb) { launch_gen_##nsb(a, a.gen_cols, s); return; }
    RBFFD_GEN_BUCKETS(RBFFD_GEN)
#undef RBFFD_GEN
}

void solve_groups(int n, int P, int precision, int *groups)
{
    groups[0] = groups[1] = groups[2] = 0;
    if (solve_tuned(n, P, precision)) {
#define RBFFD_CASE(nn, pp) \
    if (n == nn && P == pp) { groups_##nn##_##pp(precision, groups); return; }
        RBFFD_SHAPES(RBFFD_CASE)
#undef RBFFD_CASE
    }
    const int b = gen_bucket(n + P);
    groups[0] = b ? gen_lanes_for(b, kMaxRhs) : 64;
}

// ------------------------------------------------------------ grouping
namespace {
__global__ void group_count_kernel(const int *y2x, long M, int *counts)
{
    const long k = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (k < M) atomicAdd(&counts[y2x[k]], 1);
}

__global__ void group_ntask_kernel(const int *counts, long N, int ky, int *ntask)
{
    const long x = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (x < N) ntask[x] = (counts[x] + ky - 1) / ky;
    else if (x == N) ntask[x] = 0;
}

__global__ void group_scatter_kernel(const int *y2x, long M, const int *ystart, int *fill, int *ylist)
{
    const long k = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (k < M) {
        const int x = y2x[k];
        ylist[ystart[x] + atomicAdd(&fill[x], 1)] = (int)k;
    }
}

__global__ void group_tasks_kernel(const int *counts, const int *ystart, const int *tstart, long N, int ky,
                                   int *grp_x, int *grp_start, int *grp_count)
{
    const long x = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (x < N) {
        const int cnt = counts[x];
        for (int j = 0, tk = tstart[x]; j * ky < cnt; ++j, ++tk) {
            grp_x[tk] = (int)x;
            grp_start[tk] = ystart[x] + j * ky;
            grp_count[tk] = cnt - j * ky < ky ? cnt - j * ky : ky;
        }
    }
}

inline unsigned nblocks(long n) { return (unsigned)((n + 255) / 256); }
}  // namespace

size_t group_tmp_bytes(long N) { return scan_tmp_bytes(N + 1); }

void launch_group_count(const int *y2x, long M, long N, int *counts, int *fill, hipStream_t s)
{
    launch_fill_i32(counts, N + 1, 0, s);
    launch_fill_i32(fill, N, 0, s);
    group_count_kernel<<<nblocks(M), 256, 0, s>>>(y2x, M, counts);
}

void launch_group_build(const int *y2x, long M, long N, int ky, const int *counts, int *ystart, int *fill,
                        int *ntask, int *tstart, int *ylist, int *grp_x, int *grp_start, int *grp_count, void *tmp,
                        size_t tmp_bytes, hipStream_t s)
{
    launch_exclusive_scan(counts, ystart, N + 1, tmp, tmp_bytes, s);
    group_ntask_kernel<<<nblocks(N + 1), 256, 0, s>>>(counts, N, ky, ntask);
    launch_exclusive_scan(ntask, tstart, N + 1, tmp, tmp_bytes, s);
    group_scatter_kernel<<<nblocks(M), 256, 0, s>>>(y2x, M, ystart, fill, ylist);
    group_tasks_kernel<<<nblocks(N), 256, 0, s>>>(counts, ystart, tstart, N, ky, grp_x, grp_start, grp_count);
}

}  // namespace rbffd
