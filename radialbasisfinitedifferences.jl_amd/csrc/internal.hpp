// internal.hpp -- host-side declarations shared by the translation units of
// librbffd (kernel launchers, argument blocks, error plumbing).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

namespace rbffd {

constexpr int kMaxOps = 10;
constexpr int kMaxRhs = 6;   // right-hand sides per solve launch
constexpr int kMaxGenRhs = 72;  // right-hand-side columns of the generic kernel (rows x operators)

// Failures inside the library are thrown as this and turned into a status code
// plus a thread-local message at the C-ABI boundary.
struct Error {
    int code;
    std::string msg;
};

#define RBFFD_HIP(call)                                                                   \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess)                                                             \
            throw ::rbffd::Error{3, std::string(#call) + ": " + hipGetErrorString(e_)};   \
    } while (0)

// ---------------------------------------------------------------- kNN (K1)
struct GridParams {
    double x0, y0;    // lower-left corner
    double csx, csy;  // cell sizes
    int gx, gy;       // cells per axis
    int valid;        // 0: the points are not finite; queries return -1
};

// Bounding box of X (device): box[5] = (xmin, ymin, xmax, ymax, non-finite count);
// partials holds 5 values per block.
void launch_bbox(const double *X, long N, double *box, double *partials, int nblocks, hipStream_t s);
// Cell index of every point, the per-cell counts (counts must be zeroed) and
// each point's slot inside its cell (slot: N ints).
void launch_cell_count(const double *X, long N, const GridParams *g, int *cell_of, int *counts, int *slot,
                       hipStream_t s);
// Upper bound on the grid's cell count for N points.
__host__ __device__ inline long grid_max_cells(long N) { return 4 * N + 64; }
// Grid parameters computed on the device from box (see launch_bbox); bit 0 of
// *flag (may be NULL) is set when the box is not finite.
void launch_grid_params(const double *box, long N, int k, GridParams *out, int *flag, hipStream_t s);
// p[0 .. n) = v; kernels, so they can be recorded into a graph.
void launch_fill_i32(int *p, long n, int v, hipStream_t s);
void launch_fill_i64(int64_t *p, long n, int64_t v, hipStream_t s);
// Exclusive scan of counts (ncell) into start (ncell + 1); tmp as returned by scan_tmp_bytes.
size_t scan_tmp_bytes(long n);
void launch_exclusive_scan(const int *in, int *out, long n, void *tmp, size_t tmp_bytes, hipStream_t s);
// Scatter points into cell order: point i goes to start[cell_of[i]] + slot[i].
void launch_cell_scatter(const double *X, const int *xgroup, long N, const int *cell_of, const int *start,
                         const int *slot, double *cell_xy, int *cell_idx, int *cell_grp, hipStream_t s);
// Exact kNN of every query, ascending (d^2, index), over the N points of the
// cell arrays; *unfilled (may be NULL) gains the result slots left at -1 (a
// query that sees fewer than k nodes, or non-finite points).  redo (M ints, may be
// NULL) and redo_count (one int) are the workspace of the two-phase selection
// used for k > 8; without them every query runs the insertion kernel.  With
// them, rw holds the queries' k-th distances (M doubles) and the list (M ints)
// and count (one int) of queries left to the insertion kernel.
struct KnnRedo {
    double *tau;
    int *list2;
    int *count2;
};
void launch_knn_query(const GridParams *g, const int *cell_start, const double *cell_xy, const int *cell_idx,
                      const int *cell_grp, long N, const double *Q, const int *qgroup, const int *qorder,
                      long M, int k, int *idx_out, double *d2_out, int *unfilled, int *redo, int *redo_count,
                      const KnnRedo &rw, hipStream_t s);

// ------------------------------------------------------- local solves (K2)
struct SolveArgs {
    const double *X;        // N x 2 interleaved
    const double *Y;        // M x 2 interleaved
    const int *stencils;    // N x n, 0-based, first = centre
    const int *y2x;         // M: stencil of each Y row (NULL: collocated)
    const int *rows;        // task -> Y row (NULL: identity)
    long ntasks;
    int n, P, m, method;    // stencil size, monomials, PHS power, 1 or 3
    int nops;
    int op_kind[kMaxOps];
    int op_k[kMaxOps];
    int *colidx;            // M x n (NULL: skip)
    double *vals[kMaxOps];  // M x n each
    int *status;            // M (NULL: skip)
    int status_accumulate;  // keep the larger of the stored and the new status
    // mixed precision refinement
    int max_iter;
    double tol;
    int group;              // lanes per stencil of the FP64 kernel (0: default)
    // ---- the generic kernel (stencil_gen.hpp): any (n, P), and output rows
    // sharing one stencil.  Tasks of several rows: task -> X node grp_x, its
    // rows ylist[grp_start .. grp_start + grp_count); NULL: one row per task.
    const int *grp_x, *grp_start, *grp_count, *ylist;
    // grouped tasks of a shape with a specialised kernel: 1 = solve them in
    // that kernel's one-slot layout (gen_cols <= kLuGrpMaxRhs), 0 = the generic one
    int grp_tuned;
    const int *ntasks_dev;  // if set, tasks at or past *ntasks_dev do nothing
    int gen_cols;           // right-hand-side columns of the widest task
    // right-hand-side column cc: output row col_sub[cc] of the task, operator
    // col_op[cc] (index into vals) of kind col_kind[cc] and order col_k[cc]
    int8_t col_sub[kMaxGenRhs], col_op[kMaxGenRhs], col_kind[kMaxGenRhs], col_k[kMaxGenRhs];
    // systems above 128 equations: the augmented systems of mem_tasks stencils
    // at a time, mem_ld rows (padded) by n + P + nops columns each
    double *scratch;
    long mem_ld, mem_tasks;
};

// Shapes the library solves: stencils of 1..128 nodes (the neighbour
// search's limit) and systems of at most 512 equations.
bool solve_supported(int n, int P);
// Shapes with a specialised kernel (precision 1: with a mixed-precision one).
bool solve_tuned(int n, int P, int precision);
// precision: 0 = FP64 Gauss-Jordan, 1 = FP32 Gauss-Jordan inverse + FP64
// refinement (specialised shapes only; the others solve in FP64).  Tasks of
// several rows (a.grp_x) and shapes without a specialised kernel run the
// generic kernel.  a.nops <= kMaxRhs.
void launch_solve(const SolveArgs &a, int precision, hipStream_t s);
// Compiled lane-group sizes for a shape: groups[0] default, groups[1..kMaxGroups-1]
// alternatives or 0.
constexpr int kMaxGroups = 3;
void solve_groups(int n, int P, int precision, int *groups);
// Output rows per task of the generic kernel for nops operators when an X node
// has at most maxrows of them and mean_rows on average (1: no grouping).
int gen_rows_per_task(int n, int P, int nops, int maxrows, double mean_rows);
// Lanes per stencil of the generic kernel for R right-hand-side columns.
int gen_lanes_used(int n, int P, int R);
// Workspace doubles per stencil of the generic kernel for systems above 128
// equations (0 below): mem_ld rows by n + P + nops columns.
long gen_mem_ld(int n, int P);
// Y rows grouped by the X node of their stencil, in tasks of at most ky rows.
// launch_group_count: counts[x] = rows with y2x = x (counts: N + 1, fill: N,
// both cleared here).  launch_group_build: ystart, ntask, tstart (N + 1 each),
// ylist (M), grp_x, grp_start, grp_count (M each; the task count, tstart[N],
// is at most M).
size_t group_tmp_bytes(long N);
void launch_group_count(const int *y2x, long M, long N, int *counts, int *fill, hipStream_t s);
void launch_group_build(const int *y2x, long M, long N, int ky, const int *counts, int *ystart, int *fill,
                        int *ntask, int *tstart, int *ylist, int *grp_x, int *grp_start, int *grp_count, void *tmp,
                        size_t tmp_bytes, hipStream_t s);

// Compiled stencil shapes (n, P): the reference's tests (20,10) and (42,21),
// and the synthetic node-set shapes (13,6), (21,10), (31,15), (43,21), (57,28).
#define RBFFD_SHAPES(X) X(13, 6) X(20, 10) X(21, 10) X(31, 15) X(42, 21) X(43, 21) X(57, 28)
#define RBFFD_DECLARE_SHAPE(nn, pp)                                                     \
    void launch_solve_##nn##_##pp(const SolveArgs &a, int precision, hipStream_t s); \
    void launch_solve_grp_##nn##_##pp(const SolveArgs &a, hipStream_t s);            \
    void groups_##nn##_##pp(int precision, int *groups);
RBFFD_SHAPES(RBFFD_DECLARE_SHAPE)
#undef RBFFD_DECLARE_SHAPE
// Grouped tasks in a specialised kernel: right-hand-side blocks of 6, 12 and
// 24 columns, in the layout that holds the stencil's nodes in one row slot
// per lane.
constexpr int kLuGrpMaxRhs = 24;
constexpr int lu_grp_lanes(int n) { return n <= 16 ? 16 : (n <= 32 ? 32 : 64); }
// Generic kernel per padded system size NSB (R: right-hand-side columns
// needed): gen_lanes(NSB) lanes per stencil, right-hand-side blocks of 3, 6 and
// gen_rb(NSB) columns (several output rows per task where that exceeds 6).
#define RBFFD_GEN_BUCKETS(X) X(16) X(32) X(48) X(64) X(96) X(128)
constexpr int gen_lanes(int nsb) { return nsb <= 16 ? 16 : (nsb <= 32 ? 32 : 64); }
constexpr int gen_rb(int nsb) { return nsb <= 48 ? 48 : (nsb <= 64 ? 72 : 6); }
// the narrow layout (two row slots per lane) for at most kGenNarrowRhs columns; 0: none
constexpr int kGenNarrowRhs = 12;
constexpr int gen_narrow_lanes(int nsb) { return nsb == 32 ? 16 : 0; }
// lanes per stencil of the generic kernel for R right-hand-side columns
constexpr int gen_lanes_for(int nsb, int R)
{
    return (gen_narrow_lanes(nsb) && R <= kGenNarrowRhs) ? gen_narrow_lanes(nsb) : gen_lanes(nsb);
}
#define RBFFD_DECLARE_GEN(nsb) void launch_gen_##nsb(const SolveArgs &a, int R, hipStream_t s);
RBFFD_GEN_BUCKETS(RBFFD_DECLARE_GEN)
#undef RBFFD_DECLARE_GEN
void launch_gen_mem(const SolveArgs &a, hipStream_t s);
// Two wavefronts per stencil (one row per lane) for the 96- and 128-equation
// sizes, at most 24 right-hand-side columns.
void launch_gen2_96(const SolveArgs &a, int R, hipStream_t s);
void launch_gen2_128(const SolveArgs &a, int R, hipStream_t s);
constexpr int kGen2MaxRhs = 24;

// --------------------------------------------------- sparse (K3, K4)
void launch_fill_rowptr(int64_t *rowptr, long M, int n, hipStream_t s);
void launch_csc_count(const int *colidx, long nnz, int *counts, hipStream_t s);
void launch_csc_scatter(const int *colidx, long M, int n, const int *colptr, int *fill, int *perm, hipStream_t s);
void launch_csc_rank(const int *colidx, long M, int n, const int *colptr, const int *perm_in, int *perm,
                     int *rowval, hipStream_t s);
void launch_gather_f64(const double *src, const int *perm, long nnz, double *dst, hipStream_t s);
void launch_i32_to_i64(const int *src, long n, int64_t *dst, int64_t add, hipStream_t s);

// y[M x F] = A x[N x F], CSR with int32 columns.  vals_f32 selects FP32 values
// (accumulated in FP64).
// Rows run in `order` (a permutation of 0..M-1) when it is not NULL.
void launch_spmv(long M, const int64_t *rowptr, const int *col, const void *vals, int vals_f32, int F,
                 const double *x, double *y, double alpha, double beta, int uniform_n, const int *order,
                 hipStream_t s);
void launch_f64_to_f32(const double *src, long n, float *dst, hipStream_t s);
// *out = sum of x[0 .. n) by a read-only stream over cus * 8 blocks (x 16-byte aligned)
void launch_read_sum(const double *x, long n, double *out, int cus, hipStream_t s);

// Application plan for repeated products with one CSR operator
// (spmv_plan.hip): rows in blocks of kPlanRows consecutive rows of a locality
// order, each block's distinct columns listed (at most kPlanHaloCap, sorted)
// and its nonzeros given 16-bit local indices into that list.
constexpr int kPlanRows = 256;
constexpr int kPlanHaloCap = 2048;
struct SpmvPlan {
    long M = 0, N = 0, nnz = 0;
    int nblocks = 0;
    int hmax = 0;             // largest halo among the listed blocks
    int overflow_blocks = 0;  // blocks with more than kPlanHaloCap distinct columns
    int uniform = 0;          // entries of every row when all rows have as many, else 0
    bool f32 = false;
    int *prow = nullptr;      // plan row -> operator row
    int64_t *pptr = nullptr;  // M + 1, the rows in plan order
    uint16_t *plc = nullptr;  // nnz local column indices
    int *pcol = nullptr;      // nnz columns; kept only when some block overflows
    void *pval = nullptr;     // nnz values (FP64 or FP32), plan order
    int *halo = nullptr;      // nblocks x kPlanHaloCap
    int *hcount = nullptr;    // nblocks (+ 2 scratch)
};
// Allocates the plan's arrays (hipMalloc) and synchronises s.
void spmv_plan_build(SpmvPlan &P, long M, long N, const int64_t *rowptr, const int *col, const void *vals,
                     bool f32, int uniform_n, const int *order, hipStream_t s);
void spmv_plan_free(SpmvPlan &P);
// y = alpha A x + beta y; bitwise equal to launch_spmv with the default lanes.
void launch_spmv_plan(const SpmvPlan &P, int F, const double *x, double *y, double alpha, double beta,
                      hipStream_t s);
// launch_rk_stage through a plan; bitwise equal to it.
void launch_rk_stage_plan(const SpmvPlan &P, int F, const double *w, const double *u, double *out,
                          const uint8_t *fixed, double a, double b, double c, hipStream_t s);

// One SSP-RK stage out = a u + b w + c (L w) (rows with fixed != 0 keep u).
void launch_rk_stage(long M, const int64_t *rowptr, const int *col, const void *vals, int vals_f32, int F,
                     int uniform_n, const int *order, const double *w, const double *u, double *out,
                     const uint8_t *fixed, double a, double b, double c, hipStream_t s);
// Ghost-node closure: u[ghost[i]] = -(G t)_i for i < nbc (F fields), G dense
// column-major nbc x nbc, t = nbc x F; part holds closure_chunks(nbc) x nbc x F.
long closure_chunks(long nbc);
void launch_closure(long nbc, int F, const double *G, const double *t, double *part, const int *ghost, double *u,
                    hipStream_t s);
// Rows of N nodes in Morton order of their coordinates (box from launch_bbox);
// key, key_sorted, idx: N-entry workspaces; tmp of morton_tmp_bytes(N).
size_t morton_tmp_bytes(long N);
void launch_morton_order(const double *X, long N, const double *box, uint32_t *key, uint32_t *key_sorted,
                         int *idx, int *order, void *tmp, size_t tmp_bytes, hipStream_t s);
// Terms of a linear combination of value arrays sharing one pattern.
struct CombineTerms {
    int nops;
    const double *vals[kMaxOps];
    double coef[kMaxOps];
};
// out = sum_o coef[o] vals[o] (FP64, may be NULL) and its FP32 rounding (may be NULL).
void launch_combine(long nnz, const CombineTerms &t, double *out, float *out32, hipStream_t s);

// --------------------------------------------------- sparse solve (CGLS)
// Grid-wide sum target: block sums go to partial[kLsqBlocks], the counter is
// zero between launches, the total lands in *out.
constexpr int kLsqBlocks = 1024;
struct LsqReduce {
    double *partial;
    unsigned *count;
    double *out;
};
// y = A x (CSR, int64 rowptr) and |y|^2 -> *red.out
void launch_lsq_spmv(long M, long nnz, const int64_t *rowptr, const int *col, const double *val, const double *x,
                     double *y, const LsqReduce &red, hipStream_t s);
// z += a p, r -= a q with a = *gam / *qq
void launch_lsq_update(long N, long M, const double *gam, const double *qq, const double *p, const double *q,
                       double *z, double *r, hipStream_t s);
// p = s + (*gam_new / *gam) p
void launch_lsq_direction(long N, const double *gam, const double *gam_new, const double *sv, double *p,
                          hipStream_t s);
void launch_row_of(const int64_t *rowptr, long M, int *row_of, hipStream_t s);
void launch_gather_i32(const int *src, const int *idx, long n, int *dst, hipStream_t s);
void launch_col_scale(const int64_t *tptr, long N, const double *tval, double *scale, hipStream_t s);
void launch_scale_cols(const int *col, long nnz, const double *scale, const double *v, double *out, hipStream_t s);
void launch_copy_f64(const double *x, long n, double *y, hipStream_t s);
void launch_lsq_finish(long N, const double *scale, const double *z, double *x, hipStream_t s);
void launch_fill_f64(double *p, long n, double v, hipStream_t s);
// Direct least squares: D (CSR) densified into A (M x N column-major
// workspace), b copied to B (M), rocSOLVER QR solve, x = B[0:N].  *handle is
// the caller's solver handle (created on first use).  Returns gels' info
// (> 0: rank deficient).
bool dense_solver_available();
void dense_solver_destroy(void *handle);
int dense_lsq(void **handle, long M, long N, const int64_t *rowptr, const int *col, const double *vals,
              const double *b, double *x, double *A, double *B, int *info, hipStream_t s);
// Ainv = inv(A), n x n column-major, through rocSOLVER's QR solve of A X = I;
// Aw (n x n) is overwritten.  Returns gels' info (> 0: singular).
int dense_inverse(void **handle, long n, double *Aw, double *Ainv, int *info, hipStream_t s);

}  // namespace rbffd
