/*
 * rbffd.h -- C ABI of librbffd, an MI355X (gfx950) engine for RBF-FD operator
 * generation and application.  Drop-in for the operator path of
 * RadialBasisFiniteDifferences.jl: the reference's Julia functions keep their
 * signatures and ccall these entry points (citations are relative to the
 * reference root).
 *
 * Conventions
 *   - Points are interleaved Float64 pairs x0, y0, x1, y1, ... (the memory of a
 *     Vector{SVector{2,Float64}}, test/poisson_test.jl:21-22).
 *   - *_device entry points take device pointers and run on the context's
 *     stream; they synchronise the stream before returning where they must read
 *     a size or a status back (noted per function).  Host entry points take host
 *     pointers and are fully synchronous.
 *   - Every function returns RBFFD_OK (0) or an error code; rbffd_last_error()
 *     gives the message of the calling thread's last failure.
 *   - Indices inside the library are 0-based int32; the CSC outputs meant for
 *     SparseMatrixCSC{Float64,Int64} are int64 with a caller-chosen base.
 *   - One context per concurrent caller; a context is not shared between
 *     threads without external locking.
 */
#ifndef RBFFD_H
#define RBFFD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RBFFD_VERSION 1

/* status codes (the Julia shim maps them to the exceptions the reference throws) */
#define RBFFD_OK 0
#define RBFFD_ERR_ARG 1          /* invalid argument          -> ArgumentError           */
#define RBFFD_ERR_UNSUPPORTED 2  /* outside the supported range (n > 128, n + P > 512)
                                                              -> ArgumentError           */
#define RBFFD_ERR_RUNTIME 3      /* HIP runtime failure       -> ErrorException          */
#define RBFFD_ERR_SINGULAR 4     /* exactly singular stencil  -> LinearAlgebra.SingularException
                                    (inv at src/interpolationmatrix.jl:8)               */
#define RBFFD_ERR_NONFINITE 5    /* non-finite weights (NaN/Inf coordinates)            */
#define RBFFD_ERR_UNDEFINED 6    /* node in no index set      -> UndefRefError
                                    (idxs_x never filled, src/calculateneighbors.jl:8-41) */

/* operator kinds: the right-hand-side columns of the reference's local solves */
#define RBFFD_OP_E 0     /* evaluation E         src/generate_operator.jl:157,166 */
#define RBFFD_OP_DX 1    /* d/dx                 :157,161 */
#define RBFFD_OP_DY 2    /* d/dy                 :157,162 */
#define RBFFD_OP_DXX 3   /* d2/dx2               :157,163 */
#define RBFFD_OP_DYY 4   /* d2/dy2               :157,164 */
#define RBFFD_OP_DXY 5   /* d2/dxdy              :157,165 */
#define RBFFD_OP_XK 6    /* d^k/dx^k             src/hyperviscosity_operator.jl:155,159 */
#define RBFFD_OP_YK 7    /* d^k/dy^k             src/hyperviscosity_operator.jl:155,160 */
#define RBFFD_OP_LAP 8   /* Laplacian, one solve (equals Dxx + Dyy by linearity) */
#define RBFFD_OP_LAPK 9  /* Laplacian^k with the per-axis stencil scaling (no reference counterpart) */
#define RBFFD_MAX_OPS 10

/* precision of the local solves */
#define RBFFD_PREC_FP64 0   /* FP64 Gauss-Jordan elimination with partial pivoting */
#define RBFFD_PREC_MIXED 1  /* experimental: FP32 Gauss-Jordan inverse + FP64 refinement, FP64 re-solve
                               where it does not converge; built for the specialised shapes with
                               n + P <= 64, every other shape solves in FP64 */
#define RBFFD_PREC_AUTO 2   /* the faster of the two on this device: FP64 on MI355X, where FP32
                               issues at the FP64 rate without packed FP32 */

/* value types for rbffd_spmv_device */
#define RBFFD_F64 0
#define RBFFD_F32 1

typedef struct rbffd_context rbffd_context;

/* Boundary index sets of generate_operator method 2 (src/generate_operator.jl:192-193,
 * src/calculateneighbors.jl:1-97).  Marker b owns bc_idx[bc_ptr[b] .. bc_ptr[b+1])
 * and ghost_idx[ghost_ptr[b] .. ghost_ptr[b+1]). */
typedef struct rbffd_boundary {
    int32_t index_base;        /* 0 or 1 (Julia) */
    int64_t n_in;
    const int64_t *in_idx;     /* X_idx_in */
    int32_t n_markers;
    const int64_t *bc_ptr;     /* n_markers + 1 offsets into bc_idx (0-based offsets) */
    const int64_t *bc_idx;     /* X_idx_bc[b] concatenated */
    const int64_t *ghost_ptr;  /* n_markers + 1 offsets into ghost_idx */
    const int64_t *ghost_idx;  /* X_idx_bc_g[b] concatenated */
} rbffd_boundary;

/* What one generation call did. */
typedef struct rbffd_stats {
    int64_t rows;          /* output rows (= number of Y nodes) */
    int64_t singular;      /* rows whose system had an exactly zero pivot */
    int64_t nonfinite;     /* rows with non-finite weights */
    int64_t refined_fallback; /* mixed precision: rows re-solved in FP64 */
    double ms_knn, ms_solve, ms_fallback;  /* device time per phase (HIP events) */
    int64_t lanes_per_stencil; /* layout of the FP64 solve kernel (chosen by the autotuner) */
    int64_t factorizations; /* local systems factored: one per task, where a task is one output row or
                               (Y distinct from X) the rows sharing an X node's stencil */
} rbffd_stats;

/* ---------------------------------------------------------------- context */
int rbffd_version(void);
const char *rbffd_last_error(void);
int rbffd_create(int device, rbffd_context **ctx);
int rbffd_destroy(rbffd_context *ctx);
/* Run on this HIP stream from now on (NULL: the context's own stream). */
int rbffd_set_stream(rbffd_context *ctx, void *hip_stream);
int rbffd_synchronize(rbffd_context *ctx);

/* ------------------------------------------------------ device plumbing
 * No reference counterpart (the reference has no device).  For callers without
 * their own GPU allocator (the Julia shim, the Python mirror): device memory,
 * copies on the context stream (synchronous), and HIP-event timing of the
 * context stream.  rbffd_timer_stop returns the milliseconds between the last
 * rbffd_timer_start and now. */
int rbffd_device_malloc(rbffd_context *ctx, int64_t bytes, void **ptr);
int rbffd_device_free(rbffd_context *ctx, void *ptr);
int rbffd_memcpy_h2d(rbffd_context *ctx, void *dst, const void *src, int64_t bytes);
int rbffd_memcpy_d2h(rbffd_context *ctx, void *dst, const void *src, int64_t bytes);
int rbffd_memcpy_d2d(rbffd_context *ctx, void *dst, const void *src, int64_t bytes);
int rbffd_memset_device(rbffd_context *ctx, void *dst, int32_t value, int64_t bytes);
int rbffd_timer_start(rbffd_context *ctx);
int rbffd_timer_stop(rbffd_context *ctx, float *ms);
/* Name of the device, compute units, and the gfx target (e.g. "gfx950"). */
int rbffd_device_info(rbffd_context *ctx, char *name, int32_t name_len, int32_t *compute_units,
                      char *arch, int32_t arch_len);

/* ---------------------------------------------------------------- kNN (K1)
 * Replaces KDTree(X) + knn(tree, Q, k, true) (src/generate_operator.jl:43-47,
 * src/hyperviscosity_operator.jl:54-58) and the masked searches of
 * calculateneighbors (src/calculateneighbors.jl:16-94).
 * idx: M x k nearest X nodes of every query (0-based, nearest first); d2: the
 * squared distances (may be NULL).  Order: ascending dx*dx + dy*dy, ties by the
 * lower index.  xgroup/qgroup (both or neither, may be NULL): a query with
 * qgroup < 0 sees all X nodes, one with qgroup = g >= 0 sees nodes with xgroup 0
 * or g.  Synchronises the stream (the failure flags are read back).
 * RBFFD_ERR_ARG if a query sees fewer than k nodes; RBFFD_ERR_NONFINITE for
 * non-finite points. */
int rbffd_knn_device(rbffd_context *ctx, const double *X, int64_t N, const double *Q, int64_t M,
                     int32_t k, const int32_t *xgroup, const int32_t *qgroup, int32_t *idx, double *d2);

/* ------------------------------------------------ local weights (K2)
 * One output row per Y node k: the weights of the stencil of X node y2x[k]
 * (y2x NULL: Y is X and y2x[k] = k).  Replaces hot loops 1 and 2 of
 * generate_operator (src/generate_operator.jl:55-166, method 3 :388-467) and
 * hyperviscosity_operator (src/hyperviscosity_operator.jl:66-161).
 * Any stencil size 1 <= n <= 128 and polynomial degree with n + P <= 512
 * (P = (polydeg+1)(polydeg+2)/2); shapes without a specialised kernel run a
 * generic one.  With y2x, the rows that share an X node's stencil are solved
 * against one factorisation of it, as the reference factors each X node once
 * (src/generate_operator.jl:55-65,158); RBFFD_GROUP_ROWS=0 turns this off.
 * method: 1 (also method 2, same numerics) or 3.  ops/opk: nops operator kinds
 * and their derivative order k (used by XK, YK, LAPK).
 * Outputs: CSR rows of length n with sorted columns: colidx (M x n, may be
 * NULL), vals[o] (M x n, FP64) for each operator, status (M, may be NULL:
 * 0 ok, 1 singular, 2 non-finite, 3 refinement did not converge).
 * Checks stencils and y2x against N on the device first (RBFFD_ERR_ARG for an
 * index out of range), which synchronises; the solves are then asynchronous. */
int rbffd_weights_device(rbffd_context *ctx, const double *X, int64_t N, const double *Y, int64_t M,
                         const int32_t *stencils, int32_t n, const int32_t *y2x, int32_t p,
                         int32_t polydeg, int32_t method, int32_t nops, const int32_t *ops,
                         const int32_t *opk, int32_t precision, int32_t *colidx, double *const *vals,
                         int32_t *status);

/* rbffd_weights_device without allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, kernels and results as rbffd_weights_device, except that every
 * device buffer the call needs beyond its inputs and outputs (index checks,
 * the grouping of Y rows by stencil, the workspace of systems above 128
 * equations) comes from one fixed 64 MiB array in each GPU's memory, held
 * until the process ends; no layout autotuning runs.  RBFFD_ERR_ARG when a
 * call needs more than the array holds (grouping takes about 20 bytes per X
 * node and 16 per Y row).  Every call on a GPU shares that GPU's array: calls
 * that overlap on one GPU (two contexts, two threads) race.  Synchronises for
 * the index checks, like rbffd_weights_device. */
int rbffd_weights_device_noalloc(rbffd_context *ctx, const double *X, int64_t N, const double *Y, int64_t M,
                                 const int32_t *stencils, int32_t n, const int32_t *y2x, int32_t p,
                                 int32_t polydeg, int32_t method, int32_t nops, const int32_t *ops,
                                 const int32_t *opk, int32_t precision, int32_t *colidx, double *const *vals,
                                 int32_t *status);

/* --------------------------------------------- whole generation pipeline
 * kNN + nearest-X of every Y + weights, i.e. generate_operator /
 * hyperviscosity_operator up to (not including) the CSC assembly.
 * Y NULL selects method 3 (generate_operator(X, p, n, polydeg),
 * src/generate_operator.jl:354).  bnd non-NULL selects method 2.
 * stencils_out (N x n) and y2x_out (M) may be NULL.  Mixed precision re-solves
 * unconverged rows in FP64.  Synchronises; fills stats (may be NULL).
 * Returns RBFFD_ERR_SINGULAR / RBFFD_ERR_NONFINITE when rows failed (outputs are
 * still written). */
int rbffd_generate_device(rbffd_context *ctx, const double *X, int64_t N, const double *Y, int64_t M,
                          int32_t p, int32_t n, int32_t polydeg, const rbffd_boundary *bnd,
                          int32_t nops, const int32_t *ops, const int32_t *opk, int32_t precision,
                          int32_t *colidx, double *const *vals, int32_t *stencils_out,
                          int32_t *y2x_out, rbffd_stats *stats);

/* --------------------------------------------------------- CSR -> CSC (K3)
 * The SparseMatrixCSC layout of sparse(rows, cols, vals) (src/generate_operator.jl:171-182):
 * colptr (ncols + 1), rowval (M n, rows ascending inside a column), csc_vals[o]
 * (M n) for each of nops value arrays sharing the pattern.  Explicit zeros are
 * kept.  Indices offset by index_base (1 for Julia).  ncols_used (may be NULL)
 * receives 1 + the largest column referenced, the column count sparse() infers.
 * Synchronises. */
int rbffd_csr_to_csc_device(rbffd_context *ctx, int64_t M, int32_t n, int64_t ncols,
                            const int32_t *colidx, int32_t nops, const double *const *csr_vals,
                            int64_t *colptr, int64_t *rowval, double *const *csc_vals,
                            int32_t index_base, int64_t *ncols_used);

/* ----------------------------------------------------------- SpMV/SpMM (K4)
 * y = alpha A x + beta y for F fields (1, 2, 4, 8) stored node-major
 * (x[col F + f]); A is CSR with int32 columns and FP64 or FP32 values
 * (accumulated in FP64).  uniform_n > 0: every row has uniform_n entries and
 * rowptr may be NULL.  row_order (may be NULL): the rows run in this order
 * (a permutation of 0..M-1, e.g. from rbffd_locality_order_device).  Replaces SparseMatrixCSC * Vector in the reference's
 * harnesses (test/poisson_test.jl:121, examples/adv_diff_test.jl:151-152).
 * Asynchronous. */
int rbffd_spmv_device(rbffd_context *ctx, int64_t M, const int64_t *rowptr, const int32_t *col,
                      const void *vals, int32_t vals_type, int32_t F, const double *x, double *y,
                      double alpha, double beta, int32_t uniform_n, const int32_t *row_order);

/* *out = the sum of x[0 .. n) (device pointers, x 16-byte aligned): a
 * read-only stream at the device's HBM read rate, the ceiling SpMV (whose
 * traffic is almost all reads) is measured against.  No reference
 * counterpart: a measurement aid.  The last bits of the sum depend on the
 * order of the blocks' atomic adds.  Asynchronous. */
int rbffd_read_sum_device(rbffd_context *ctx, const double *x, int64_t n, double *out);

/* An application plan for repeated products with one operator.  The plan
 * copies the CSR into blocks of 256 consecutive rows of row_order (NULL: the
 * natural order; a locality order from rbffd_locality_order_device keeps each
 * block a compact patch), lists each block's distinct columns and refers to
 * them by 16-bit local index.  A product then loads a block's x rows into LDS
 * with coalesced loads and reads them there per nonzero, instead of gathering
 * every nonzero's x row from memory.  Arguments as rbffd_spmv_device; results
 * are bitwise equal to it (default lane layout).  The plan takes about 10 B
 * per nonzero with FP64 values (6 with FP32), plus 4 more when a block has
 * over 2048 distinct columns, plus 8 KiB per block, and keeps no pointer to
 * the caller's arrays.  No reference counterpart: the reference's harnesses
 * apply one operator many times (examples/adv_diff_test.jl:151-152), and the
 * plan moves the cost of that reuse to one set-up.  Creation synchronises;
 * products are asynchronous.  info (may be NULL) receives the row blocks,
 * the largest listed halo, the blocks over 2048 columns and the nonzeros. */
typedef struct rbffd_spmv_plan rbffd_spmv_plan;
int rbffd_spmv_plan_create(rbffd_context *ctx, int64_t M, int64_t N, const int64_t *rowptr, const int32_t *col,
                           const void *vals, int32_t vals_type, int32_t uniform_n, const int32_t *row_order,
                           rbffd_spmv_plan **plan, int64_t *info);
int rbffd_spmv_plan_apply(rbffd_context *ctx, const rbffd_spmv_plan *plan, int32_t F, const double *x, double *y,
                          double alpha, double beta);
int rbffd_spmv_plan_destroy(rbffd_spmv_plan *plan);

/* A locality order of N nodes (Morton order of their coordinates): with it as
 * row_order, the rows of one workgroup are a compact patch of the domain and
 * share most of the x values they gather.  Results are identical with or
 * without an order.  No reference counterpart.  Asynchronous. */
int rbffd_locality_order_device(rbffd_context *ctx, const double *X, int64_t N, int32_t *order);

/* ---------------------------------------------------- operator algebra
 * out = sum_o coef[o] vals[o] over nnz entries of operators that share one
 * pattern (the operators of one generation call do), e.g. a fused
 * advection-diffusion operator from Dx, Dy, Lap and Lap^k.  vals: host array
 * of nops device pointers; coef: host array.  out (FP64) and out32 (its FP32
 * rounding) may each be NULL.  Replaces the SparseMatrixCSC sums of the
 * harness (examples/adv_diff_test.jl:101,151-152).  Asynchronous. */
int rbffd_combine_device(rbffd_context *ctx, int64_t nnz, int32_t nops, const double *const *vals,
                         const double *coef, double *out, float *out32);

/* ------------------------------------------------------ sparse solve
 * x = argmin ||D x - b||_2 for a sparse D (M x N, M >= N; M = N solves the
 * square system), CSR with int64 rowptr, int32 columns and FP64 values, all
 * device pointers.  Replaces `D \ f` of the reference's PDE harness
 * (test/poisson_test.jl:121, test/mesh_import_test.jl:150) on the
 * collocation matrix assembled at test/poisson_test.jl:75-113, where Julia
 * runs a sparse QR (M > N) or LU (M = N) factorisation.
 *   RBFFD_SOLVE_QR:   D densified on the device and solved by rocSOLVER's QR
 *                     least squares (dgels; librocsolver.so.0 is opened at
 *                     run time, RBFFD_ERR_UNSUPPORTED without it).  Needs
 *                     M N < 2^31; a rank-deficient D gives RBFFD_ERR_SINGULAR.
 *                     The start value of x is not used.
 *   RBFFD_SOLVE_CGLS: conjugate gradients on the column-equilibrated normal
 *                     equations, every iteration on the device; x holds the
 *                     start (zeros for none).  Stops once
 *                     ||C D^T (b - D x)|| <= tol ||C D^T (b - D x0)|| (C the
 *                     column scaling) or after max_iters iterations.
 *   RBFFD_SOLVE_AUTO: QR when the dense copy has at most 2^28 entries and
 *                     rocSOLVER loads, CGLS otherwise.
 * *iters and *relres (each may be NULL) receive the CGLS iteration count and
 * the ratio above (0 and 0 for QR).  Needs M, N, nnz < 2^31.  Synchronises. */
#define RBFFD_SOLVE_AUTO 0
#define RBFFD_SOLVE_CGLS 1
#define RBFFD_SOLVE_QR 2
int rbffd_lsq_solve_device(rbffd_context *ctx, int64_t M, int64_t N, const int64_t *rowptr, const int32_t *col,
                           const double *vals, const double *b, double *x, int32_t method, double tol,
                           int64_t max_iters, int64_t *iters, double *relres);

/* ----------------------------------------------------- time stepping
 * `steps` steps of the three-stage strong-stability-preserving Runge-Kutta
 * method (Shu-Osher form) for du/dt = L u on F fields stored node-major, each
 * stage one fused pass y = a u + b w + c (L w) over the CSR operator L (N x N,
 * FP64 or FP32 values, FP64 accumulation).  Rows with fixed[i] != 0 keep their
 * value (Dirichlet rows; fixed may be NULL).  row_order as in
 * rbffd_spmv_device.  u (N F) is advanced in place;
 * work holds 2 N F doubles.  Replaces the explicit time integration of the
 * harness (examples/adv_diff_test.jl:190-198, an adaptive SSPRK43 there).
 * Asynchronous. */
int rbffd_ssprk3_device(rbffd_context *ctx, int64_t N, const int64_t *rowptr, const int32_t *col, const void *vals,
                        int32_t vals_type, int32_t uniform_n, const int32_t *row_order, int32_t F, double *u,
                        double *work, const uint8_t *fixed, double dt, int64_t steps);

/* rbffd_ssprk3_device over an application plan of L (N x N): the stages
 * stage each block's rows of the stage input in LDS as rbffd_spmv_plan_apply
 * does.  Results are bitwise equal to rbffd_ssprk3_device with the plan's
 * operator and row order.  Asynchronous. */
int rbffd_ssprk3_plan_device(rbffd_context *ctx, const rbffd_spmv_plan *plan, int32_t F, double *u, double *work,
                             const uint8_t *fixed, double dt, int64_t steps);

/* Ghost-node closure of one boundary marker (examples/adv_diff_test.jl:107-146,
 * 168-182): the marker's ghost values are set so that its boundary rows of the
 * normal-derivative operator D (Dx on left/right markers, Dy on top/bottom)
 * vanish,
 *     u[ghost] = -inv(D[bc, ghost]) * (D[bc, not ghost] * u).
 * All pointers are device pointers. */
typedef struct rbffd_ghost_closure {
    int64_t nbc;            /* boundary nodes of the marker = its ghost nodes */
    const int64_t *rowptr;  /* nbc + 1: CSR of D[bc, :] without this marker's ghost columns (w_bc_int) */
    const int32_t *col;     /* global node indices */
    const double *vals;
    const double *ginv;     /* nbc x nbc column-major inv(D[bc, ghost]) (w_bc_inv) */
    const int32_t *ghost;   /* the marker's ghost nodes, in the column order of D[bc, ghost] */
} rbffd_ghost_closure;

/* rbffd_ssprk3_device with ghost-node closures: the closures run in the
 * given order on u before the first step and on the output of every stage,
 * so every stage reads consistent ghost values (the reference refreshes them
 * inside each right-hand-side evaluation, examples/adv_diff_test.jl:168-182).
 * closures: host array of nclosures entries.  Sums run in a fixed order:
 * results are bitwise reproducible.  Asynchronous after the argument checks
 * (which read back a few flags). */
int rbffd_ssprk3_closure_device(rbffd_context *ctx, int64_t N, const int64_t *rowptr, const int32_t *col,
                                const void *vals, int32_t vals_type, int32_t uniform_n, const int32_t *row_order,
                                int32_t F, double *u, double *work, const uint8_t *fixed, int32_t nclosures,
                                const rbffd_ghost_closure *closures, double dt, int64_t steps);

/* Ainv = inv(A) for a dense n x n column-major device matrix, by rocSOLVER's
 * QR solve of A X = I (librocsolver.so.0 opened at run time;
 * RBFFD_ERR_UNSUPPORTED without it).  Replaces `inv(Array(w_bc_g[m]))` of the
 * ghost closure (examples/adv_diff_test.jl:118,124,130,136).  A singular A
 * gives RBFFD_ERR_SINGULAR.  Needs n^2 < 2^31.  Synchronises. */
int rbffd_dense_inverse_device(rbffd_context *ctx, int64_t n, const double *A, double *Ainv);

/* ------------------------------------------------------ uneven batch
 * Many independent node sets, each with its own stencil size, polynomial
 * degree and PHS power, generated in one call (Y = X for every set; method 1,
 * or method 2 where xgroup/qgroup are given).  No reference counterpart as one
 * call: it batches what a caller does with one generate_operator /
 * hyperviscosity_operator per set (src/generate_operator.jl:29,192;
 * src/hyperviscosity_operator.jl:26,177).  Every pointer in a set is a device
 * pointer; outputs are the CSR rows of rbffd_weights_device. */
typedef struct rbffd_batch_set {
    const double *X;                /* N x 2 points */
    int64_t N;
    int32_t p, n, polydeg;          /* PHS power, stencil size, polynomial degree */
    const int32_t *xgroup, *qgroup; /* N each, or NULL (see rbffd_boundary_groups) */
    int32_t *colidx;                /* N x n, columns sorted per row */
    double *vals[RBFFD_MAX_OPS];    /* N x n per operator */
    int32_t *status;                /* N, or NULL */
} rbffd_batch_set;

/* Host: the group ids of method 2 for N nodes (xgroup, qgroup: N each), the
 * visibility rule of calculateneighbors (src/calculateneighbors.jl:16-42,83-94).
 * RBFFD_ERR_UNDEFINED if a node is in no index set. */
int rbffd_boundary_groups(const rbffd_boundary *bnd, int64_t N, int32_t *xgroup, int32_t *qgroup);

/* Generate every set.  The sets are spread over parallel branches (kNN and
 * solves per set).  use_graph != 0: the first call with a given argument list
 * runs eagerly (the layout autotuner runs there) and records the whole batch
 * as a hipGraph; later calls with identical arguments (same pointers and
 * sizes) replay the graph.  Synchronises.  set_status (host, nsets, may be
 * NULL): 0 ok, 1 singular system, 2 non-finite input or weights, 3 a query sees
 * fewer than n nodes.  Returns the error of the first failing set (outputs of
 * the other sets are valid).  stats: rows, singular, nonfinite, and ms_solve =
 * device time of the whole batch. */
int rbffd_generate_batch_device(rbffd_context *ctx, int32_t nsets, const rbffd_batch_set *sets, int32_t nops,
                                const int32_t *ops, const int32_t *opk, int32_t use_graph, int32_t *set_status,
                                rbffd_stats *stats);

/* ------------------------------------------------ host entry point (Julia)
 * generate_operator / hyperviscosity_operator with host arrays, returning the
 * SparseMatrixCSC{Float64,Int64} arrays of every requested operator
 * (src/generate_operator.jl:29,192,354; src/hyperviscosity_operator.jl:26,177,314).
 * Y NULL: method 3.  bnd non-NULL: method 2.  Outputs (caller-allocated, sizes
 * known before the call): colptr (N + 1), rowval (M n), nzval[o] (M n); all
 * operators share colptr/rowval.  index_base 1 gives Julia's 1-based arrays.
 * *ncols receives the column count sparse() infers (max referenced column);
 * the row count is M.  Synchronous. */
int rbffd_generate_operator(rbffd_context *ctx, const double *X, int64_t N, const double *Y, int64_t M,
                            int32_t p, int32_t n, int32_t polydeg, const rbffd_boundary *bnd,
                            int32_t nops, const int32_t *ops, const int32_t *opk, int32_t precision,
                            int32_t index_base, int64_t *colptr, int64_t *rowval, double *const *nzval,
                            int64_t *ncols, rbffd_stats *stats);

#ifdef __cplusplus
}
#endif

#endif /* RBFFD_H */
