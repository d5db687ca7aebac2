/*
 * rbffd_oracle.c -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * A CPU restatement, in plain C, of the RBF-FD operator generation of
 * RadialBasisFiniteDifferences.jl (the reference is Julia and cannot run here;
 * it ships no golden weights).  Used only by tests/, by the smoke check of
 * __graft_entry__.py and by the CPU-baseline leg of bench.py, always as the
 * checker or the baseline, never as the measured or shipped path.
 *
 * Contents
 *   orc_monomials     monomial exponents in the basis order    src/polynomialbasis.jl:8-9
 *   orc_phs_deriv     closed-form PHS derivatives               src/rbfbasis.jl, src/rbfbasis_k.jl
 *   orc_knn_brute     exact kNN by brute force, (d^2, index) order   NearestNeighbors knn as called
 *                     at src/generate_operator.jl:43-47, src/calculateneighbors.jl:16-94
 *   orc_knn_grid      the same result set from a uniform-grid search (pinned against brute force)
 *   orc_weights       per-stencil weights, FP64 or __float128      src/generate_operator.jl,
 *                     src/hyperviscosity_operator.jl (methods 1/2 and 3)
 *   orc_spmv_csr      y = A x for a CSR matrix (FP64, row order)
 *
 * Build: see oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off).
 */
#include <math.h>
#include <quadmath.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

enum {
    ORC_OP_E = 0, ORC_OP_DX = 1, ORC_OP_DY = 2, ORC_OP_DXX = 3, ORC_OP_DYY = 4,
    ORC_OP_DXY = 5, ORC_OP_XK = 6, ORC_OP_YK = 7, ORC_OP_LAP = 8, ORC_OP_LAPK = 9
};

static double fact(int n)
{
    double r = 1;
    for (int i = 2; i <= n; ++i) r *= i;
    return r;
}
static double binom(int n, int k) { return fact(n) / (fact(k) * fact(n - k)); }

/* ---- FP64 instantiation ---------------------------------------------------*/
#define REAL double
#define RS(name) name##_d
#define RSQRT sqrt
#define RFABS fabs
#include "oracle_core.h"
#undef REAL
#undef RS
#undef RSQRT
#undef RFABS

/* ---- __float128 instantiation (high-precision oracle) ----------------------*/
#define REAL __float128
#define RS(name) name##_q
#define RSQRT sqrtq
#define RFABS fabsq
#include "oracle_core.h"
#undef REAL
#undef RS
#undef RSQRT
#undef RFABS

/*
 * Monomial basis order.  The reference builds monomials([x, y], 0:polydeg) and
 * reverses them (src/polynomialbasis.jl:8-9).  The order of DynamicPolynomials
 * 0.5.2 cannot be checked here (the package is absent); the basis used is graded,
 * increasing total degree, and inside a degree increasing power of x:
 *   1, y, x, y^2, xy, x^2, y^3, ...
 * The stencil weights do not depend on this order (a column permutation of the
 * polynomial block only permutes the unused Lagrange multipliers).
 */
int orc_monomials(int polydeg, int *ea, int *eb)
{
    int q = 0;
    for (int t = 0; t <= polydeg; ++t)
        for (int a = 0; a <= t; ++a) {
            if (ea) ea[q] = a;
            if (eb) eb[q] = t - a;
            ++q;
        }
    return q;
}

double orc_phs_deriv(int m, int a, int b, double vx, double vy)
{
    return phs_deriv_d(m, a, b, vx, vy);
}

double orc_phs_deriv_q(int m, int a, int b, double vx, double vy)
{
    return (double)phs_deriv_q(m, a, b, (__float128)vx, (__float128)vy);
}

/* An operator applied to the PHS at v, per-axis scales (sx, sy) folded in for
 * the Laplacian-type operators. */
double orc_op_phs(int op, int k, int m, double vx, double vy, double sx, double sy)
{
    return op_phs_d(op, k, m, vx, vy, sx, sy);
}

/* ---------------------------------------------------------------------------
 * kNN.  Visibility (src/calculateneighbors.jl:16-42, 83-94): a query with
 * qgroup < 0 sees every X node; a query with qgroup = g >= 0 sees X nodes with
 * xgroup == 0 and with xgroup == g.  (Boundary marker b and its ghost nodes
 * query with g = b; ghost nodes of marker b carry xgroup = b; interior and
 * boundary nodes carry 0; nodes in no index set carry -2.)  xgroup/qgroup NULL
 * means everything is visible.
 * Order: ascending squared Euclidean distance dx*dx + dy*dy (no fused
 * multiply-add), ties broken by the lower index.
 * -------------------------------------------------------------------------*/
static inline int key_less(double d, int i, double d2, int i2)
{
    return d < d2 || (d == d2 && i < i2);
}

static inline void topk_insert(double *kd, int *ki, int k, double d, int id)
{
    if (!key_less(d, id, kd[k - 1], ki[k - 1])) return;
    int j = k - 1;
    while (j > 0 && key_less(d, id, kd[j - 1], ki[j - 1])) {
        kd[j] = kd[j - 1]; ki[j] = ki[j - 1]; --j;
    }
    kd[j] = d; ki[j] = id;
}

static inline int visible(const int *xg, const int *qg, long xi, long qi)
{
    if (!xg || !qg) return 1;
    int g = qg[qi];
    if (g < 0) return 1;
    int h = xg[xi];
    return h == 0 || h == g;
}

void orc_knn_brute(const double *X, long N, const double *Q, long M, int k,
                   const int *xgroup, const int *qgroup, int *idx, double *d2out, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
    {
        double *kd = malloc(sizeof(double) * k);
        int *ki = malloc(sizeof(int) * k);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 64)
#endif
        for (long q = 0; q < M; ++q) {
            for (int j = 0; j < k; ++j) { kd[j] = INFINITY; ki[j] = -1; }
            double qx = Q[2 * q], qy = Q[2 * q + 1];
            for (long i = 0; i < N; ++i) {
                if (!visible(xgroup, qgroup, i, q)) continue;
                double dx = X[2 * i] - qx, dy = X[2 * i + 1] - qy;
                double dd = dx * dx;
                double ee = dy * dy;
                double d = dd + ee;
                if (ki[k - 1] < 0 || key_less(d, (int)i, kd[k - 1], ki[k - 1]))
                    topk_insert(kd, ki, k, d, (int)i);
            }
            for (int j = 0; j < k; ++j) {
                idx[q * k + j] = ki[j];
                if (d2out) d2out[q * k + j] = kd[j];
            }
        }
        free(kd); free(ki);
    }
}

/* Exact kNN through a uniform grid: rings of cells around the query cell,
 * cells farther than the current k-th distance skipped, stop once the
 * unvisited region is provably farther.  Same result as orc_knn_brute. */
void orc_knn_grid(const double *X, long N, const double *Q, long M, int k,
                  const int *xgroup, const int *qgroup, int *idx, double *d2out, int nthreads)
{
    double x0 = INFINITY, y0 = INFINITY, x1 = -INFINITY, y1 = -INFINITY;
    for (long i = 0; i < N; ++i) {
        if (X[2 * i] < x0) x0 = X[2 * i];
        if (X[2 * i] > x1) x1 = X[2 * i];
        if (X[2 * i + 1] < y0) y0 = X[2 * i + 1];
        if (X[2 * i + 1] > y1) y1 = X[2 * i + 1];
    }
    double w = x1 - x0, h = y1 - y0;
    /* about 2 points per cell; an axis of zero extent gets one unit cell */
    double cs = (w > 0 && h > 0) ? sqrt(w * h * 2.0 / (double)N) : (w > h ? w : h) * 2.0 / (double)N;
    if (!(cs > 0)) cs = 1;
    long gx = w > 0 ? (long)(w / cs) + 1 : 1, gy = h > 0 ? (long)(h / cs) + 1 : 1;
    if (gx > 1 << 15) gx = 1 << 15;
    if (gy > 1 << 15) gy = 1 << 15;
    double csx = w > 0 ? w / (double)gx * (1 + 1e-12) : 1.0;
    double csy = h > 0 ? h / (double)gy * (1 + 1e-12) : 1.0;
    long ncell = gx * gy;
    long *start = calloc(ncell + 1, sizeof(long));
    int *cell_pts = malloc(sizeof(int) * (N > 0 ? N : 1));
    long *fill = calloc(ncell, sizeof(long));
    for (long i = 0; i < N; ++i) {
        long cx = (long)((X[2 * i] - x0) / csx), cy = (long)((X[2 * i + 1] - y0) / csy);
        if (cx >= gx) cx = gx - 1;
        if (cy >= gy) cy = gy - 1;
        start[cy * gx + cx + 1]++;
    }
    for (long c = 0; c < ncell; ++c) start[c + 1] += start[c];
    for (long i = 0; i < N; ++i) {
        long cx = (long)((X[2 * i] - x0) / csx), cy = (long)((X[2 * i + 1] - y0) / csy);
        if (cx >= gx) cx = gx - 1;
        if (cy >= gy) cy = gy - 1;
        long c = cy * gx + cx;
        cell_pts[start[c] + fill[c]++] = (int)i;
    }
    free(fill);
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
    {
        double *kd = malloc(sizeof(double) * k);
        int *ki = malloc(sizeof(int) * k);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 256)
#endif
        for (long q = 0; q < M; ++q) {
            for (int j = 0; j < k; ++j) { kd[j] = INFINITY; ki[j] = -1; }
            double qx = Q[2 * q], qy = Q[2 * q + 1];
            /* far-away queries are clamped next to the grid (the search then
             * covers every cell before it stops) */
            double fx = floor((qx - x0) / csx), fy = floor((qy - y0) / csy);
            long qcx = (long)fmin(fmax(fx, -2.0), (double)gx + 1.0);
            long qcy = (long)fmin(fmax(fy, -2.0), (double)gy + 1.0);
            long rmax = gx + gy + 4;
            for (long r = 0; r <= rmax; ++r) {
                for (long cy = qcy - r; cy <= qcy + r; ++cy) {
                    if (cy < 0 || cy >= gy) continue;
                    int edge_row = (cy == qcy - r || cy == qcy + r);
                    for (long cx = qcx - r; cx <= qcx + r; cx += (edge_row ? 1 : 2 * r)) {
                        if (cx >= 0 && cx < gx) {
                            /* distance from the query to this cell rectangle */
                            double rx0 = x0 + cx * csx, rx1 = rx0 + csx;
                            double ry0 = y0 + cy * csy, ry1 = ry0 + csy;
                            double ddx = qx < rx0 ? rx0 - qx : (qx > rx1 ? qx - rx1 : 0);
                            double ddy = qy < ry0 ? ry0 - qy : (qy > ry1 ? qy - ry1 : 0);
                            double md = (ddx * ddx + ddy * ddy) * (1 - 1e-9);
                            if (!(ki[k - 1] >= 0 && md > kd[k - 1])) {
                                long c = cy * gx + cx;
                                for (long p = start[c]; p < start[c + 1]; ++p) {
                                    int i = cell_pts[p];
                                    if (!visible(xgroup, qgroup, i, q)) continue;
                                    double dx = X[2 * i] - qx, dy = X[2 * i + 1] - qy;
                                    double dd = dx * dx;
                                    double ee = dy * dy;
                                    double d = dd + ee;
                                    if (ki[k - 1] < 0 || key_less(d, i, kd[k - 1], ki[k - 1]))
                                        topk_insert(kd, ki, k, d, i);
                                }
                            }
                        }
                        if (r == 0) break;
                    }
                }
                /* everything not yet visited lies outside the block of rings <= r */
                double bx0 = qx - (x0 + (qcx - r) * csx), bx1 = (x0 + (qcx + r + 1) * csx) - qx;
                double by0 = qy - (y0 + (qcy - r) * csy), by1 = (y0 + (qcy + r + 1) * csy) - qy;
                double b = bx0;
                if (bx1 < b) b = bx1;
                if (by0 < b) b = by0;
                if (by1 < b) b = by1;
                if (ki[k - 1] >= 0 && b > 0 && b * b * (1 - 1e-9) > kd[k - 1]) break;
                int covers = (qcx - r <= 0) && (qcy - r <= 0) && (qcx + r >= gx - 1) && (qcy + r >= gy - 1);
                if (covers) break;
            }
            for (int j = 0; j < k; ++j) {
                idx[q * k + j] = ki[j];
                if (d2out) d2out[q * k + j] = kd[j];
            }
        }
        free(kd); free(ki);
    }
    free(start); free(cell_pts);
}

/* ---------------------------------------------------------------------------
 * Weights for a set of Y rows.
 *   X (N x 2), Y (M x 2) interleaved Float64 (SVector{2,Float64} layout).
 *   stencils: N x n neighbour lists, 0-based, first entry = stencil centre.
 *   y2x: nearest X node of each Y row (NULL: collocated, y2x[k] = k).
 *   rows/nrows: the Y rows to compute (NULL: all M).
 *   method: 1 (also 2: same numerics, different neighbours) or 3.
 *   real_kind: 0 = double, 1 = __float128.
 *   solve_kind: 0 = inv(A) * RHS as the reference, 1 = LU solve.
 *   W: nrows x nops x n weights in stencil order.
 *   scale: nrows x 2 (may be NULL).  status: nrows (0 ok, 1 singular, 2 non-finite).
 * Returns the number of rows with a non-zero status.
 * -------------------------------------------------------------------------*/
long orc_weights(const double *X, long N, const double *Y, long M,
                 const int *stencils, int n, const int *y2x,
                 int m, int polydeg, int method,
                 int nops, const int *ops, const int *opk,
                 int real_kind, int solve_kind,
                 const long *rows, long nrows,
                 double *W, double *scale, int *status, int nthreads)
{
    (void)N;
    int ea[64], eb[64];
    int P = orc_monomials(polydeg, ea, eb);
    int ns = n + P;
    if (!rows) nrows = M;
    long nbad = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel reduction(+ : nbad)
#endif
    {
        size_t wsz = (size_t)(2 * ns * ns + 2 * ns * nops + 4 * n + 16);
        double *wd = real_kind ? NULL : malloc(sizeof(double) * wsz);
        __float128 *wq = real_kind ? malloc(sizeof(__float128) * wsz) : NULL;
        int *ipiv = malloc(sizeof(int) * ns);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 16)
#endif
        for (long t = 0; t < nrows; ++t) {
            long yk = rows ? rows[t] : t;
            long xc = y2x ? y2x[yk] : yk;
            const int *st = stencils + xc * (long)n;
            double *Wt = W + t * (long)nops * n;
            double *sc = scale ? scale + 2 * t : NULL;
            int s;
            if (real_kind)
                s = stencil_weights_q(X, Y, yk, st, n, m, P, ea, eb, method, nops, ops, opk,
                                      solve_kind, wq, ipiv, Wt, sc);
            else
                s = stencil_weights_d(X, Y, yk, st, n, m, P, ea, eb, method, nops, ops, opk,
                                      solve_kind, wd, ipiv, Wt, sc);
            if (status) status[t] = s;
            if (s) nbad++;
        }
        free(wd); free(wq); free(ipiv);
    }
    return nbad;
}

/* y = A x, CSR with 0-based int32 columns, F fields stored node-major (x[col*F + f]). */
void orc_spmv_csr(long nrows, const long *rowptr, const int *col, const double *val,
                  int F, const double *x, double *y, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
#endif
    for (long r = 0; r < nrows; ++r) {
        for (int f = 0; f < F; ++f) {
            double acc = 0;
            for (long p = rowptr[r]; p < rowptr[r + 1]; ++p) acc += val[p] * x[(long)col[p] * F + f];
            y[r * F + f] = acc;
        }
    }
}

int orc_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
