// sparse.hip -- K3 (CSR -> CSC for the SparseMatrixCSC the reference returns)
// and K4 (operator application, CSR SpMV / multi-field SpMM).
//
// K3 restates the output of sparse(I, J, V) (src/generate_operator.jl:171-182):
// column pointers from a counting pass, entries scattered by column, each
// column's rows put in ascending order, every operator's values gathered
// through one permutation (all operators share one pattern).
//
// K4 replaces the SparseMatrixCSC * Vector products of the reference's
// harnesses (test/poisson_test.jl:121, examples/adv_diff_test.jl:151-152).
// A group of G lanes owns one row: the row's values and columns are read as
// contiguous runs (coalesced), x is gathered per column, and the G partial
// sums are reduced across lanes.  Fields are stored node-major (x[col*F + f])
// so one gather brings all F fields of a node.
#include "device_common.hpp"
#include "internal.hpp"

namespace rbffd {

__global__ void fill_rowptr_kernel(int64_t *rowptr, long M, int n)
{
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i <= M) rowptr[i] = i * (int64_t)n;
}

void launch_fill_rowptr(int64_t *rowptr, long M, int n, hipStream_t s)
{
    fill_rowptr_kernel<<<(unsigned)((M + 1 + 255) / 256), 256, 0, s>>>(rowptr, M, n);
}

__global__ void csc_count_kernel(const int *colidx, long nnz, int *counts)
{
    long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (p < nnz) atomicAdd(&counts[colidx[p]], 1);
}

void launch_csc_count(const int *colidx, long nnz, int *counts, hipStream_t s)
{
    if (nnz <= 0) return;
    csc_count_kernel<<<(unsigned)((nnz + 255) / 256), 256, 0, s>>>(colidx, nnz, counts);
}

__global__ void csc_scatter_kernel(const int *colidx, long M, int n, const int *colptr, int *fill,
                                   int *perm, int *rowval)
{
    long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (p >= M * (long)n) return;
    int c = colidx[p];
    int pos = colptr[c] + atomicAdd(&fill[c], 1);
    perm[pos] = (int)p;
    rowval[pos] = (int)(p / n);
}

void launch_csc_scatter(const int *colidx, long M, int n, const int *colptr, int *fill, int *perm,
                        int *rowval, hipStream_t s)
{
    long nnz = M * (long)n;
    if (nnz <= 0) return;
    csc_scatter_kernel<<<(unsigned)((nnz + 255) / 256), 256, 0, s>>>(colidx, M, n, colptr, fill, perm,
                                                                      rowval);
}

// One thread per column: insertion sort of its (row, perm) pairs.  Columns hold
// a few dozen entries, rows inside a column are distinct.
__global__ void csc_sort_kernel(const int *colptr, long ncols, int *rowval, int *perm)
{
    long c = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (c >= ncols) return;
    const int b = colptr[c], e = colptr[c + 1];
    for (int i = b + 1; i < e; ++i) {
        const int r = rowval[i], q = perm[i];
        int j = i - 1;
        while (j >= b && rowval[j] > r) {
            rowval[j + 1] = rowval[j];
            perm[j + 1] = perm[j];
            --j;
        }
        rowval[j + 1] = r;
        perm[j + 1] = q;
    }
}

void launch_csc_sort_columns(const int *colptr, long ncols, int *rowval, int *perm, hipStream_t s)
{
    if (ncols <= 0) return;
    csc_sort_kernel<<<(unsigned)((ncols + 255) / 256), 256, 0, s>>>(colptr, ncols, rowval, perm);
}

__global__ void gather_f64_kernel(const double *src, const int *perm, long nnz, double *dst)
{
    long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (p < nnz) dst[p] = src[perm[p]];
}

void launch_gather_f64(const double *src, const int *perm, long nnz, double *dst, hipStream_t s)
{
    if (nnz <= 0) return;
    gather_f64_kernel<<<(unsigned)((nnz + 255) / 256), 256, 0, s>>>(src, perm, nnz, dst);
}

__global__ void i32_to_i64_kernel(const int *src, long n, int64_t *dst, int64_t add)
{
    long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (p < n) dst[p] = (int64_t)src[p] + add;
}

void launch_i32_to_i64(const int *src, long n, int64_t *dst, int64_t add, hipStream_t s)
{
    if (n <= 0) return;
    i32_to_i64_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(src, n, dst, add);
}

__global__ void f64_to_f32_kernel(const double *src, long n, float *dst)
{
    long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (p < n) dst[p] = (float)src[p];
}

void launch_f64_to_f32(const double *src, long n, float *dst, hipStream_t s)
{
    if (n <= 0) return;
    f64_to_f32_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(src, n, dst);
}

// ------------------------------------------------------------------ SpMV/SpMM
template <typename TV, int G, int F, bool UNIFORM>
__global__ __launch_bounds__(256) void spmv_kernel(long M, const int64_t *__restrict__ rowptr,
                                                   const int *__restrict__ col, const TV *__restrict__ val,
                                                   int n_uniform, const double *__restrict__ x,
                                                   double *__restrict__ y, double alpha, double beta)
{
    const long gid = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const long row = gid / G;
    const int sub = (int)(gid % G);
    // rows past the end still take part in the shuffles below
    const bool live = row < M;
    long p0 = 0, p1 = 0;
    if (live) {
        if (UNIFORM) { p0 = row * (long)n_uniform; p1 = p0 + n_uniform; }
        else { p0 = rowptr[row]; p1 = rowptr[row + 1]; }
    }
    double acc[F];
#pragma unroll
    for (int f = 0; f < F; ++f) acc[f] = 0.0;
    for (long p = p0 + sub; p < p1; p += G) {
        const double v = (double)val[p];
        const long cidx = (long)col[p] * F;
#pragma unroll
        for (int f = 0; f < F; ++f) acc[f] = __builtin_fma(v, x[cidx + f], acc[f]);
    }
#pragma unroll
    for (int o = G / 2; o >= 1; o >>= 1) {
#pragma unroll
        for (int f = 0; f < F; ++f) acc[f] += __shfl_xor(acc[f], o, G);
    }
    if (live) {
        // the G lanes of a row share the F stores
#pragma unroll
        for (int f = 0; f < F; ++f) {
            if (f % G == sub) {
                double r = alpha * acc[f];
                if (beta != 0.0) r = __builtin_fma(beta, y[row * F + f], r);
                y[row * F + f] = r;
            }
        }
    }
}

template <typename TV, int G, int F>
static void spmv_launch_t(long M, const int64_t *rowptr, const int *col, const TV *val, int nu,
                          const double *x, double *y, double alpha, double beta, hipStream_t s)
{
    const long threads = M * G;
    const unsigned blocks = (unsigned)((threads + 255) / 256);
    if (nu > 0)
        spmv_kernel<TV, G, F, true><<<blocks, 256, 0, s>>>(M, rowptr, col, val, nu, x, y, alpha, beta);
    else
        spmv_kernel<TV, G, F, false><<<blocks, 256, 0, s>>>(M, rowptr, col, val, nu, x, y, alpha, beta);
}

template <typename TV>
static void spmv_dispatch(long M, const int64_t *rowptr, const int *col, const TV *val, int F, int nu,
                          const double *x, double *y, double alpha, double beta, hipStream_t s)
{
    switch (F) {
    case 1: spmv_launch_t<TV, 16, 1>(M, rowptr, col, val, nu, x, y, alpha, beta, s); break;
    case 2: spmv_launch_t<TV, 16, 2>(M, rowptr, col, val, nu, x, y, alpha, beta, s); break;
    case 4: spmv_launch_t<TV, 16, 4>(M, rowptr, col, val, nu, x, y, alpha, beta, s); break;
    case 8: spmv_launch_t<TV, 16, 8>(M, rowptr, col, val, nu, x, y, alpha, beta, s); break;
    default: throw Error{2, "spmv: fields must be 1, 2, 4 or 8"};
    }
}

void launch_spmv(long M, const int64_t *rowptr, const int *col, const void *vals, int vals_f32, int F,
                 const double *x, double *y, double alpha, double beta, int uniform_n, hipStream_t s)
{
    if (M <= 0) return;
    if (vals_f32)
        spmv_dispatch<float>(M, rowptr, col, (const float *)vals, F, uniform_n, x, y, alpha, beta, s);
    else
        spmv_dispatch<double>(M, rowptr, col, (const double *)vals, F, uniform_n, x, y, alpha, beta, s);
}

}  // namespace rbffd
