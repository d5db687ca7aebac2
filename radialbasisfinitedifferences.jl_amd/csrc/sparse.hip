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
// sums are reduced across lanes by DPP.  Fields are stored node-major
// (x[col*F + f]) so one gather brings all F fields of a node.
#include <algorithm>
#include <cstdlib>

#include <hipcub/hipcub.hpp>

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

// Entries of each column, in arbitrary order (atomic slots).
__global__ void csc_scatter_kernel(const int *colidx, long nnz, const int *colptr, int *fill, int *perm)
{
    long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (p >= nnz) return;
    int c = colidx[p];
    int pos = colptr[c] + atomicAdd(&fill[c], 1);
    perm[pos] = (int)p;
}

void launch_csc_scatter(const int *colidx, long M, int n, const int *colptr, int *fill, int *perm, hipStream_t s)
{
    long nnz = M * (long)n;
    if (nnz <= 0) return;
    csc_scatter_kernel<<<(unsigned)((nnz + 255) / 256), 256, 0, s>>>(colidx, nnz, colptr, fill, perm);
}

// Final order inside each column: an entry's slot is the number of entries of
// its column with a smaller entry index (entry p lies in row p / n, so rows
// come out ascending).  One thread per entry counts over its column's
// unsorted segment; columns hold a few dozen entries.
__global__ void csc_rank_kernel(const int *colidx, long nnz, int n, const int *colptr, const int *perm_in,
                                int *perm, int *rowval)
{
    const long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (p >= nnz) return;
    const int c = colidx[p];
    const int b = colptr[c], e = colptr[c + 1];
    int rank = 0;
    for (int i = b; i < e; ++i) rank += perm_in[i] < (int)p ? 1 : 0;
    perm[b + rank] = (int)p;
    rowval[b + rank] = (int)(p / n);
}

void launch_csc_rank(const int *colidx, long M, int n, const int *colptr, const int *perm_in, int *perm,
                     int *rowval, hipStream_t s)
{
    const long nnz = M * (long)n;
    if (nnz <= 0) return;
    csc_rank_kernel<<<(unsigned)((nnz + 255) / 256), 256, 0, s>>>(colidx, nnz, n, colptr, perm_in, perm, rowval);
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

__global__ void fill_i64_kernel(int64_t *p, long n, int64_t v)
{
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

void launch_fill_i64(int64_t *p, long n, int64_t v, hipStream_t s)
{
    if (n <= 0) return;
    fill_i64_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(p, n, v);
}

// ------------------------------------------------------------------ SpMV/SpMM
template <typename TV, int G, int F, bool UNIFORM>
__global__ __launch_bounds__(256) void spmv_kernel(long M, const int64_t *__restrict__ rowptr,
                                                   const int *__restrict__ col, const TV *__restrict__ val,
                                                   int n_uniform, const int *__restrict__ order,
                                                   const double *__restrict__ x, double *__restrict__ y, double alpha,
                                                   double beta)
{
    const long gid = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const int sub = (int)(gid % G);
    // rows past the end still take part in the shuffles below
    const bool live = gid / G < M;
    // rows run in `order` (a locality order) when one is given
    const long row = !live ? M : (order ? (long)order[gid / G] : gid / G);
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
    for (int f = 0; f < F; ++f) acc[f] = group_sum_f64<G>(acc[f]);
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

// SpMM with the fields split over lanes: 16 lanes per row, FL lanes share a
// nonzero (FL divides F) and take F/FL contiguous fields each, so a nonzero's
// x row is one contiguous run across them; the 16/FL nonzero phases are
// summed by DPP row rotations, which leave the row's total in the phase-0
// lanes.
template <typename TV, int F, int FL, bool UNIFORM>
__global__ __launch_bounds__(256) void spmm_split_kernel(long M, const int64_t *__restrict__ rowptr,
                                                         const int *__restrict__ col, const TV *__restrict__ val,
                                                         int n_uniform, const int *__restrict__ order,
                                                         const double *__restrict__ x, double *__restrict__ y,
                                                         double alpha, double beta)
{
    constexpr int G = 16, FPL = F / FL, NZP = G / FL;
    static_assert(F % FL == 0 && FL <= 8, "field split");
    const long gid = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const int sub = (int)(gid % G);
    const int fq = sub % FL, nzp = sub / FL;
    // rows past the end still take part in the rotations below
    const bool live = gid / G < M;
    const long row = !live ? M : (order ? (long)order[gid / G] : gid / G);
    long p0 = 0, p1 = 0;
    if (live) {
        if (UNIFORM) { p0 = row * (long)n_uniform; p1 = p0 + n_uniform; }
        else { p0 = rowptr[row]; p1 = rowptr[row + 1]; }
    }
    double acc[FPL];
#pragma unroll
    for (int f = 0; f < FPL; ++f) acc[f] = 0.0;
    for (long p = p0 + nzp; p < p1; p += NZP) {
        const double v = (double)val[p];
        const double *xr = x + (long)col[p] * F + fq * FPL;
#pragma unroll
        for (int f = 0; f < FPL; ++f) acc[f] = __builtin_fma(v, xr[f], acc[f]);
    }
#pragma unroll
    for (int f = 0; f < FPL; ++f) {
        if constexpr (NZP >= 8) acc[f] += dpp_f64<0x120 | (FL)>(acc[f]);      // row_ror
        if constexpr (NZP >= 4) acc[f] += dpp_f64<0x120 | (G / 4)>(acc[f]);
        if constexpr (NZP >= 2) acc[f] += dpp_f64<0x120 | (G / 2)>(acc[f]);
    }
    if (live && nzp == 0) {
#pragma unroll
        for (int f = 0; f < FPL; ++f) {
            const long o = row * F + fq * FPL + f;
            double r = alpha * acc[f];
            if (beta != 0.0) r = __builtin_fma(beta, y[o], r);
            y[o] = r;
        }
    }
}

template <typename TV, int F, int FL>
static void spmm_split_launch_t(long M, const int64_t *rowptr, const int *col, const TV *val, int nu,
                                const int *order, const double *x, double *y, double alpha, double beta,
                                hipStream_t s)
{
    const long threads = M * 16;
    const unsigned blocks = (unsigned)((threads + 255) / 256);
    if (nu > 0)
        spmm_split_kernel<TV, F, FL, true><<<blocks, 256, 0, s>>>(M, rowptr, col, val, nu, order, x, y, alpha, beta);
    else
        spmm_split_kernel<TV, F, FL, false><<<blocks, 256, 0, s>>>(M, rowptr, col, val, nu, order, x, y, alpha, beta);
}

template <typename TV, int G, int F>
static void spmv_launch_t(long M, const int64_t *rowptr, const int *col, const TV *val, int nu, const int *order,
                          const double *x, double *y, double alpha, double beta, hipStream_t s)
{
    const long threads = M * G;
    const unsigned blocks = (unsigned)((threads + 255) / 256);
    if (nu > 0)
        spmv_kernel<TV, G, F, true><<<blocks, 256, 0, s>>>(M, rowptr, col, val, nu, order, x, y, alpha, beta);
    else
        spmv_kernel<TV, G, F, false><<<blocks, 256, 0, s>>>(M, rowptr, col, val, nu, order, x, y, alpha, beta);
}

// Lanes per row: 16 (one field or several), or RBFFD_SPMV_LANES /
// RBFFD_SPMM_LANES = 4 | 8 | 16 | 32.
static int lanes_env(const char *name, int dflt)
{
    const char *e = std::getenv(name);
    const int v = e ? std::atoi(e) : dflt;
    return (v == 4 || v == 8 || v == 16 || v == 32) ? v : dflt;
}

template <typename TV, int F>
static void spmv_lanes_t(int G, long M, const int64_t *rowptr, const int *col, const TV *val, int nu,
                         const int *order, const double *x, double *y, double alpha, double beta, hipStream_t s)
{
    switch (G) {
    case 4: spmv_launch_t<TV, 4, F>(M, rowptr, col, val, nu, order, x, y, alpha, beta, s); break;
    case 8: spmv_launch_t<TV, 8, F>(M, rowptr, col, val, nu, order, x, y, alpha, beta, s); break;
    case 32: spmv_launch_t<TV, 32, F>(M, rowptr, col, val, nu, order, x, y, alpha, beta, s); break;
    default: spmv_launch_t<TV, 16, F>(M, rowptr, col, val, nu, order, x, y, alpha, beta, s); break;
    }
}

template <typename TV>
static void spmv_dispatch(long M, const int64_t *rowptr, const int *col, const TV *val, int F, int nu,
                          const int *order, const double *x, double *y, double alpha, double beta, hipStream_t s)
{
    static const int g1 = lanes_env("RBFFD_SPMV_LANES", 16);
    static const int gf = lanes_env("RBFFD_SPMM_LANES", 16);
    // lanes sharing a nonzero: RBFFD_SPMM_FL = 1 | 2 | 4 | 8; by default 2
    // from 4 fields up (8 fields over 1M nodes, n = 31: 189 us against 212 us with
    // every field in every lane; 4 and 8 lanes were slower)
    static const int fl_env = [] {
        const char *e = std::getenv("RBFFD_SPMM_FL");
        const int v = e ? std::atoi(e) : 0;
        return (v == 1 || v == 2 || v == 4 || v == 8) ? v : 0;
    }();
    const int fl = fl_env ? fl_env : (F >= 4 ? 2 : 1);
    if (F >= 2 && gf == 16 && fl > 1 && fl <= F) {
        auto go = [&](auto fc, auto flc) {
            spmm_split_launch_t<TV, decltype(fc)::value, decltype(flc)::value>(M, rowptr, col, val, nu, order, x, y,
                                                                               alpha, beta, s);
        };
        using I2 = std::integral_constant<int, 2>;
        using I4 = std::integral_constant<int, 4>;
        using I8 = std::integral_constant<int, 8>;
        if (F == 2) go(I2{}, I2{});
        else if (F == 4) { if (fl == 2) go(I4{}, I2{}); else go(I4{}, I4{}); }
        else if (F == 8) { if (fl == 2) go(I8{}, I2{}); else if (fl == 4) go(I8{}, I4{}); else go(I8{}, I8{}); }
        else throw Error{2, "spmv: fields must be 1, 2, 4 or 8"};
        return;
    }
    switch (F) {
    case 1: spmv_lanes_t<TV, 1>(g1, M, rowptr, col, val, nu, order, x, y, alpha, beta, s); break;
    case 2: spmv_lanes_t<TV, 2>(gf, M, rowptr, col, val, nu, order, x, y, alpha, beta, s); break;
    case 4: spmv_lanes_t<TV, 4>(gf, M, rowptr, col, val, nu, order, x, y, alpha, beta, s); break;
    case 8: spmv_lanes_t<TV, 8>(gf, M, rowptr, col, val, nu, order, x, y, alpha, beta, s); break;
    default: throw Error{2, "spmv: fields must be 1, 2, 4 or 8"};
    }
}

void launch_spmv(long M, const int64_t *rowptr, const int *col, const void *vals, int vals_f32, int F,
                 const double *x, double *y, double alpha, double beta, int uniform_n, const int *order,
                 hipStream_t s)
{
    if (M <= 0) return;
    if (vals_f32)
        spmv_dispatch<float>(M, rowptr, col, (const float *)vals, F, uniform_n, order, x, y, alpha, beta, s);
    else
        spmv_dispatch<double>(M, rowptr, col, (const double *)vals, F, uniform_n, order, x, y, alpha, beta, s);
}

// ---- locality order: rows sorted by the Morton code of their node
// (coordinates quantised to 16 bits over the bounding box)
__device__ __forceinline__ uint32_t spread_bits(uint32_t v)
{
    v &= 0xFFFFu;
    v = (v | (v << 8)) & 0x00FF00FFu;
    v = (v | (v << 4)) & 0x0F0F0F0Fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}

__global__ void morton_kernel(const double *X, long N, const double *box, uint32_t *key, int *idx)
{
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double w = box[2] - box[0], h = box[3] - box[1];
    const double fx = w > 0 ? (X[2 * i] - box[0]) / w : 0.0;
    const double fy = h > 0 ? (X[2 * i + 1] - box[1]) / h : 0.0;
    const uint32_t qx = (uint32_t)fmin(fmax(fx * 65535.0, 0.0), 65535.0);
    const uint32_t qy = (uint32_t)fmin(fmax(fy * 65535.0, 0.0), 65535.0);
    key[i] = spread_bits(qx) | (spread_bits(qy) << 1);
    idx[i] = (int)i;
}

size_t morton_tmp_bytes(long N)
{
    size_t bytes = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                             (const int *)nullptr, (int *)nullptr, (int)N, 0, 32);
    return bytes;
}

void launch_morton_order(const double *X, long N, const double *box, uint32_t *key, uint32_t *key_sorted,
                         int *idx, int *order, void *tmp, size_t tmp_bytes, hipStream_t s)
{
    if (N <= 0) return;
    morton_kernel<<<(unsigned)((N + 255) / 256), 256, 0, s>>>(X, N, box, key, idx);
    RBFFD_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, key, key_sorted, idx, order, (int)N, 0, 32, s));
}

// ------------------------------------------------------------ time stepping
// One Shu-Osher stage: out = a u + b w + c (L w) for F fields, L in CSR with
// G lanes per row (as spmv_kernel).  Rows with fixed[r] != 0 keep u.  out may
// be u (each value is read and written by the same lane); out must not be w.
template <typename TV, int G, int F, int FL, bool UNIFORM>
__global__ __launch_bounds__(256) void rk_stage_kernel(long M, const int64_t *__restrict__ rowptr,
                                                       const int *__restrict__ col, const TV *__restrict__ val,
                                                       int n_uniform, const int *__restrict__ order,
                                                       const double *__restrict__ w, const double *u, double *out,
                                                       const uint8_t *__restrict__ fixed, double a, double b, double c)
{
    // FL > 1: the fields split over FL lanes per nonzero, as spmm_split_kernel
    constexpr int FPL = F / FL, NZP = G / FL;
    static_assert(F % FL == 0 && (FL == 1 || G == 16), "field split");
    const long gid = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const int sub = (int)(gid % G);
    const int fq = sub % FL, nzp = sub / FL;
    const bool live = gid / G < M;
    const long row = !live ? M : (order ? (long)order[gid / G] : gid / G);
    long p0 = 0, p1 = 0;
    if (live) {
        if (UNIFORM) { p0 = row * (long)n_uniform; p1 = p0 + n_uniform; }
        else { p0 = rowptr[row]; p1 = rowptr[row + 1]; }
    }
    double acc[FPL];
#pragma unroll
    for (int f = 0; f < FPL; ++f) acc[f] = 0.0;
    for (long p = p0 + nzp; p < p1; p += NZP) {
        const double v = (double)val[p];
        const long cidx = (long)col[p] * F + fq * FPL;
#pragma unroll
        for (int f = 0; f < FPL; ++f) acc[f] = __builtin_fma(v, w[cidx + f], acc[f]);
    }
    if constexpr (FL == 1) {
#pragma unroll
        for (int f = 0; f < F; ++f) acc[f] = group_sum_f64<G>(acc[f]);
    } else {
#pragma unroll
        for (int f = 0; f < FPL; ++f) {
            if constexpr (NZP >= 8) acc[f] += dpp_f64<0x120 | (FL)>(acc[f]);  // row_ror
            if constexpr (NZP >= 4) acc[f] += dpp_f64<0x120 | (G / 4)>(acc[f]);
            if constexpr (NZP >= 2) acc[f] += dpp_f64<0x120 | (G / 2)>(acc[f]);
        }
    }
    if (live) {
        const bool hold = fixed && fixed[row];
        if constexpr (FL == 1) {
#pragma unroll
            for (int f = 0; f < F; ++f) {
                if (f % G == sub) {
                    const long i = row * F + f;
                    const double ui = u[i];
                    out[i] = hold ? ui : a * ui + b * w[i] + c * acc[f];
                }
            }
        } else if (nzp == 0) {
#pragma unroll
            for (int f = 0; f < FPL; ++f) {
                const long i = row * F + fq * FPL + f;
                const double ui = u[i];
                out[i] = hold ? ui : a * ui + b * w[i] + c * acc[f];
            }
        }
    }
}

template <typename TV, int F>
static void rk_stage_t(long M, const int64_t *rowptr, const int *col, const TV *val, int nu, const int *order,
                       const double *w, const double *u, double *out, const uint8_t *fixed, double a, double b,
                       double c, hipStream_t s)
{
    constexpr int G = 16;
    // two lanes per nonzero from 4 fields up, as the SpMM
    constexpr int FL = F >= 4 ? 2 : 1;
    const unsigned blocks = (unsigned)((M * G + 255) / 256);
    if (nu > 0)
        rk_stage_kernel<TV, G, F, FL, true><<<blocks, 256, 0, s>>>(M, rowptr, col, val, nu, order, w, u, out, fixed,
                                                                   a, b, c);
    else
        rk_stage_kernel<TV, G, F, FL, false><<<blocks, 256, 0, s>>>(M, rowptr, col, val, nu, order, w, u, out,
                                                                    fixed, a, b, c);
}

template <typename TV>
static void rk_stage_f(int F, long M, const int64_t *rowptr, const int *col, const TV *val, int nu,
                       const int *order, const double *w, const double *u, double *out, const uint8_t *fixed,
                       double a, double b, double c, hipStream_t s)
{
    switch (F) {
    case 1: rk_stage_t<TV, 1>(M, rowptr, col, val, nu, order, w, u, out, fixed, a, b, c, s); break;
    case 2: rk_stage_t<TV, 2>(M, rowptr, col, val, nu, order, w, u, out, fixed, a, b, c, s); break;
    case 4: rk_stage_t<TV, 4>(M, rowptr, col, val, nu, order, w, u, out, fixed, a, b, c, s); break;
    case 8: rk_stage_t<TV, 8>(M, rowptr, col, val, nu, order, w, u, out, fixed, a, b, c, s); break;
    default: throw Error{2, "time stepping: fields must be 1, 2, 4 or 8"};
    }
}

void launch_rk_stage(long M, const int64_t *rowptr, const int *col, const void *vals, int vals_f32, int F,
                     int uniform_n, const int *order, const double *w, const double *u, double *out,
                     const uint8_t *fixed, double a, double b, double c, hipStream_t s)
{
    if (M <= 0) return;
    if (vals_f32)
        rk_stage_f<float>(F, M, rowptr, col, (const float *)vals, uniform_n, order, w, u, out, fixed, a, b, c, s);
    else
        rk_stage_f<double>(F, M, rowptr, col, (const double *)vals, uniform_n, order, w, u, out, fixed, a, b, c,
                           s);
}

// out = sum_o coef[o] vals[o] (FP64), and its FP32 rounding when out32 is set.
__global__ void combine_kernel(long nnz, CombineTerms t, double *out, float *out32)
{
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= nnz) return;
    double r = 0.0;
    for (int o = 0; o < t.nops; ++o) r = __builtin_fma(t.coef[o], t.vals[o][i], r);
    if (out) out[i] = r;
    if (out32) out32[i] = (float)r;
}

void launch_combine(long nnz, const CombineTerms &t, double *out, float *out32, hipStream_t s)
{
    if (nnz <= 0) return;
    combine_kernel<<<(unsigned)((nnz + 255) / 256), 256, 0, s>>>(nnz, t, out, out32);
}

// ---- ghost-node closure (examples/adv_diff_test.jl:168-182):
// u[ghost] = -G t with G = inv(D[bc, ghost]) dense column-major (nbc x nbc)
// and t = D[bc, not ghost] u (nbc x F).  The product runs in chunks of
// kClosureChunk columns of G; the chunk sums are added in chunk order, so the
// result does not depend on scheduling.
constexpr int kClosureChunk = 256;

long closure_chunks(long nbc) { return (nbc + kClosureChunk - 1) / kClosureChunk; }

template <int F>
__global__ __launch_bounds__(256) void closure_gemv_kernel(long nbc, const double *__restrict__ G,
                                                           const double *__restrict__ t, double *__restrict__ part)
{
    __shared__ double ts[kClosureChunk * F];
    const long j0 = (long)blockIdx.y * kClosureChunk;
    const int nj = (int)((nbc - j0) < kClosureChunk ? (nbc - j0) : kClosureChunk);
    for (int e = threadIdx.x; e < nj * F; e += blockDim.x) ts[e] = t[j0 * F + e];
    __syncthreads();
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= nbc) return;
    double acc[F];
#pragma unroll
    for (int f = 0; f < F; ++f) acc[f] = 0.0;
    for (int j = 0; j < nj; ++j) {
        const double g = G[(j0 + j) * nbc + i];
#pragma unroll
        for (int f = 0; f < F; ++f) acc[f] = __builtin_fma(g, ts[j * F + f], acc[f]);
    }
    double *o = part + ((long)blockIdx.y * nbc + i) * F;
#pragma unroll
    for (int f = 0; f < F; ++f) o[f] = acc[f];
}

__global__ void closure_scatter_kernel(long nbc, int F, long nchunks, const double *__restrict__ part,
                                       const int *__restrict__ ghost, double *__restrict__ u)
{
    const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (e >= nbc * F) return;
    double s = 0.0;
    for (long c = 0; c < nchunks; ++c) s += part[c * nbc * F + e];
    const long i = e / F;
    u[(long)ghost[i] * F + (e - i * F)] = -s;
}

void launch_closure(long nbc, int F, const double *G, const double *t, double *part, const int *ghost, double *u,
                    hipStream_t s)
{
    if (nbc <= 0) return;
    const long nc = closure_chunks(nbc);
    dim3 grid((unsigned)((nbc + 255) / 256), (unsigned)nc);
    switch (F) {
    case 1: closure_gemv_kernel<1><<<grid, 256, 0, s>>>(nbc, G, t, part); break;
    case 2: closure_gemv_kernel<2><<<grid, 256, 0, s>>>(nbc, G, t, part); break;
    case 4: closure_gemv_kernel<4><<<grid, 256, 0, s>>>(nbc, G, t, part); break;
    case 8: closure_gemv_kernel<8><<<grid, 256, 0, s>>>(nbc, G, t, part); break;
    default: throw Error{2, "closure: fields must be 1, 2, 4 or 8"};
    }
    closure_scatter_kernel<<<(unsigned)((nbc * F + 255) / 256), 256, 0, s>>>(nbc, F, nc, part, ghost, u);
}

// ------------------------------------------------------------ read ceiling
// *out = sum of x[0 .. n): a read-only stream (16-byte loads, four in flight
// per thread, one atomic per block) that measures the HBM read rate SpMV's
// traffic (almost all reads) can reach.  *out is cleared first; the sum's last
// bits depend on the order of the blocks' atomics.
__global__ void clear_f64_kernel(double *out) { *out = 0.0; }

__global__ __launch_bounds__(256) void read_sum_kernel(const double2 *x, long n2, const double *tail, int ntail,
                                                        double *out)
{
    double acc = 0.0;
    const long stride = (long)gridDim.x * blockDim.x;
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    for (; i + 3 * stride < n2; i += 4 * stride) {
        const double2 a = x[i], b = x[i + stride], c = x[i + 2 * stride], d = x[i + 3 * stride];
        acc += (a.x + a.y) + (b.x + b.y) + (c.x + c.y) + (d.x + d.y);
    }
    for (; i < n2; i += stride) acc += x[i].x + x[i].y;
    if (blockIdx.x == 0 && (int)threadIdx.x < ntail) acc += tail[threadIdx.x];
    __shared__ double part[4];
    for (int o = 32; o >= 1; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(out, (part[0] + part[1]) + (part[2] + part[3]));
}

void launch_read_sum(const double *x, long n, double *out, int cus, hipStream_t s)
{
    clear_f64_kernel<<<1, 1, 0, s>>>(out);
    if (n <= 0) return;
    const long n2 = n / 2;
    const unsigned blocks = (unsigned)std::max(1L, std::min<long>((long)cus * 8, (n2 + 255) / 256));
    read_sum_kernel<<<blocks, 256, 0, s>>>((const double2 *)x, n2, x + 2 * n2, (int)(n - 2 * n2), out);
}

}  // namespace rbffd
