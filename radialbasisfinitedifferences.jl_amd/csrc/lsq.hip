// lsq.hip -- device kernels of the sparse solve D x = b behind `D \ f` in the
// reference's PDE harness (test/poisson_test.jl:75-121,
// test/mesh_import_test.jl:150): a least-squares solve when D has more rows
// than columns (oversampled collocation, Y != X), a plain solve when square.
//
// The solver is CGLS (conjugate gradients on the normal equations,
// Hestenes-Stiefel form) on the column-equilibrated matrix D~ = D C,
// C = diag(1 / ||D e_j||), with x = x0 + C z.  One iteration is four kernels:
//   1. q = D~ p                     and |q|^2
//   2. z += a p, r -= a q           a = gamma / |q|^2
//   3. s = D~^T r                   and gamma' = |s|^2
//   4. p = s + (gamma' / gamma) p
// D~^T is stored as a CSR of its own, so both products are row gathers with
// no atomics.  Squared norms use a two-level sum in a fixed order (block sums,
// then the last block to finish adds them in block order), so a solve is
// bitwise reproducible.  The scalars stay in device memory: no iteration
// waits for the host.
//
// For systems whose dense copy fits the device, the direct method writes D
// into a column-major dense array and calls rocSOLVER's QR least-squares
// routine (gels), which gives `\`'s answer to rounding for both shapes.
// rocSOLVER is opened at run time, so the library keeps linking only the HIP
// runtime; without it the direct method reports that it is unsupported.
#include <dlfcn.h>

#include <mutex>

#include <rocblas/rocblas.h>      // types and prototypes only
#include <rocsolver/rocsolver.h>  // (the symbols come from dlsym)

#include "device_common.hpp"
#include "internal.hpp"

namespace rbffd {

__device__ __forceinline__ double wave_sum_f64(double v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Sum of v over the grid (256 threads a block).  Each block stores its sum in
// partial[blockIdx.x]; the last block to arrive adds the block sums in block
// order, writes the total to *out and re-arms the counter.
__device__ void grid_sum_finish(double v, const LsqReduce &r)
{
    __shared__ double wsum[4];
    __shared__ int last;
    const int w = threadIdx.x >> 6;
    v = wave_sum_f64(v);
    if ((threadIdx.x & 63) == 0) wsum[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        r.partial[blockIdx.x] = (wsum[0] + wsum[1]) + (wsum[2] + wsum[3]);
        __threadfence();
        const unsigned t = atomicAdd(r.count, 1u);
        last = t == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    double s = 0.0;
    for (unsigned i = threadIdx.x; i < gridDim.x; i += blockDim.x)
        s += __hip_atomic_load(&r.partial[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s = wave_sum_f64(s);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) wsum[w] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        *r.out = (wsum[0] + wsum[1]) + (wsum[2] + wsum[3]);
        atomicExch(r.count, 0u);
    }
}

// y = A x over CSR rows, G lanes a row, and |y|^2 into *red.out.  Waves walk
// the rows in passes of 64 / G rows (a wave-uniform loop, so the lane-group
// sums always run with the whole wave).
template <int G>
__global__ __launch_bounds__(256) void lsq_spmv_kernel(long M, const int64_t *__restrict__ rowptr,
                                                       const int *__restrict__ col, const double *__restrict__ val,
                                                       const double *__restrict__ x, double *__restrict__ y,
                                                       LsqReduce red)
{
    constexpr int RPW = 64 / G;
    const int lane = threadIdx.x & 63, sub = lane % G;
    const long wave = (blockIdx.x * (long)blockDim.x + threadIdx.x) >> 6;
    const long stride = (long)gridDim.x * (blockDim.x >> 6) * RPW;
    double sq = 0.0;
    for (long base = wave * RPW; base < M; base += stride) {
        const long row = base + lane / G;
        const bool live = row < M;
        long p0 = 0, p1 = 0;
        if (live) { p0 = rowptr[row]; p1 = rowptr[row + 1]; }
        double acc = 0.0;
        for (long p = p0 + sub; p < p1; p += G) acc = __builtin_fma(val[p], x[col[p]], acc);
        acc = group_sum_f64<G>(acc);
        if (live && sub == 0) {
            y[row] = acc;
            sq = __builtin_fma(acc, acc, sq);
        }
    }
    grid_sum_finish(sq, red);
}

static int lsq_lanes(long rows, long nnz) { return (rows > 0 && nnz / rows < 12) ? 8 : 16; }

static unsigned lsq_spmv_blocks(long M, int G)
{
    const long b = (M * G + 255) / 256;
    return (unsigned)std::max<long>(1, std::min<long>(kLsqBlocks, b));
}

void launch_lsq_spmv(long M, long nnz, const int64_t *rowptr, const int *col, const double *val, const double *x,
                     double *y, const LsqReduce &red, hipStream_t s)
{
    // M == 0 still launches one block so that *red.out is written (0)
    if (lsq_lanes(M, nnz) == 8)
        lsq_spmv_kernel<8><<<lsq_spmv_blocks(M, 8), 256, 0, s>>>(M, rowptr, col, val, x, y, red);
    else
        lsq_spmv_kernel<16><<<lsq_spmv_blocks(M, 16), 256, 0, s>>>(M, rowptr, col, val, x, y, red);
}

static unsigned blocks_of(long n) { return (unsigned)std::max<long>(1, std::min<long>(65535L * 16, (n + 255) / 256)); }

// z += a p (N entries), r -= a q (M entries), a = gamma / |q|^2
__global__ void lsq_update_kernel(long N, long M, const double *__restrict__ gam, const double *__restrict__ qq,
                                  const double *__restrict__ p, const double *__restrict__ q, double *__restrict__ z,
                                  double *__restrict__ r)
{
    const double a = *qq > 0.0 ? *gam / *qq : 0.0;
    const long n = N > M ? N : M;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        if (i < N) z[i] = __builtin_fma(a, p[i], z[i]);
        if (i < M) r[i] = __builtin_fma(-a, q[i], r[i]);
    }
}

void launch_lsq_update(long N, long M, const double *gam, const double *qq, const double *p, const double *q,
                       double *z, double *r, hipStream_t s)
{
    lsq_update_kernel<<<blocks_of(std::max(N, M)), 256, 0, s>>>(N, M, gam, qq, p, q, z, r);
}

// p = s + (gamma' / gamma) p
__global__ void lsq_direction_kernel(long N, const double *__restrict__ gam, const double *__restrict__ gam_new,
                                     const double *__restrict__ sv, double *__restrict__ p)
{
    const double b = *gam > 0.0 ? *gam_new / *gam : 0.0;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x)
        p[i] = __builtin_fma(b, p[i], sv[i]);
}

void launch_lsq_direction(long N, const double *gam, const double *gam_new, const double *sv, double *p,
                          hipStream_t s)
{
    if (N > 0) lsq_direction_kernel<<<blocks_of(N), 256, 0, s>>>(N, gam, gam_new, sv, p);
}

// ---- set-up and finish
// row_of[p] = row of CSR entry p
__global__ void row_of_kernel(const int64_t *rowptr, long M, int *row_of)
{
    const long r = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (r >= M) return;
    for (long p = rowptr[r]; p < rowptr[r + 1]; ++p) row_of[p] = (int)r;
}

void launch_row_of(const int64_t *rowptr, long M, int *row_of, hipStream_t s)
{
    if (M > 0) row_of_kernel<<<(unsigned)((M + 255) / 256), 256, 0, s>>>(rowptr, M, row_of);
}

// dst[i] = src[idx[i]]
__global__ void gather_i32_kernel(const int *src, const int *idx, long n, int *dst)
{
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[idx[i]];
}

void launch_gather_i32(const int *src, const int *idx, long n, int *dst, hipStream_t s)
{
    if (n > 0) gather_i32_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(src, idx, n, dst);
}

// scale[j] = 1 / ||row j of the transpose|| (0 for an empty or all-zero column)
__global__ void col_scale_kernel(const int64_t *tptr, long N, const double *tval, double *scale)
{
    const long j = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (j >= N) return;
    double s = 0.0;
    for (long p = tptr[j]; p < tptr[j + 1]; ++p) s = __builtin_fma(tval[p], tval[p], s);
    scale[j] = s > 0.0 ? 1.0 / __builtin_sqrt(s) : 0.0;
}

void launch_col_scale(const int64_t *tptr, long N, const double *tval, double *scale, hipStream_t s)
{
    if (N > 0) col_scale_kernel<<<(unsigned)((N + 255) / 256), 256, 0, s>>>(tptr, N, tval, scale);
}

// out[p] = v[p] * scale[col[p]]
__global__ void scale_cols_kernel(const int *col, long nnz, const double *scale, const double *v, double *out)
{
    const long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (p < nnz) out[p] = v[p] * scale[col[p]];
}

void launch_scale_cols(const int *col, long nnz, const double *scale, const double *v, double *out, hipStream_t s)
{
    if (nnz > 0) scale_cols_kernel<<<(unsigned)((nnz + 255) / 256), 256, 0, s>>>(col, nnz, scale, v, out);
}

// y = x (n entries)
__global__ void copy_f64_kernel(const double *x, long n, double *y)
{
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) y[i] = x[i];
}

void launch_copy_f64(const double *x, long n, double *y, hipStream_t s)
{
    if (n > 0) copy_f64_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(x, n, y);
}

// x[j] += scale[j] * z[j]
__global__ void lsq_finish_kernel(long N, const double *scale, const double *z, double *x)
{
    const long j = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (j < N) x[j] = __builtin_fma(scale[j], z[j], x[j]);
}

void launch_lsq_finish(long N, const double *scale, const double *z, double *x, hipStream_t s)
{
    if (N > 0) lsq_finish_kernel<<<(unsigned)((N + 255) / 256), 256, 0, s>>>(N, scale, z, x);
}

__global__ void fill_f64_kernel(double *p, long n, double v)
{
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

void launch_fill_f64(double *p, long n, double v, hipStream_t s)
{
    if (n > 0) fill_f64_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(p, n, v);
}

// ---- direct method
// A (column-major, leading dimension M) += the entries of CSR row r
__global__ void densify_kernel(long M, const int64_t *rowptr, const int *col, const double *val, double *A)
{
    const long r = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (r >= M) return;
    for (long p = rowptr[r]; p < rowptr[r + 1]; ++p) A[(long)col[p] * M + r] += val[p];
}

namespace {

struct SolverLib {
    decltype(&rocblas_create_handle) create = nullptr;
    decltype(&rocblas_destroy_handle) destroy = nullptr;
    decltype(&rocblas_set_stream) set_stream = nullptr;
    decltype(&rocsolver_dgels) dgels = nullptr;
};

const SolverLib *solver_lib()
{
    static std::once_flag once;
    static SolverLib lib;
    std::call_once(once, [] {
        void *hb = dlopen("librocblas.so.5", RTLD_NOW | RTLD_LOCAL);
        if (!hb) hb = dlopen("/opt/rocm/lib/librocblas.so.5", RTLD_NOW | RTLD_LOCAL);
        void *hs = dlopen("librocsolver.so.0", RTLD_NOW | RTLD_LOCAL);
        if (!hs) hs = dlopen("/opt/rocm/lib/librocsolver.so.0", RTLD_NOW | RTLD_LOCAL);
        if (!hb || !hs) return;
        SolverLib l;
        l.create = (decltype(l.create))dlsym(hb, "rocblas_create_handle");
        l.destroy = (decltype(l.destroy))dlsym(hb, "rocblas_destroy_handle");
        l.set_stream = (decltype(l.set_stream))dlsym(hb, "rocblas_set_stream");
        l.dgels = (decltype(l.dgels))dlsym(hs, "rocsolver_dgels");
        if (l.create && l.destroy && l.set_stream && l.dgels) lib = l;
    });
    return lib.dgels ? &lib : nullptr;
}

}  // namespace

bool dense_solver_available() { return solver_lib() != nullptr; }

void dense_solver_destroy(void *handle)
{
    const SolverLib *L = solver_lib();
    if (L && handle) (void)L->destroy((rocblas_handle)handle);
}

int dense_lsq(void **handle, long M, long N, const int64_t *rowptr, const int *col, const double *vals,
              const double *b, double *x, double *A, double *B, int *info, hipStream_t s)
{
    const SolverLib *L = solver_lib();
    if (!L) throw Error{2, "solve: the direct method needs rocSOLVER (librocsolver.so.0), which did not load"};
    if (!*handle) {
        rocblas_handle h = nullptr;
        if (L->create(&h) != rocblas_status_success) throw Error{3, "solve: rocblas_create_handle failed"};
        *handle = h;
    }
    rocblas_handle h = (rocblas_handle)*handle;
    if (L->set_stream(h, s) != rocblas_status_success) throw Error{3, "solve: rocblas_set_stream failed"};
    launch_fill_f64(A, M * N, 0.0, s);
    densify_kernel<<<(unsigned)((M + 255) / 256), 256, 0, s>>>(M, rowptr, col, vals, A);
    launch_copy_f64(b, M, B, s);
    RBFFD_HIP(hipGetLastError());
    const rocblas_status st = L->dgels(h, rocblas_operation_none, (rocblas_int)M, (rocblas_int)N, 1, A,
                                       (rocblas_int)M, B, (rocblas_int)M, info);
    if (st != rocblas_status_success) throw Error{3, "solve: rocsolver_dgels failed (status " + std::to_string((int)st) + ")"};
    launch_copy_f64(B, N, x, s);
    int hinfo = 0;
    RBFFD_HIP(hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, s));
    RBFFD_HIP(hipStreamSynchronize(s));
    return hinfo;
}

// A X = I by QR: X = inv(A)
__global__ void identity_kernel(long n, double *X)
{
    const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (e < n * n) X[e] = (e % (n + 1) == 0) ? 1.0 : 0.0;
}

int dense_inverse(void **handle, long n, double *Aw, double *Ainv, int *info, hipStream_t s)
{
    const SolverLib *L = solver_lib();
    if (!L) throw Error{2, "inverse: needs rocSOLVER (librocsolver.so.0), which did not load"};
    if (!*handle) {
        rocblas_handle h = nullptr;
        if (L->create(&h) != rocblas_status_success) throw Error{3, "inverse: rocblas_create_handle failed"};
        *handle = h;
    }
    rocblas_handle h = (rocblas_handle)*handle;
    if (L->set_stream(h, s) != rocblas_status_success) throw Error{3, "inverse: rocblas_set_stream failed"};
    identity_kernel<<<(unsigned)((n * n + 255) / 256), 256, 0, s>>>(n, Ainv);
    RBFFD_HIP(hipGetLastError());
    const rocblas_status st = L->dgels(h, rocblas_operation_none, (rocblas_int)n, (rocblas_int)n, (rocblas_int)n, Aw,
                                       (rocblas_int)n, Ainv, (rocblas_int)n, info);
    if (st != rocblas_status_success) throw Error{3, "inverse: rocsolver_dgels failed (status " + std::to_string((int)st) + ")"};
    int hinfo = 0;
    RBFFD_HIP(hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, s));
    RBFFD_HIP(hipStreamSynchronize(s));
    return hinfo;
}

}  // namespace rbffd
