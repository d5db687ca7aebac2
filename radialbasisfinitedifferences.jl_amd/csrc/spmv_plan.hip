// spmv_plan.hip -- K4 for repeated products with one operator: an
// application plan that stages each workgroup's x rows in LDS.
//
// The plain kernels (sparse.hip) gather a node's F fields from HBM/L2 once
// per nonzero, so a product with 8 fields issues about 4 vector-memory loads
// per nonzero and is bound by the texture address path of the L1, not by HBM.
// A plan cuts the rows, in a locality order (Morton order of the rows' nodes
// for RBF-FD operators), into blocks of kPlanRows consecutive rows and lists
// each block's distinct columns (its halo, sorted).  Its nonzeros then refer to
// halo slots by 16-bit local index.  A product loads the block's halo rows of
// x into LDS with coalesced 16-byte loads (sorted halo indices are runs of
// consecutive nodes on structured node sets) and reads x from LDS per nonzero.
//
// Each row's nonzeros keep their order, and lanes take them and reduce across
// DPP exactly as spmv_kernel / spmm_split_kernel / rk_stage_kernel do, so a
// plan product is bitwise equal to the plain one with the default lane layout
// (16 lanes a row; two lanes a nonzero from 4 fields up).
//
// Replaces SparseMatrixCSC * Vector of the reference's harnesses
// (test/poisson_test.jl:121, examples/adv_diff_test.jl:151-152) where one
// operator is applied many times.
#include <algorithm>
#include <climits>
#include <cstdint>

#include <hipcub/hipcub.hpp>

#include "device_common.hpp"
#include "internal.hpp"

namespace rbffd {

namespace {

constexpr int kHashSlots = 2 * kPlanHaloCap;  // open addressing, load factor <= 1/2

__global__ void plan_rowlen_kernel(long M, const int64_t *__restrict__ rowptr, int nu, const int *__restrict__ order,
                                   int64_t *__restrict__ len)
{
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i > M) return;
    if (i == M) {
        len[i] = 0;
        return;
    }
    const long r = order ? (long)order[i] : i;
    len[i] = nu > 0 ? (int64_t)nu : rowptr[r + 1] - rowptr[r];
}

// pcol/pval of plan row i = the entries of row order[i], in their order;
// 16 lanes a row
template <typename TV>
__global__ void plan_copy_kernel(long M, const int64_t *__restrict__ rowptr, int nu, const int *__restrict__ order,
                                 const int *__restrict__ col, const TV *__restrict__ val,
                                 const int64_t *__restrict__ pptr, int *__restrict__ pcol, TV *__restrict__ pval)
{
    const long gid = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const long i = gid / 16;
    if (i >= M) return;
    const long r = order ? (long)order[i] : i;
    const long s0 = nu > 0 ? r * (long)nu : rowptr[r];
    const long d0 = pptr[i], len = pptr[i + 1] - d0;
    for (long e = gid % 16; e < len; e += 16) {
        pcol[d0 + e] = col[s0 + e];
        pval[d0 + e] = val[s0 + e];
    }
}

__device__ __forceinline__ unsigned plan_hash(int key)
{
    return ((unsigned)key * 2654435761u) >> (32 - 12);  // 12 bits = log2(kHashSlots)
}

// One workgroup per block of rows: the distinct columns in a hash set in LDS,
// sorted (bitonic, in LDS), then each nonzero's local index.  A block with
// more than kPlanHaloCap distinct columns only records its count; products
// read its x through pcol.
__global__ __launch_bounds__(256) void plan_halo_kernel(long M, const int64_t *__restrict__ pptr,
                                                        const int *__restrict__ pcol, uint16_t *__restrict__ plc,
                                                        int *__restrict__ halo, int *__restrict__ hcount)
{
    static_assert(kHashSlots == 4096, "plan_hash takes 12 bits");
    __shared__ int keys[kHashSlots];
    __shared__ int rank[kHashSlots];
    __shared__ int uniq[kPlanHaloCap];
    __shared__ int nuniq;
    const int tid = threadIdx.x;
    const long r0 = (long)blockIdx.x * kPlanRows;
    const long r1 = r0 + kPlanRows < M ? r0 + kPlanRows : M;
    const long p0 = pptr[r0], p1 = pptr[r1];
    for (int h = tid; h < kHashSlots; h += 256) keys[h] = -1;
    if (tid == 0) nuniq = 0;
    __syncthreads();
    // every probe sequence ends: at most kHashSlots probes, and a full table
    // only happens past kPlanHaloCap distinct keys, which is an overflow anyway
    for (long p = p0 + tid; p < p1; p += 256) {
        const int key = pcol[p];
        unsigned h = plan_hash(key);
        for (int probe = 0; probe < kHashSlots; ++probe) {
            const int old = atomicCAS(&keys[h], -1, key);
            if (old == -1) {
                const int u = atomicAdd(&nuniq, 1);
                if (u < kPlanHaloCap) uniq[u] = key;
                break;
            }
            if (old == key) break;
            h = (h + 1) & (kHashSlots - 1);
        }
    }
    __syncthreads();
    const int H = nuniq;
    if (tid == 0) hcount[blockIdx.x] = H;
    if (H > kPlanHaloCap) return;  // uniform: every lane read the same H
    // bitonic sort of uniq, padded to a power of two with INT_MAX
    int P2 = 1;
    while (P2 < H) P2 <<= 1;
    for (int i = H + tid; i < P2; i += 256) uniq[i] = INT_MAX;
    __syncthreads();
    for (int k = 2; k <= P2; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = tid; i < P2; i += 256) {
                const int l = i ^ j;
                if (l > i) {
                    const int a = uniq[i], b = uniq[l];
                    const bool up = (i & k) == 0;
                    if ((a > b) == up) {
                        uniq[i] = b;
                        uniq[l] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
    int *const hb = halo + (long)blockIdx.x * kPlanHaloCap;
    for (int i = tid; i < H; i += 256) {
        const int key = uniq[i];
        hb[i] = key;
        unsigned h = plan_hash(key);
        while (keys[h] != key) h = (h + 1) & (kHashSlots - 1);  // key is in the table
        rank[h] = i;
    }
    __syncthreads();
    for (long p = p0 + tid; p < p1; p += 256) {
        const int key = pcol[p];
        unsigned h = plan_hash(key);
        while (keys[h] != key) h = (h + 1) & (kHashSlots - 1);
        plc[p] = (uint16_t)rank[h];
    }
}

__global__ void iota_kernel(int *p, long n)
{
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) p[i] = (int)i;
}

__global__ void max_le_kernel(const int *__restrict__ v, int n, int cap, int *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && v[i] <= cap) atomicMax(out, v[i]);
    if (i < n && v[i] > cap) atomicAdd(out + 1, 1);
}

// Epilogues: MODE 0 y = alpha acc (+ beta y); MODE 1 one SSP-RK stage,
// out = a u + b w + c acc (rows with fixed != 0 keep u).
struct PlanOut {
    double *y;
    double alpha, beta;
    const double *u;
    const uint8_t *fixed;
    double a, b, c;
};

// LDS stride of a staged x row, in doubles: F + 2 from 2 fields up, so rows
// that a wavefront reads at once spread over the banks (at a stride of F
// doubles, 8 fields, its 32 rows hit only 4 bank offsets) and stay 16-byte
// aligned for the staging stores
template <int F>
constexpr int plan_lds_stride() { return F == 1 ? 1 : F + 2; }

// 16 lanes a row, FL lanes a nonzero (FL = 1: every field in every lane), NT
// lanes a block and RPP rows per lane group, so a block's kPlanRows rows run
// in kPlanRows * 16 / (RPP NT) passes.
// Rows of nu entries each (nu > 0) need no row pointers.  xs: the block's halo
// rows of x in LDS when the halo fits (staged).
template <typename TV, int F, int FL, int MODE, int NT, int RPP>
__global__ __launch_bounds__(NT) void plan_apply_kernel(long M, int nu, const int *__restrict__ prow,
                                                        const int64_t *__restrict__ pptr,
                                                        const uint16_t *__restrict__ plc,
                                                        const int *__restrict__ pcol, const TV *__restrict__ pval,
                                                        const int *__restrict__ halo, const int *__restrict__ hcount,
                                                        int lds_cap, bool x16, const double *__restrict__ x,
                                                        PlanOut o)
{
    constexpr int G = 16, FPL = F / FL, NZP = G / FL, SX = plan_lds_stride<F>();
    static_assert(F % FL == 0 && FL <= 2, "field split");
    extern __shared__ double xs[];
    const int tid = threadIdx.x;
    const long r0 = (long)blockIdx.x * kPlanRows;
    const int nr = (int)(M - r0 < kPlanRows ? M - r0 : kPlanRows);
    const int H = hcount[blockIdx.x];
    const bool listed = H <= kPlanHaloCap, staged = H <= lds_cap;
    const int *const hb = halo + (long)blockIdx.x * kPlanHaloCap;
    if (staged) {
        if (F == 1 || !x16) {
            for (int e = tid; e < H * F; e += NT) {
                const int h = e / F, f = e - h * F;
                xs[h * SX + f] = x[(long)hb[h] * F + f];
            }
        } else {
            constexpr int C = F / 2;  // 16-byte pieces of a node's row (x 16-byte aligned)
            for (int e = tid; e < H * C; e += NT) {
                const int h = e / C, part = e - h * C;
                reinterpret_cast<double2 *>(xs + h * SX)[part] =
                    reinterpret_cast<const double2 *>(x + (long)hb[h] * F)[part];
            }
        }
    }
    __syncthreads();
    const int sub = tid & (G - 1), grp = tid / G;
    const int fq = sub % FL, nzp = sub / FL;
    // Each group takes RPP rows a pass (i, i + NT/G, ...), so a lane has RPP
    // times the loads in flight; with rows of nu entries the rows' nonzeros run
    // in one loop.  Every lane runs every pass: the reductions need the whole
    // group.
    constexpr int STRIDE = NT / G;
    for (int i0 = 0; i0 < nr; i0 += RPP * STRIDE) {
        long p0[RPP], p1[RPP], row[RPP];
        bool live[RPP];
#pragma unroll
        for (int h = 0; h < RPP; ++h) {
            const int i = i0 + grp + h * STRIDE;
            live[h] = i < nr;
            p0[h] = p1[h] = row[h] = 0;
            if (live[h]) {
                p0[h] = nu > 0 ? (r0 + i) * (long)nu : pptr[r0 + i];
                p1[h] = nu > 0 ? p0[h] + nu : pptr[r0 + i + 1];
                row[h] = prow[r0 + i];
            }
        }
        double acc[RPP][FPL];
#pragma unroll
        for (int h = 0; h < RPP; ++h)
#pragma unroll
            for (int f = 0; f < FPL; ++f) acc[h][f] = 0.0;
        // the same nonzeros in the same order as the plain kernels
        if (staged && nu > 0) {
            for (int j = nzp; j < nu; j += NZP) {
                double v[RPP];
                int c[RPP];
#pragma unroll
                for (int h = 0; h < RPP; ++h) {
                    v[h] = live[h] ? (double)pval[p0[h] + j] : 0.0;
                    c[h] = live[h] ? (int)plc[p0[h] + j] : 0;
                }
#pragma unroll
                for (int h = 0; h < RPP; ++h) {
                    if (!live[h]) continue;
                    const double *xr = xs + c[h] * SX + fq * FPL;
#pragma unroll
                    for (int f = 0; f < FPL; ++f) acc[h][f] = __builtin_fma(v[h], xr[f], acc[h][f]);
                }
            }
        } else {
#pragma unroll
            for (int h = 0; h < RPP; ++h) {
                if (staged) {
                    for (long p = p0[h] + nzp; p < p1[h]; p += NZP) {
                        const double v = (double)pval[p];
                        const double *xr = xs + (int)plc[p] * SX + fq * FPL;
#pragma unroll
                        for (int f = 0; f < FPL; ++f) acc[h][f] = __builtin_fma(v, xr[f], acc[h][f]);
                    }
                } else {
                    for (long p = p0[h] + nzp; p < p1[h]; p += NZP) {
                        const double v = (double)pval[p];
                        const long cc = listed ? (long)hb[plc[p]] : (long)pcol[p];
                        const double *xr = x + cc * F + fq * FPL;
#pragma unroll
                        for (int f = 0; f < FPL; ++f) acc[h][f] = __builtin_fma(v, xr[f], acc[h][f]);
                    }
                }
            }
        }
#pragma unroll
        for (int h = 0; h < RPP; ++h) {
            if constexpr (FL == 1) {
#pragma unroll
                for (int f = 0; f < F; ++f) acc[h][f] = group_sum_f64<G>(acc[h][f]);
            } else {
#pragma unroll
                for (int f = 0; f < FPL; ++f) {
                    if constexpr (NZP >= 8) acc[h][f] += dpp_f64<0x120 | (FL)>(acc[h][f]);  // row_ror
                    if constexpr (NZP >= 4) acc[h][f] += dpp_f64<0x120 | (G / 4)>(acc[h][f]);
                    if constexpr (NZP >= 2) acc[h][f] += dpp_f64<0x120 | (G / 2)>(acc[h][f]);
                }
            }
        }
#pragma unroll
        for (int h = 0; h < RPP; ++h) {
            if (!live[h]) continue;
            const long rw = row[h];
            auto put = [&](int f, double sum) {
                const long e = rw * F + f;
                if constexpr (MODE == 0) {
                    double r = o.alpha * sum;
                    if (o.beta != 0.0) r = __builtin_fma(o.beta, o.y[e], r);
                    o.y[e] = r;
                } else {
                    const double ui = o.u[e];
                    o.y[e] = (o.fixed && o.fixed[rw]) ? ui : o.a * ui + o.b * x[e] + o.c * sum;
                }
            };
            if constexpr (FL == 1) {
#pragma unroll
                for (int f = 0; f < F; ++f)
                    if (f % G == sub) put(f, acc[h][f]);
            } else if (nzp == 0) {
#pragma unroll
                for (int f = 0; f < FPL; ++f) put(fq * FPL + f, acc[h][f]);
            }
        }
    }
}

template <typename TV, int F, int MODE>
void plan_apply_t(const SpmvPlan &P, const double *x, const PlanOut &o, hipStream_t s)
{
    constexpr int FL = F >= 4 ? 2 : 1, SX = plan_lds_stride<F>();
    // LDS for the largest listed halo, at most 64 KiB a workgroup
    const int cap = std::min(P.hmax, (int)(65536 / (sizeof(double) * SX)));
    const size_t lds = sizeof(double) * SX * (size_t)std::max(cap, 1);
    const bool x16 = ((uintptr_t)x & 15) == 0;
    // 1024 lanes a block from 4 fields up (at 8 fields 161-168 us per product
    // over 1M rows, against 318 us with 256 lanes a block), 256 below
    constexpr int NT = F >= 4 ? 1024 : 256;
    // rows a lane group takes per pass: two up to 2 fields (one field: 85-86 us
    // against 95), one from 4 fields up (two rows: 8 fields no faster, and an
    // FP32 SSP-RK3 run 17% slower)
    constexpr int RPP = F >= 4 ? 1 : 2;
    plan_apply_kernel<TV, F, FL, MODE, NT, RPP><<<P.nblocks, NT, lds, s>>>(
        P.M, P.uniform, P.prow, P.pptr, P.plc, P.pcol, (const TV *)P.pval, P.halo, P.hcount, cap, x16, x, o);
}

template <int MODE>
void plan_apply(const SpmvPlan &P, int F, const double *x, const PlanOut &o, hipStream_t s)
{
    auto go = [&](auto tv) {
        using TV = decltype(tv);
        switch (F) {
        case 1: plan_apply_t<TV, 1, MODE>(P, x, o, s); break;
        case 2: plan_apply_t<TV, 2, MODE>(P, x, o, s); break;
        case 4: plan_apply_t<TV, 4, MODE>(P, x, o, s); break;
        case 8: plan_apply_t<TV, 8, MODE>(P, x, o, s); break;
        default: throw Error{2, "spmv plan: fields must be 1, 2, 4 or 8"};
        }
    };
    if (P.f32) go(float{});
    else go(double{});
}

}  // namespace

void spmv_plan_free(SpmvPlan &P)
{
    for (void *q : {(void *)P.prow, (void *)P.pptr, (void *)P.plc, (void *)P.pcol, P.pval, (void *)P.halo,
                    (void *)P.hcount})
        if (q) (void)hipFree(q);
    P = SpmvPlan{};
}

void spmv_plan_build(SpmvPlan &P, long M, long N, const int64_t *rowptr, const int *col, const void *vals,
                     bool f32, int uniform_n, const int *order, hipStream_t s)
{
    spmv_plan_free(P);
    P.M = M;
    P.N = N;
    P.f32 = f32;
    P.nblocks = (int)((M + kPlanRows - 1) / kPlanRows);
    P.uniform = uniform_n > 0 ? uniform_n : 0;
    if (M <= 0) return;  // products of an empty operator do nothing
    const size_t vb = f32 ? sizeof(float) : sizeof(double);
    try {
        RBFFD_HIP(hipMalloc(&P.pptr, sizeof(int64_t) * (M + 1)));
        RBFFD_HIP(hipMalloc(&P.prow, sizeof(int) * std::max<long>(M, 1)));
        int64_t *len = nullptr;
        RBFFD_HIP(hipMalloc(&len, sizeof(int64_t) * (M + 1)));
        const unsigned b1 = (unsigned)((M + 1 + 255) / 256);
        plan_rowlen_kernel<<<b1, 256, 0, s>>>(M, rowptr, uniform_n, order, len);
        size_t tb = 0;
        RBFFD_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, len, P.pptr, (int)(M + 1), s));
        void *tmp = nullptr;
        RBFFD_HIP(hipMalloc(&tmp, std::max<size_t>(tb, 16)));
        RBFFD_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, tb, len, P.pptr, (int)(M + 1), s));
        if (order)
            RBFFD_HIP(hipMemcpyAsync(P.prow, order, sizeof(int) * M, hipMemcpyDeviceToDevice, s));
        else
            iota_kernel<<<(unsigned)((M + 255) / 256), 256, 0, s>>>(P.prow, M);
        int64_t nnz = 0;
        RBFFD_HIP(hipMemcpyAsync(&nnz, P.pptr + M, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        RBFFD_HIP(hipStreamSynchronize(s));
        (void)hipFree(len);
        (void)hipFree(tmp);
        P.nnz = nnz;
        RBFFD_HIP(hipMalloc(&P.pcol, sizeof(int) * std::max<int64_t>(nnz, 1)));
        RBFFD_HIP(hipMalloc(&P.pval, vb * std::max<int64_t>(nnz, 1)));
        RBFFD_HIP(hipMalloc(&P.plc, sizeof(uint16_t) * std::max<int64_t>(nnz, 1)));
        RBFFD_HIP(hipMalloc(&P.halo, sizeof(int) * (size_t)P.nblocks * kPlanHaloCap));
        RBFFD_HIP(hipMalloc(&P.hcount, sizeof(int) * (P.nblocks + 2)));
        const unsigned b16 = (unsigned)((M * 16 + 255) / 256);
        if (f32)
            plan_copy_kernel<float><<<b16, 256, 0, s>>>(M, rowptr, uniform_n, order, col, (const float *)vals, P.pptr,
                                                        P.pcol, (float *)P.pval);
        else
            plan_copy_kernel<double><<<b16, 256, 0, s>>>(M, rowptr, uniform_n, order, col, (const double *)vals,
                                                         P.pptr, P.pcol, (double *)P.pval);
        plan_halo_kernel<<<P.nblocks, 256, 0, s>>>(M, P.pptr, P.pcol, P.plc, P.halo, P.hcount);
        // largest listed halo, and the count of blocks past the cap
        int *mx = P.hcount + P.nblocks;
        launch_fill_i32(mx, 2, 0, s);
        max_le_kernel<<<(unsigned)((P.nblocks + 255) / 256), 256, 0, s>>>(P.hcount, P.nblocks, kPlanHaloCap, mx);
        RBFFD_HIP(hipGetLastError());
        int h[2] = {0, 0};
        RBFFD_HIP(hipMemcpyAsync(h, mx, sizeof(h), hipMemcpyDeviceToHost, s));
        RBFFD_HIP(hipStreamSynchronize(s));
        P.hmax = h[0];
        P.overflow_blocks = h[1];
        if (P.overflow_blocks == 0) {  // every block reads x through its halo list
            (void)hipFree(P.pcol);
            P.pcol = nullptr;
        }
    } catch (...) {
        spmv_plan_free(P);
        throw;
    }
}

void launch_spmv_plan(const SpmvPlan &P, int F, const double *x, double *y, double alpha, double beta,
                      hipStream_t s)
{
    if (P.M <= 0) return;
    PlanOut o{};
    o.y = y;
    o.alpha = alpha;
    o.beta = beta;
    plan_apply<0>(P, F, x, o, s);
}

void launch_rk_stage_plan(const SpmvPlan &P, int F, const double *w, const double *u, double *out,
                          const uint8_t *fixed, double a, double b, double c, hipStream_t s)
{
    if (P.M <= 0) return;
    PlanOut o{};
    o.y = out;
    o.u = u;
    o.fixed = fixed;
    o.a = a;
    o.b = b;
    o.c = c;
    plan_apply<1>(P, F, w, o, s);
}

}  // namespace rbffd
