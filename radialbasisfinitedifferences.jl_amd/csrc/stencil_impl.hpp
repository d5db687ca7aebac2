// stencil_impl.hpp -- K2, mixed-precision path and the per-shape launchers.
//
// Replaces hot loops 1 and 2 of the reference (src/generate_operator.jl:55-65
// and :89-166, method 3 at :388-467; src/hyperviscosity_operator.jl:66-161)
// when FP32 factorisation with FP64 refinement is requested.
// One wavefront owns one output row; lane i holds row i of the saddle-point
// system in FP32 together with the FP64 row of A and the FP64 right-hand side.
// FP32 Gauss-Jordan with partial pivoting on [A | I | B] (pivot row broadcast
// through LDS) gives an FP32 inverse and a first solution; FP64 refinement
// x += inv (b - A x) then runs until the update reaches the FP64 floor or stops
// shrinking.  Rows whose last update is above kAcceptRatio get status 3 and
// are re-solved by the FP64 kernel (stencil_lu.hpp).
//
// Weights are scaled as the reference does and written straight into CSR rows
// whose columns are sorted (rank of each stencil index inside the stencil).
#pragma once
#include "device_common.hpp"
#include "internal.hpp"
#include "stencil_lu.hpp"


namespace rbffd {

// refinement: accept a row whose last update is below this, relative to the
// operator's largest weight
constexpr double kAcceptRatio = 1e-11;

template <int A_, int B_>
struct cmax { static constexpr int v = A_ > B_ ? A_ : B_; };

__device__ __forceinline__ double recip_fast(double p)
{
    double r = __builtin_amdgcn_rcp(p);
    double e = __builtin_fma(-p, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-p, r, 1.0);
    return __builtin_fma(r, e, r);
}
__device__ __forceinline__ float recip_fast(float p) { return __builtin_amdgcn_rcpf(p); }

// NN: stencil size, NP: monomials, R: right-hand sides, MIXED: precision 1.
template <int NN, int NP, int R, bool MIXED, int WPB>
struct SolveShape {
    static constexpr int NS = NN + NP;
    static constexpr int ROWS = (NS + 63) / 64;
    static constexpr int AUG = MIXED ? NS : 0;        // identity columns
    static constexpr int WIDTH = NS + AUG + R;        // augmented row
    static constexpr int BUF = cmax<NN * NP, 2 * NS * R>::v;
};

template <typename TF, int NN, int NP, int R, bool MIXED, int WPB>
struct alignas(16) WaveLDS {
    using S = SolveShape<NN, NP, R, MIXED, WPB>;
    double xi[NN][2];
    double buf[S::BUF];
    TF piv[S::WIDTH + 4];
    int g[NN];
};

template <typename TF, int NN, int NP, int R, bool MIXED, int WPB>
__global__ __launch_bounds__(64 * WPB) void solve_kernel(const SolveArgs a)
{
    using S = SolveShape<NN, NP, R, MIXED, WPB>;
    constexpr int NS = S::NS, ROWS = S::ROWS, WIDTH = S::WIDTH, AUG = S::AUG;
    static_assert(NN <= 64, "stencil size above 64 needs another lane mapping");
    static_assert(NS <= 128, "system size above 128 needs another lane mapping");
    static_assert(!MIXED || ROWS == 1, "mixed precision keeps one row per lane");
    __shared__ WaveLDS<TF, NN, NP, R, MIXED, WPB> lds_all[WPB];

    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const long task = (long)blockIdx.x * WPB + w;
    if (task >= a.ntasks) return;  // whole wavefront; no block barriers below
    auto &L = lds_all[w];

    const long yk = a.rows ? (long)a.rows[task] : task;
    const long xs = a.y2x ? (long)a.y2x[yk] : yk;
    // a failed neighbour search leaves -1 entries; clamp so every read stays in X
    const int gidx = lane < NN ? max(a.stencils[xs * NN + lane], 0) : 0;
    const int c = __builtin_amdgcn_readfirstlane(gidx);

    // ---- stencil frame (src/scalestencil.jl:10-21, src/generate_operator.jl:100-134)
    double sx = 1.0, sy = 1.0, xix, xiy, yx, yy, ux, uy;
    {
#pragma clang fp contract(off)
        const double cx = a.X[2 * (long)c], cy = a.X[2 * (long)c + 1];
        const double px = lane < NN ? a.X[2 * (long)gidx] : cx;
        const double py = lane < NN ? a.X[2 * (long)gidx + 1] : cy;
        const double dx = px - cx, dy = py - cy;
        if (a.method == 3) {
            // src/generate_operator.jl:393-410: shift only, centre := (eps, eps);
            // RBF RHS at +xi (:433), polynomial RHS at the centre (:429-430)
            xix = lane == 0 ? kEps : dx;
            xiy = lane == 0 ? kEps : dy;
            yx = kEps; yy = kEps;
            ux = xix; uy = xiy;
        } else {
            const double mx = wave_max_nonneg_f64(lane < NN ? __builtin_fabs(dx) : 0.0);
            const double my = wave_max_nonneg_f64(lane < NN ? __builtin_fabs(dy) : 0.0);
            sx = 1.0 / mx;
            sy = 1.0 / my;
            xix = dx * sx;
            xiy = dy * sy;
            yx = (a.Y[2 * yk] - cx) * sx;
            yy = (a.Y[2 * yk + 1] - cy) * sy;
            ux = yx - xix;
            uy = yy - xiy;
            if (ux == 0.0) ux = kEps;
            if (uy == 0.0) uy = kEps;
        }
    }

    // ---- publish xi and the monomial block for the transposed rows
    if (lane < NN) {
        L.xi[lane][0] = xix;
        L.xi[lane][1] = xiy;
        L.g[lane] = gidx;
        constexpr int DEG = mono_deg(NP - 1);
        double pxp[DEG + 1], pyp[DEG + 1];
        pxp[0] = 1.0; pyp[0] = 1.0;
#pragma unroll
        for (int d = 1; d <= DEG; ++d) { pxp[d] = pxp[d - 1] * xix; pyp[d] = pyp[d - 1] * xiy; }
#pragma unroll
        for (int q = 0; q < NP; ++q) L.buf[lane * NP + q] = pxp[mono_ea(q)] * pyp[mono_eb(q)];
    }
    __builtin_amdgcn_wave_barrier();

    // ---- assemble rows (src/interpolationmatrix.jl:2-5) and right-hand sides (:157)
    TF row[ROWS][WIDTH];
    double A64[MIXED ? NS : 1];
    double b64[MIXED ? R : 1];
    const int m = a.m;
#pragma unroll
    for (int rr = 0; rr < ROWS; ++rr) {
        const int i = lane + 64 * rr;
        if (i < NN) {
#pragma unroll
            for (int j = 0; j < NN; ++j) {
                const double ddx = xix - L.xi[j][0], ddy = xiy - L.xi[j][1];
                const double v = phs_r2(m, ddx * ddx + ddy * ddy);
                row[rr][j] = (TF)v;
                if constexpr (MIXED) A64[j] = v;
            }
#pragma unroll
            for (int q = 0; q < NP; ++q) {
                const double v = L.buf[i * NP + q];
                row[rr][NN + q] = (TF)v;
                if constexpr (MIXED) A64[NN + q] = v;
            }
        } else if (i < NS) {
            const int q = i - NN;
#pragma unroll
            for (int j = 0; j < NN; ++j) {
                const double v = L.buf[j * NP + q];
                row[rr][j] = (TF)v;
                if constexpr (MIXED) A64[j] = v;
            }
#pragma unroll
            for (int j = NN; j < NS; ++j) {
                row[rr][j] = (TF)0;
                if constexpr (MIXED) A64[j] = 0.0;
            }
        } else {
#pragma unroll
            for (int j = 0; j < NS; ++j) row[rr][j] = (TF)0;
        }
        if constexpr (MIXED) {
#pragma unroll
            for (int j = 0; j < AUG; ++j) row[rr][NS + j] = (TF)(j == i ? 1 : 0);
        }
        // right-hand sides: RBF rows at u (eps-substituted), polynomial rows at
        // the shifted and scaled Y point (src/generate_operator.jl:68-79)
        const int q = i - NN;
        int ea = 0, eb = 0;
        if (i >= NN && i < NS) {
            int t = 0;
            while ((t + 1) * (t + 2) / 2 <= q) ++t;
            ea = q - t * (t + 1) / 2;
            eb = t - ea;
        }
#pragma unroll
        for (int cc = 0; cc < R; ++cc) {
            double v = 0.0;
            if (cc < a.nops) {
                const int op = a.op_kind[cc], kk = a.op_k[cc];
                if (i < NN) v = op_phs_fast(op, kk, m, ux, uy, sx, sy);
                else if (i < NS) v = op_mono(op, kk, ea, eb, yx, yy, sx, sy);
            }
            row[rr][NS + AUG + cc] = (TF)v;
            if constexpr (MIXED) b64[cc] = v;
        }
    }

    // ---- Gauss-Jordan with partial pivoting
    uint32_t done = 0;
#pragma unroll
    for (int rr = 0; rr < ROWS; ++rr)
        if (lane + 64 * rr >= NS) done |= 1u << rr;
    int pstep[ROWS];
    TF pval[ROWS];
#pragma unroll
    for (int rr = 0; rr < ROWS; ++rr) { pstep[rr] = 0; pval[rr] = (TF)1; }
    bool singular = false;

#pragma unroll
    for (int k = 0; k < NS; ++k) {
        uint32_t key = 0;
#pragma unroll
        for (int rr = 0; rr < ROWS; ++rr) {
            if (!((done >> rr) & 1u)) {
                const float f = __builtin_fabsf((float)row[rr][k]);
                const uint32_t kr = (__float_as_uint(f) & 0xFFFFFF00u) | 0x80u | (uint32_t)(127 - (lane + 64 * rr));
                key = key > kr ? key : kr;
            }
        }
        const uint32_t kmax = wave_max_u32(key);
        const int v = 127 - (int)(kmax & 127u);
        const int plane = v & 63, psel = v >> 6;
        const bool iam = lane == plane;
        if (iam) {
#pragma unroll
            for (int j = k; j < WIDTH; ++j) {
                TF e = row[0][j];
                if constexpr (ROWS > 1) e = psel ? row[1][j] : e;
                L.piv[j] = e;
            }
#pragma unroll
            for (int rr = 0; rr < ROWS; ++rr)
                if (psel == rr) { pstep[rr] = k; pval[rr] = row[rr][k]; done |= 1u << rr; }
        }
        __builtin_amdgcn_wave_barrier();
        const TF p = L.piv[k];
        singular |= (p == (TF)0);
        const TF rp = recip_fast(p);
#pragma unroll
        for (int rr = 0; rr < ROWS; ++rr) {
            const bool active = (lane + 64 * rr < NS) && !(iam && psel == rr);
            if (active) {
                const TF l = row[rr][k] * rp;
#pragma unroll
                for (int j = k + 1; j < WIDTH; ++j) row[rr][j] = __builtin_fma(-l, L.piv[j], row[rr][j]);
            }
        }
        __builtin_amdgcn_wave_barrier();
    }

    // ---- solution: x_{pstep} = (augmented rhs) / pivot
    double xsol[ROWS][R];
#pragma unroll
    for (int rr = 0; rr < ROWS; ++rr) {
        const double d = (double)pval[rr];
#pragma unroll
        for (int cc = 0; cc < R; ++cc) xsol[rr][cc] = (double)row[rr][NS + AUG + cc] / d;
    }

    bool converged = true;
    if constexpr (MIXED) {
        // inverse row of the pivot step, FP64 copy of inv = row[NS..2NS) / pivot
        double invr[NS];
        const double rpv = 1.0 / (double)pval[0];
#pragma unroll
        for (int j = 0; j < NS; ++j) invr[j] = (double)row[0][NS + j] * rpv;
        double *xb = L.buf;            // NS x R: x by unknown
        double *rb = L.buf + NS * R;   // NS x R: residual by equation
        const bool real_row = lane < NS;
        converged = false;
        double prev_ratio = 1e300;
        for (int it = 0; it < a.max_iter; ++it) {
            if (real_row) {
#pragma unroll
                for (int cc = 0; cc < R; ++cc) xb[pstep[0] * R + cc] = xsol[0][cc];
            }
            __builtin_amdgcn_wave_barrier();
            double res[R];
#pragma unroll
            for (int cc = 0; cc < R; ++cc) res[cc] = b64[cc];
#pragma unroll
            for (int j = 0; j < NS; ++j) {
#pragma unroll
                for (int cc = 0; cc < R; ++cc) res[cc] = __builtin_fma(-A64[j], xb[j * R + cc], res[cc]);
            }
            if (real_row) {
#pragma unroll
                for (int cc = 0; cc < R; ++cc) rb[lane * R + cc] = res[cc];
            }
            __builtin_amdgcn_wave_barrier();
            // largest update relative to the largest weight, per operator
            double ratio = 0.0;
#pragma unroll
            for (int cc = 0; cc < R; ++cc) {
                double dcc = 0.0;
#pragma unroll
                for (int j = 0; j < NS; ++j) dcc = __builtin_fma(invr[j], rb[j * R + cc], dcc);
                double dm = 0.0, xm = 0.0;
                if (real_row) {
                    xsol[0][cc] += dcc;
                    // weights only: rows >= NN are Lagrange multipliers
                    if (pstep[0] < NN) {
                        dm = __builtin_fabs(dcc);
                        xm = __builtin_fabs(xsol[0][cc]);
                    }
                }
                const double gd = wave_max_nonneg_f64(dm);
                const double gx = wave_max_nonneg_f64(xm);
                ratio = fmax(ratio, gx > 0.0 ? gd / gx : 0.0);
            }
            __builtin_amdgcn_wave_barrier();
            // stop at the FP64 floor, or when the updates stop shrinking; the
            // error left is about the size of the last update
            if (ratio <= a.tol) { converged = true; break; }
            if (ratio > 0.5 * prev_ratio) { converged = ratio <= kAcceptRatio; break; }
            prev_ratio = ratio;
            if (it == a.max_iter - 1) converged = ratio <= kAcceptRatio;
        }
    }

    // ---- move x to the lane of its stencil node
    double *xb = L.buf;
#pragma unroll
    for (int rr = 0; rr < ROWS; ++rr) {
        if (lane + 64 * rr < NS) {
#pragma unroll
            for (int cc = 0; cc < R; ++cc) xb[pstep[rr] * R + cc] = xsol[rr][cc];
        }
    }
    __builtin_amdgcn_wave_barrier();

    uint32_t bad = 0;
    if (lane < NN) {
        int rank = 0;
#pragma unroll
        for (int j = 0; j < NN; ++j) rank += L.g[j] < gidx ? 1 : 0;
        const long base = yk * NN + rank;
        if (a.colidx) a.colidx[base] = gidx;
#pragma unroll
        for (int cc = 0; cc < R; ++cc) {
            if (cc < a.nops) {
                const double f = a.method == 3 ? 1.0 : op_post_scale(a.op_kind[cc], a.op_k[cc], sx, sy);
                const double wv = f * xb[lane * R + cc];
                if (!__builtin_isfinite(wv)) bad = 1;
                a.vals[cc][base] = wv;
            }
        }
    }
    bad = wave_or_u32(bad);
    if (a.status && lane == 0) {
        int st = 0;
        if (singular) st = 1;
        else if (bad) st = 2;
        else if (!converged) st = 3;
        if (a.status_accumulate) a.status[yk] = max(a.status[yk], st);
        else a.status[yk] = st;
    }
}

// ------------------------------------------------------------ launchers
constexpr int kWPB = 4;

template <typename TF, int NN, int NP, int R, bool MIXED>
void launch_t(const SolveArgs &a, hipStream_t s)
{
    const unsigned blocks = (unsigned)((a.ntasks + kWPB - 1) / kWPB);
    solve_kernel<TF, NN, NP, R, MIXED, kWPB><<<blocks, 64 * kWPB, 0, s>>>(a);
}

// Wavefronts per block for a kernel using `bytes` of LDS per wavefront: the
// most wavefronts resident in a CU's 160 KiB of LDS, blocks within 64 KiB,
// larger blocks on ties.
constexpr int kLdsPerCU = 160 * 1024;
constexpr int resident_waves(int bytes, int wpb)
{
    return bytes * wpb > 65536 ? 0 : (kLdsPerCU / (bytes * wpb)) * wpb;
}
constexpr int lu_wpb(int bytes)
{
    int best = 1;
    for (int w = 2; w <= 4; ++w)
        if (resident_waves(bytes, w) >= resident_waves(bytes, best)) best = w;
    return best;
}

// Wavefronts per SIMD the FP64 Gauss-Jordan kernel is compiled for: three
// (168 registers) where the row registers, ROWS x WIDTH doubles, leave the
// headroom the rest of the kernel needs at G lanes per stencil (measured on
// the compiled shapes: 34 registers at 64 lanes, 44 at 16), else no bound.
template <int NN, int NP, int R, int G>
constexpr int lu_min_waves()
{
    using T = LuShape<NN, NP, R, G>;
    constexpr int head = G == 16 ? 44 : (G == 32 ? 40 : 34);
    return T::ROWS * T::WIDTH * 2 + head <= 168 ? 3 : 1;
}

// FP64 Gauss-Jordan kernel: G lanes per stencil.
template <int NN, int NP, int R, int G>
void launch_lu(const SolveArgs &a, hipStream_t s)
{
    constexpr int bytes = (int)sizeof(LuLDS<NN, NP, R, G>);
    // the 16-lane kernel fills a CU with 8 one-wavefront blocks as well as
    // with 4 two-wavefront ones, and single wavefronts free their slots
    // one at a time (7.51 against 7.58 ms for 1M stencils of n = 31)
    constexpr int wpb = G == 16 ? 1 : lu_wpb(bytes);
    constexpr int per_block = wpb * (64 / G);
    const unsigned blocks = (unsigned)((a.ntasks + per_block - 1) / per_block);
    solve_lu_kernel<NN, NP, R, G, wpb, lu_min_waves<NN, NP, R, G>()><<<blocks, 64 * wpb, 0, s>>>(a);
}

// Grouped tasks (Y rows sharing a stencil, a.grp_x) in the specialised
// kernel, one row slot per lane.
template <int NN, int NP, int R>
void launch_lu_grp(const SolveArgs &a, hipStream_t s)
{
    constexpr int G = lu_grp_lanes(NN);
    constexpr int bytes = (int)(sizeof(LuLDS<NN, NP, R, G>) + sizeof(LuRows<R, 64 / G>));
    constexpr int wpb = G == 16 ? 1 : lu_wpb(bytes);
    constexpr int per_block = wpb * (64 / G);
    const unsigned blocks = (unsigned)((a.ntasks + per_block - 1) / per_block);
    solve_lu_kernel<NN, NP, R, G, wpb, lu_min_waves<NN, NP, R, G>(), true><<<blocks, 64 * wpb, 0, s>>>(a);
}

template <int NN, int NP>
void launch_shape_grp(const SolveArgs &a, hipStream_t s)
{
    static_assert(kLuGrpMaxRhs == 24, "right-hand-side blocks");
    if (a.gen_cols <= 6) launch_lu_grp<NN, NP, 6>(a, s);
    else if (a.gen_cols <= 12) launch_lu_grp<NN, NP, 12>(a, s);
    else if (a.gen_cols <= 24) launch_lu_grp<NN, NP, 24>(a, s);
    else throw Error{2, "grouped tasks of at most 24 right-hand-side columns in a specialised kernel"};
}

// Lane-group sizes compiled for a stencil size: g0 is the default, the others
// (0 = none) are candidates for the autotuner.  A lane holds ceil(n / G) row
// slots of n + P + r doubles; a layout is compiled where those stay within
// 200 registers at r = 3 (no spill with the rest of the kernel) and where at
// least half of a group's lanes hold a node.  g0 is the smallest such G
// (most stencils per wavefront), 64 when none fits.
template <int NN, int NS>
struct LuGroups {
    static constexpr bool fits(int g) { return ((NN + g - 1) / g) * (NS + 3) * 2 <= 200; }
    static constexpr bool used(int g) { return fits(g) && 2 * NN > g; }
    static constexpr int g0 = used(16) ? 16 : (used(32) ? 32 : 64);
    static constexpr int g1 = g0 < 32 && used(32) ? 32 : (g0 < 64 && used(64) ? 64 : 0);
    static constexpr int g2 = g0 < 32 && g1 == 32 && used(64) ? 64 : 0;
};

template <int NN, int NP, int R>
void launch_prec(const SolveArgs &a, int precision, hipStream_t s)
{
    constexpr int NS = NN + NP;
    if (precision == 1) {
        if constexpr (NS <= 64) launch_t<float, NN, NP, R, true>(a, s);
        else throw Error{2, "mixed precision needs n + P <= 64"};
    } else {
        constexpr int g0 = LuGroups<NN, NS>::g0, g1 = LuGroups<NN, NS>::g1, g2 = LuGroups<NN, NS>::g2;
        if constexpr (g1 != 0) {
            if (a.group == g1) { launch_lu<NN, NP, R, g1>(a, s); return; }
        }
        if constexpr (g2 != 0) {
            if (a.group == g2) { launch_lu<NN, NP, R, g2>(a, s); return; }
        }
        launch_lu<NN, NP, R, g0>(a, s);
    }
}

// Right-hand-side widths compiled per shape; more operators run in chunks.
template <int NN, int NP>
void launch_shape(const SolveArgs &a, int precision, hipStream_t s)
{
    if (a.nops <= 1) launch_prec<NN, NP, 1>(a, precision, s);
    else if (a.nops <= 2) launch_prec<NN, NP, 2>(a, precision, s);
    else if (a.nops <= 3) launch_prec<NN, NP, 3>(a, precision, s);
    else if (a.nops <= kMaxRhs) launch_prec<NN, NP, kMaxRhs>(a, precision, s);
    else throw Error{2, "at most 6 right-hand sides per launch"};
}

template <int NN, int NP>
void shape_groups(int precision, int *groups)
{
    constexpr int NS = NN + NP;
    groups[0] = precision == 1 ? 64 : LuGroups<NN, NS>::g0;
    groups[1] = precision == 1 ? 0 : LuGroups<NN, NS>::g1;
    groups[2] = precision == 1 ? 0 : LuGroups<NN, NS>::g2;
}

}  // namespace rbffd

// One translation unit per stencil shape keeps the build parallel.
#define RBFFD_DEFINE_SHAPE(nn, pp)                                                             \
    namespace rbffd {                                                                          \
    void launch_solve_##nn##_##pp(const SolveArgs &a, int precision, hipStream_t s)            \
    {                                                                                          \
        launch_shape<nn, pp>(a, precision, s);                                                 \
    }                                                                                          \
    void launch_solve_grp_##nn##_##pp(const SolveArgs &a, hipStream_t s) { launch_shape_grp<nn, pp>(a, s); } \
    void groups_##nn##_##pp(int precision, int *groups) { shape_groups<nn, pp>(precision, groups); } \
    }


