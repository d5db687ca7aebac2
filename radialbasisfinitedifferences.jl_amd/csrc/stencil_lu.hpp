// stencil_lu.hpp -- K2, FP64 path: batched assembly and LU solve of the local
// RBF-FD systems, G lanes per stencil.
//
// Replaces hot loops 1 and 2 of the reference (src/generate_operator.jl:55-65
// and :89-166, method 3 at :388-467; src/hyperviscosity_operator.jl:66-161).
//
// Layout.  A wavefront holds 64/G stencils; lane t of a group owns rows
// t, t+G, t+2G, ... of its stencil's saddle-point system
//     A = [Phi P; P^T 0],  Phi_ij = |xi_i - xi_j|^m,  P_iq = x^a y^b (xi_i),
// each row extended by its right-hand-side entries, all in registers.
// LU with partial pivoting: a DPP max over the group picks the pivot row, the
// pivot lane publishes that row as row k of a packed U factor in LDS, and
// every lane eliminates its own rows with one LDS broadcast read feeding ROWS
// FMAs.  Rows never move (the permutation stays implicit), and a column leaves
// the registers once it is eliminated.  Forward substitution rides along in
// the right-hand-side columns; back substitution runs from U in LDS with a
// DPP broadcast of each solved unknown.  LU (rather than Gauss-Jordan) keeps
// the forward error at the level of a backward-stable FP64 solve.
//
// Weights are scaled as the reference does (src/generate_operator.jl:161-166)
// and written straight into CSR rows with sorted columns.
#pragma once
#include <type_traits>

#include "device_common.hpp"
#include "internal.hpp"

namespace rbffd {

template <int A_, int B_>
struct cmax_ { static constexpr int v = A_ > B_ ? A_ : B_; };

// elimination steps per block of equal static row width
constexpr int kStepBlock = 8;

template <int B, int NB, class F>
__device__ __forceinline__ void for_blocks(F &&f)
{
    if constexpr (B < NB) {
        f(std::integral_constant<int, B>{});
        for_blocks<B + 1, NB>(f);
    }
}

// Maximum over each aligned group of G lanes; every lane gets its group's value.
template <int G>
__device__ __forceinline__ uint32_t group_max_u32(uint32_t v)
{
    static_assert(G == 16 || G == 32 || G == 64, "group size");
    v = max(v, dpp_u32<0xB1>(v));   // quad_perm [1,0,3,2]
    v = max(v, dpp_u32<0x4E>(v));   // quad_perm [2,3,0,1]
    v = max(v, dpp_u32<0x141>(v));  // row_half_mirror
    v = max(v, dpp_u32<0x140>(v));  // row_mirror
    if constexpr (G >= 32) {
        // swap with the lane 16 apart inside each 32-lane half (bitmask swizzle, xor 0x10)
        v = max(v, (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, 0x401F));
    }
    if constexpr (G == 64) {
        uint32_t a = (uint32_t)__builtin_amdgcn_readlane((int)v, 0);
        uint32_t b = (uint32_t)__builtin_amdgcn_readlane((int)v, 32);
        v = a > b ? a : b;
    }
    return v;
}

template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v)
{
    int2 p = *reinterpret_cast<int2 *>(&v);
    p.x = __builtin_amdgcn_update_dpp(0, p.x, CTRL, 0xF, 0xF, false);
    p.y = __builtin_amdgcn_update_dpp(0, p.y, CTRL, 0xF, 0xF, false);
    return *reinterpret_cast<double *>(&p);
}

// Value of lane `src` of each G-lane group, in every lane of the group.  src
// is a constant after unrolling: 16-lane groups use DPP row_newbcast (each DPP
// row is one group), wider groups a lane shuffle.
template <int G>
__device__ __forceinline__ double group_bcast(double v, int src)
{
    if constexpr (G == 16) {
        switch (src) {
        case 0: return dpp_f64<0x150>(v);
        case 1: return dpp_f64<0x151>(v);
        case 2: return dpp_f64<0x152>(v);
        case 3: return dpp_f64<0x153>(v);
        case 4: return dpp_f64<0x154>(v);
        case 5: return dpp_f64<0x155>(v);
        case 6: return dpp_f64<0x156>(v);
        case 7: return dpp_f64<0x157>(v);
        case 8: return dpp_f64<0x158>(v);
        case 9: return dpp_f64<0x159>(v);
        case 10: return dpp_f64<0x15A>(v);
        case 11: return dpp_f64<0x15B>(v);
        case 12: return dpp_f64<0x15C>(v);
        case 13: return dpp_f64<0x15D>(v);
        case 14: return dpp_f64<0x15E>(v);
        default: return dpp_f64<0x15F>(v);
        }
    } else {
        return __shfl(v, src, G);
    }
}

template <int G>
__device__ __forceinline__ double group_max_nonneg_f64(double v)
{
    uint64_t b = (uint64_t)__double_as_longlong(v);
#pragma unroll
    for (int o = G / 2; o >= 1; o >>= 1) {
        uint64_t ob = (uint64_t)__shfl_xor((unsigned long long)b, o, G);
        b = ob > b ? ob : b;
    }
    return __longlong_as_double((long long)b);
}

// r^e for a non-negative integer e below 32 that is the same in every lane
__device__ __forceinline__ double upow(double b, int e)
{
    double p = 1.0;
    if (e & 1) p = b;
    b *= b;
    if (e & 2) p *= b;
    b *= b;
    if (e & 4) p *= b;
    b *= b;
    if (e & 8) p *= b;
    b *= b;
    if (e & 16) p *= b;
    return p;
}

// |v|^m for the matrix block: odd m takes one square root
__device__ __forceinline__ double phs_r2(int m, double r2)
{
    double p = upow(r2, m >> 1);
    return (m & 1) ? p * __builtin_sqrt(r2) : p;
}

// Operator applied to the PHS at v: closed forms for the first- and
// second-order operators, the general series for the others.
__device__ inline double op_phs_fast(int op, int k, int m, double vx, double vy, double sx, double sy)
{
    if (op > OP_DXY && op != OP_LAP) return op_phs(op, k, m, vx, vy, sx, sy);
    const double r2 = vx * vx + vy * vy;
    // E uses the matrix entries' own formula: for collocated rows b = A e_c holds
    // bit for bit, so the solve returns the identity row to rounding
    if (op == OP_E) return phs_r2(m, r2);
    const double rm2 = (m >= 2) ? phs_r2(m - 2, r2) : phs_r2(m, r2) / r2;  // r^(m-2)
    const double rm4 = rm2 / r2;                                            // r^(m-4)
    const double dm = (double)m;
    switch (op) {
    case OP_E:   return rm2 * r2;
    case OP_DX:  return dm * vx * rm2;
    case OP_DY:  return dm * vy * rm2;
    case OP_DXX: return dm * rm2 + dm * (dm - 2.0) * vx * vx * rm4;
    case OP_DYY: return dm * rm2 + dm * (dm - 2.0) * vy * vy * rm4;
    case OP_DXY: return dm * (dm - 2.0) * vx * vy * rm4;
    default: {  // OP_LAP with per-axis scales
        const double fxx = dm * rm2 + dm * (dm - 2.0) * vx * vx * rm4;
        const double fyy = dm * rm2 + dm * (dm - 2.0) * vy * vy * rm4;
        return sx * sx * fxx + sy * sy * fyy;
    }
    }
}

template <int NN, int NP, int R, int G>
struct LuShape {
    static constexpr int NS = NN + NP;
    static constexpr int ROWS = (NS + G - 1) / G;  // rows per lane
    static constexpr int WIDTH = NS + R;
    static constexpr int S = 64 / G;               // stencils per wavefront
    static constexpr int NODES = (NN + G - 1) / G; // stencil nodes per lane
    // U factor, packed by pivot step: row k holds columns k..WIDTH-1
    static constexpr int UTOT = NS * WIDTH - NS * (NS - 1) / 2;
    // +kStepBlock: a step writes up to kStepBlock-1 stale values past its row;
    // +2 keeps the groups' slots off one LDS bank
    static constexpr int USTRIDE = cmax_<UTOT, NN * NP>::v + kStepBlock + 2;
    static constexpr int NBLK = (NS + kStepBlock - 1) / kStepBlock;
    static constexpr __host__ __device__ int uoff(int k) { return k * WIDTH - k * (k - 1) / 2; }
};

template <int NN, int NP, int R, int G>
struct alignas(16) LuLDS {
    using T = LuShape<NN, NP, R, G>;
    double xi[T::S][NN][2];
    double u[T::S][T::USTRIDE];  // monomial block during assembly, then the U factor
    int g[T::S][NN];
};

template <int NN, int NP, int R, int G, int WPB>
__global__ __launch_bounds__(64 * WPB) void solve_lu_kernel(const SolveArgs a)
{
    using T = LuShape<NN, NP, R, G>;
    constexpr int NS = T::NS, ROWS = T::ROWS, WIDTH = T::WIDTH, S = T::S, NODES = T::NODES;
    static_assert(NS <= 128, "system size above 128");
    __shared__ LuLDS<NN, NP, R, G> lds_all[WPB];

    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int grp = lane / G;
    const int t = lane % G;
    auto &L = lds_all[w];
    // every lane of the wavefront stays until the end (shuffles span the wave);
    // a group without a task works on a copy of task 0 and stores nothing
    const long task0 = ((long)blockIdx.x * WPB + w) * S;
    if (task0 >= a.ntasks) return;  // whole wavefront
    long task = task0 + grp;
    const bool live = task < a.ntasks;
    if (!live) task = task0;

    const long yk = a.rows ? (long)a.rows[task] : task;
    const long xs = a.y2x ? (long)a.y2x[yk] : yk;
    const int m = a.m;

    // ---- stencil frame (src/scalestencil.jl:10-21, src/generate_operator.jl:100-134)
    int gid[NODES];
    double dxn[NODES], dyn[NODES];
    const int c = a.stencils[xs * NN];
    double sx = 1.0, sy = 1.0, yx, yy;
    {
#pragma clang fp contract(off)
        const double cx = a.X[2 * (long)c], cy = a.X[2 * (long)c + 1];
        double mx = 0.0, my = 0.0;
#pragma unroll
        for (int r = 0; r < NODES; ++r) {
            const int i = t + G * r;
            gid[r] = i < NN ? a.stencils[xs * NN + i] : c;
            dxn[r] = a.X[2 * (long)gid[r]] - cx;
            dyn[r] = a.X[2 * (long)gid[r] + 1] - cy;
            mx = fmax(mx, __builtin_fabs(dxn[r]));
            my = fmax(my, __builtin_fabs(dyn[r]));
        }
        if (a.method == 3) {
            yx = kEps;
            yy = kEps;
        } else {
            mx = group_max_nonneg_f64<G>(mx);
            my = group_max_nonneg_f64<G>(my);
            sx = 1.0 / mx;
            sy = 1.0 / my;
            yx = (a.Y[2 * yk] - cx) * sx;
            yy = (a.Y[2 * yk + 1] - cy) * sy;
        }
#pragma unroll
        for (int r = 0; r < NODES; ++r) {
            const int i = t + G * r;
            double xi_x, xi_y;
            if (a.method == 3) {
                // src/generate_operator.jl:393-410: centre := (eps, eps)
                xi_x = i == 0 ? kEps : dxn[r];
                xi_y = i == 0 ? kEps : dyn[r];
            } else {
                xi_x = dxn[r] * sx;
                xi_y = dyn[r] * sy;
            }
            dxn[r] = xi_x;
            dyn[r] = xi_y;
            if (i < NN) {
                L.xi[grp][i][0] = xi_x;
                L.xi[grp][i][1] = xi_y;
                L.g[grp][i] = gid[r];
            }
        }
    }
    // monomial block P (node-major) for the transposed rows
    constexpr int DEG = mono_deg(NP - 1);
#pragma unroll
    for (int r = 0; r < NODES; ++r) {
        const int i = t + G * r;
        if (i < NN) {
            double pxp[DEG + 1], pyp[DEG + 1];
            pxp[0] = 1.0;
            pyp[0] = 1.0;
#pragma unroll
            for (int d = 1; d <= DEG; ++d) { pxp[d] = pxp[d - 1] * dxn[r]; pyp[d] = pyp[d - 1] * dyn[r]; }
#pragma unroll
            for (int q = 0; q < NP; ++q) L.u[grp][i * NP + q] = pxp[mono_ea(q)] * pyp[mono_eb(q)];
        }
    }
    __builtin_amdgcn_wave_barrier();

    // ---- rows of [A | B] (src/interpolationmatrix.jl:2-5, src/generate_operator.jl:157)
    double row[ROWS][WIDTH];
#pragma unroll
    for (int rr = 0; rr < ROWS; ++rr) {
        const int i = t + G * rr;
        if (i < NN) {
            const double xix = L.xi[grp][i][0], xiy = L.xi[grp][i][1];
#pragma unroll
            for (int j = 0; j < NN; ++j) {
                const double ddx = xix - L.xi[grp][j][0], ddy = xiy - L.xi[grp][j][1];
                row[rr][j] = phs_r2(m, ddx * ddx + ddy * ddy);
            }
#pragma unroll
            for (int q = 0; q < NP; ++q) row[rr][NN + q] = L.u[grp][i * NP + q];
            // RBF rows of the right-hand sides at u = Yscaled - xi (exact zeros -> eps)
            double ux, uy;
            {
#pragma clang fp contract(off)
                if (a.method == 3) {
                    ux = xix;
                    uy = xiy;
                } else {
                    ux = yx - xix;
                    uy = yy - xiy;
                    if (ux == 0.0) ux = kEps;
                    if (uy == 0.0) uy = kEps;
                }
            }
#pragma unroll
            for (int cc = 0; cc < R; ++cc)
                row[rr][NS + cc] = cc < a.nops ? op_phs_fast(a.op_kind[cc], a.op_k[cc], m, ux, uy, sx, sy) : 0.0;
        } else if (i < NS) {
            const int q = i - NN;
#pragma unroll
            for (int j = 0; j < NN; ++j) row[rr][j] = L.u[grp][j * NP + q];
#pragma unroll
            for (int j = NN; j < NS; ++j) row[rr][j] = 0.0;
            int tt = 0;
            while ((tt + 1) * (tt + 2) / 2 <= q) ++tt;
            const int ea = q - tt * (tt + 1) / 2, eb = tt - ea;
#pragma unroll
            for (int cc = 0; cc < R; ++cc)
                row[rr][NS + cc] = cc < a.nops ? op_mono(a.op_kind[cc], a.op_k[cc], ea, eb, yx, yy, sx, sy) : 0.0;
        } else {
#pragma unroll
            for (int j = 0; j < WIDTH; ++j) row[rr][j] = 0.0;
        }
    }
    __builtin_amdgcn_wave_barrier();

    // ---- LU with partial pivoting, forward substitution in the RHS columns.
    // Rows slide one place per step (row[j] <- row[j+1] - l u[j+1]), so the
    // pivot column is always register 0 and every register index is static.
    // Steps run as a real loop in blocks of kStepBlock with a static width that
    // shrinks per block: compact code, at most kStepBlock-1 stale columns of
    // extra work per step.
    uint32_t done = 0;
#pragma unroll
    for (int rr = 0; rr < ROWS; ++rr)
        if (t + G * rr >= NS) done |= 1u << rr;
    bool singular = false;
    double *U = L.u[grp];
    int k = 0;
    for_blocks<0, T::NBLK>([&](auto bc) {
        constexpr int b = decltype(bc)::value;
        constexpr int WB = WIDTH - kStepBlock * b;  // >= WIDTH - k for every k of the block
        const int kend = NS < kStepBlock * (b + 1) ? NS : kStepBlock * (b + 1);
#pragma nounroll
        for (; k < kend; ++k) {
            uint32_t key = 0;
#pragma unroll
            for (int rr = 0; rr < ROWS; ++rr) {
                if (!((done >> rr) & 1u)) {
                    const float f = __builtin_fabsf((float)row[rr][0]);
                    const uint32_t kr = (__float_as_uint(f) & 0xFFFFFF00u) | 0x80u | (uint32_t)(127 - (t + G * rr));
                    key = key > kr ? key : kr;
                }
            }
            const uint32_t kmax = group_max_u32<G>(key);
            const int v = 127 - (int)(kmax & 127u);
            const bool iam = t == v % G;
            const int psel = v / G;
            double *piv = U + T::uoff(k);  // piv[j] = U(k, k + j)
#pragma unroll
            for (int rr = 0; rr < ROWS; ++rr) {
                if (iam && psel == rr) {
#pragma unroll
                    for (int j = 0; j < WB; ++j) piv[j] = row[rr][j];
                    done |= 1u << rr;
                }
            }
            // the other lanes read what the pivot lane just wrote
            __builtin_amdgcn_wave_barrier();
            const double p = piv[0];
            singular |= (p == 0.0);
            double rp = __builtin_amdgcn_rcp(p);
            rp = __builtin_fma(rp, __builtin_fma(-p, rp, 1.0), rp);
            rp = __builtin_fma(rp, __builtin_fma(-p, rp, 1.0), rp);
            double l[ROWS];
#pragma unroll
            for (int rr = 0; rr < ROWS; ++rr) l[rr] = ((done >> rr) & 1u) ? 0.0 : row[rr][0] * rp;
#pragma unroll
            for (int j = 0; j + 1 < WB; ++j) {
                const double pj = piv[j + 1];
#pragma unroll
                for (int rr = 0; rr < ROWS; ++rr) row[rr][j] = __builtin_fma(-l[rr], pj, row[rr][j + 1]);
            }
            // the next step's pivot lane overwrites the stale tail read here
            __builtin_amdgcn_wave_barrier();
        }
    });
    __builtin_amdgcn_wave_barrier();

    // ---- back substitution from the U factor in LDS.  Lane t owns the
    // unknowns i = t + G s: it holds y_i, forms x_i = y_i / U_ii when its turn
    // comes, and a group broadcast of x_k lets every lane update the y_i it
    // owns (i < k).  The lane that owns x_i is the one that writes node i.
    double yv[ROWS][R], xv[ROWS][R], rdiag[ROWS];
    int ubase[ROWS];
#pragma unroll
    for (int s2 = 0; s2 < ROWS; ++s2) {
        const int i = t + G * s2;
        const int ic = i < NS ? i : NS - 1;
        ubase[s2] = T::uoff(ic) - ic;  // U(i, j) = U[ubase + j]
        rdiag[s2] = 1.0 / U[ubase[s2] + ic];
#pragma unroll
        for (int cc = 0; cc < R; ++cc) {
            yv[s2][cc] = U[ubase[s2] + NS + cc];
            xv[s2][cc] = 0.0;
        }
    }
#pragma nounroll
    for (int kk = NS - 1; kk >= 0; --kk) {
        const int sk = kk / G, tk = kk % G;
        double uik[ROWS];
#pragma unroll
        for (int s2 = 0; s2 < ROWS; ++s2) {
            const int i = t + G * s2;
            uik[s2] = (i < kk) ? U[ubase[s2] + kk] : 0.0;
        }
#pragma unroll
        for (int cc = 0; cc < R; ++cc) {
            double own = 0.0;
#pragma unroll
            for (int s2 = 0; s2 < ROWS; ++s2) own = (s2 == sk) ? yv[s2][cc] * rdiag[s2] : own;
            const double xk = __shfl(own, tk, G);
#pragma unroll
            for (int s2 = 0; s2 < ROWS; ++s2) {
                xv[s2][cc] = (s2 == sk && t == tk) ? xk : xv[s2][cc];
                yv[s2][cc] = __builtin_fma(-uik[s2], xk, yv[s2][cc]);
            }
        }
    }

    // ---- CSR output: sorted columns by rank inside the stencil
    uint32_t bad = 0;
#pragma unroll
    for (int r = 0; r < NODES; ++r) {
        const int i = t + G * r;
        if (i < NN && live) {
            int rank = 0;
#pragma unroll
            for (int j = 0; j < NN; ++j) rank += L.g[grp][j] < gid[r] ? 1 : 0;
            const long base = yk * NN + rank;
            if (a.colidx) a.colidx[base] = gid[r];
#pragma unroll
            for (int cc = 0; cc < R; ++cc) {
                if (cc < a.nops) {
                    const double f = a.method == 3 ? 1.0 : op_post_scale(a.op_kind[cc], a.op_k[cc], sx, sy);
                    const double wv = f * xv[r][cc];
                    if (!__builtin_isfinite(wv)) bad = 1;
                    a.vals[cc][base] = wv;
                }
            }
        }
    }
    const uint32_t sflag = group_max_u32<G>((singular ? 2u : 0u) | bad);
    if (a.status && t == 0 && live) {
        const int st = (sflag & 2u) ? 1 : ((sflag & 1u) ? 2 : 0);
        if (a.status_accumulate) a.status[yk] = max(a.status[yk], st);
        else a.status[yk] = st;
    }
}

}  // namespace rbffd
