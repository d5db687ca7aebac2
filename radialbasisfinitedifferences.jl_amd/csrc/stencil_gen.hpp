// stencil_gen.hpp -- K2 for every stencil shape, and for output rows that
// share a stencil (Y distinct from X).
//
// Replaces hot loops 1 and 2 of the reference for any stencil size n and
// polynomial degree (src/generate_operator.jl:55-166, method 3 at :388-467;
// src/hyperviscosity_operator.jl:66-161; the reference accepts any n and
// polydeg, src/generate_operator.jl:29, src/polynomialbasis.jl:6).
//
// Padding.  The local system of nsys = n + P equations is padded to a
// compiled size NSB with identity equations decoupled from it: equation e >=
// nsys is x_e = 0, and column e is zero in every real equation.  One kernel
// therefore serves every shape up to NSB; the padding rows pivot on their own
// columns in the last steps (the real rows are zero there, so their
// multipliers are zero) and leave the solution unchanged.
//
// Factor once.  A task is one stencil together with the Y rows that use it:
// the reference factors each X node's system once and applies it to every Y
// row mapped to that node (src/generate_operator.jl:55-65, :158).  Here the
// right-hand sides of up to R / nops Y rows ride along in one elimination.
// Each right-hand-side column sees the same operations wherever it sits, so
// a row's weights do not depend on which rows share its task.
//
// Layout and elimination follow the specialised kernel (stencil_lu.hpp): G
// lanes per stencil, ROWS row slots of [A | B] per lane in registers, Gauss-
// Jordan with implicit partial pivoting, the pivot row published through LDS
// and spread by DPP broadcasts (read with v_readlane at 64 lanes), the next
// pivot search overlapping each step.
// Column order: w (n), lambda (P), padding.  Matrix entries come from LDS
// tables (scaled coordinates and the monomial block, stride P), so register
// indices stay static while n and P vary at run time.
//
// Systems above 128 equations (where a row slot id no longer fits the pivot
// key) run in solve_mem_kernel: one wavefront per stencil, the augmented
// system in a global workspace, column-major so a column of the lanes' rows is
// one coalesced access.
#pragma once
#include "stencil_lu.hpp"

namespace rbffd {

// Exponents (a, b) of monomial q (x^a y^b, graded by total degree) for a q
// known only at run time.
__device__ __forceinline__ void mono_exponents(int q, int &ea, int &eb)
{
    int tt = 0;
    while ((tt + 1) * (tt + 2) / 2 <= q) ++tt;
    ea = q - tt * (tt + 1) / 2;
    eb = tt - ea;
}

template <int NSB, int R, int G>
struct GenShape {
    static constexpr int ROWS = (NSB + G - 1) / G;  // row slots per lane
    static constexpr int WIDTH = NSB + R;
    static constexpr int S = 64 / G;                // stencils per wavefront
    static constexpr int SB = step_block<G>();
    static constexpr int UBUF = 16 * ((WIDTH + 15) / 16);
    static constexpr int PM = NSB * NSB / 4;        // n P <= (n + P)^2 / 4
    // highest monomial degree whose basis still leaves room for one node
    static constexpr int DEGB = mono_deg(NSB - 2);
    static constexpr int QB = (DEGB + 1) * (DEGB + 2) / 2;
};

template <int NSB, int R, int G>
struct alignas(16) GenLDS {
    using T = GenShape<NSB, R, G>;
    double u[T::S][T::UBUF + 2];  // the current pivot row
    double xi[T::S][NSB][2];      // scaled stencil coordinates
    double pm[T::S][T::PM];       // P(node j, monomial q) at j * P + q
    double ys[T::S][R][2];        // the task's Y points in the stencil frame
    double rp[T::S];              // reciprocal of the current pivot
    int g[T::S][NSB];             // global node indices
    int rank[T::S][NSB];          // rank of node i's global index inside the stencil
    int yrow[T::S][R];            // the task's output rows
};

template <int NSB, int R, int G, int WPB>
__global__ __launch_bounds__(64 * WPB, 1) void solve_gen_kernel(const SolveArgs a)
{
    using T = GenShape<NSB, R, G>;
    constexpr int ROWS = T::ROWS, WIDTH = T::WIDTH, S = T::S;
    static_assert(NSB <= 128 && G * ROWS <= 128, "row slot ids take 7 bits of the pivot key");
    __shared__ GenLDS<NSB, R, G> lds_all[WPB];

    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int grp = lane / G;
    const int t = lane % G;
    auto &L = lds_all[w];
    long ntasks = a.ntasks;
    if (a.ntasks_dev) ntasks = min(ntasks, (long)*a.ntasks_dev);
    // every lane of the wavefront stays until the end (shuffles span the wave);
    // a group without a task works on a copy of task 0 and stores nothing
    const long task0 = ((long)blockIdx.x * WPB + w) * S;
    if (task0 >= ntasks) return;  // whole wavefront
    long task = task0 + grp;
    const bool live = task < ntasks;
    if (!live) task = task0;

    const int n = a.n, P = a.P, nsys = n + P, m = a.m, nops = a.nops;
    const bool method3 = a.method == 3;
    // the task's stencil and output rows
    long xs, ybase;
    int ny;
    if (a.grp_x) {
        xs = a.grp_x[task];
        ybase = a.grp_start[task];
        ny = a.grp_count[task];
    } else {
        ybase = a.rows ? (long)a.rows[task] : task;
        xs = a.y2x ? (long)a.y2x[ybase] : ybase;
        ny = 1;
    }
    // ncols <= R and ny <= 64 (one status bit per output row): the launcher's choice
    const int ncols = ny * nops;

    // ---- stencil frame (src/scalestencil.jl:10-21, src/generate_operator.jl:100-134)
    int gid[ROWS];
    double dxn[ROWS], dyn[ROWS];
    // a failed neighbour search leaves -1 entries; clamp so every read stays in X
    const int c = max(a.stencils[xs * n], 0);
    double sx = 1.0, sy = 1.0, cx, cy;
    {
#pragma clang fp contract(off)
        cx = a.X[2 * (long)c];
        cy = a.X[2 * (long)c + 1];
        double mx = 0.0, my = 0.0;
#pragma unroll
        for (int r = 0; r < ROWS; ++r) {
            const int i = t + G * r;
            gid[r] = i < n ? max(a.stencils[xs * n + i], 0) : c;
            dxn[r] = a.X[2 * (long)gid[r]] - cx;
            dyn[r] = a.X[2 * (long)gid[r] + 1] - cy;
            mx = fmax(mx, __builtin_fabs(dxn[r]));
            my = fmax(my, __builtin_fabs(dyn[r]));
        }
        if (!method3) {
            mx = group_max_nonneg_f64<G>(mx);
            my = group_max_nonneg_f64<G>(my);
            sx = 1.0 / mx;
            sy = 1.0 / my;
        }
#pragma unroll
        for (int r = 0; r < ROWS; ++r) {
            const int i = t + G * r;
            double xi_x, xi_y;
            if (method3) {
                // src/generate_operator.jl:393-410: centre := (eps, eps)
                xi_x = i == 0 ? kEps : dxn[r];
                xi_y = i == 0 ? kEps : dyn[r];
            } else {
                xi_x = dxn[r] * sx;
                xi_y = dyn[r] * sy;
            }
            dxn[r] = xi_x;
            dyn[r] = xi_y;
            if (i < n) {
                L.xi[grp][i][0] = xi_x;
                L.xi[grp][i][1] = xi_y;
                L.g[grp][i] = gid[r];
            }
        }
        // the task's Y points in the stencil frame (src/generate_operator.jl:112-124);
        // method 3 evaluates the polynomial right-hand side at the centre (:429-430)
        for (int s = t; s < ny; s += G) {
            const long yk = a.grp_x ? (long)a.ylist[ybase + s] : ybase;
            L.yrow[grp][s] = (int)yk;
            if (method3) {
                L.ys[grp][s][0] = kEps;
                L.ys[grp][s][1] = kEps;
            } else {
                L.ys[grp][s][0] = (a.Y[2 * yk] - cx) * sx;
                L.ys[grp][s][1] = (a.Y[2 * yk + 1] - cy) * sy;
            }
        }
    }
    // the monomial block, node-major with stride P (src/polynomialblock.jl:23-30)
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
        const int i = t + G * r;
        if (i < n) {
            double pxp[T::DEGB + 1], pyp[T::DEGB + 1];
            pxp[0] = 1.0;
            pyp[0] = 1.0;
#pragma unroll
            for (int d = 1; d <= T::DEGB; ++d) { pxp[d] = pxp[d - 1] * dxn[r]; pyp[d] = pyp[d - 1] * dyn[r]; }
#pragma unroll
            for (int q = 0; q < T::QB; ++q)
                if (q < P) L.pm[grp][i * P + q] = pxp[mono_ea(q)] * pyp[mono_eb(q)];
        }
    }
    __builtin_amdgcn_wave_barrier();
    // each node's output column: the rank of its global index inside the
    // stencil (CSR rows come out with sorted columns, as sparse() sorts)
    {
        int rk[ROWS];
#pragma unroll
        for (int r = 0; r < ROWS; ++r) rk[r] = 0;
        for (int j = 0; j < n; ++j) {
            const int gj = L.g[grp][j];
#pragma unroll
            for (int r = 0; r < ROWS; ++r) rk[r] += gj < gid[r] ? 1 : 0;
        }
#pragma unroll
        for (int r = 0; r < ROWS; ++r)
            if (t + G * r < n) L.rank[grp][t + G * r] = rk[r];
    }

    // ---- right-hand sides, while few registers are live.  Column cc belongs
    // to output row col_sub[cc] of the task and operator col_op[cc].  RBF rows
    // at u = Yscaled - xi with exact zeros -> eps (src/generate_operator.jl:125-154;
    // method 3 at +xi, :433), polynomial rows at Yscaled (:68-79)
    double rhs[ROWS][R];
#pragma unroll
    for (int rr = 0; rr < ROWS; ++rr) {
        const int e = t + G * rr;
#pragma unroll
        for (int cc = 0; cc < R; ++cc) rhs[rr][cc] = 0.0;
        if (e < n) {
            double ux = 0.0, uy = 0.0;
            PhsAt q{};
#pragma unroll
            for (int cc = 0; cc < R; ++cc) {
                if (cc < ncols) {
                    if (a.col_op[cc] == 0) {  // the first operator of a new output row
#pragma clang fp contract(off)
                        if (method3) {
                            ux = dxn[rr];
                            uy = dyn[rr];
                        } else {
                            ux = L.ys[grp][a.col_sub[cc]][0] - dxn[rr];
                            uy = L.ys[grp][a.col_sub[cc]][1] - dyn[rr];
                            if (ux == 0.0) ux = kEps;
                            if (uy == 0.0) uy = kEps;
                        }
                        q = phs_at(m, ux, uy);
                    }
                    rhs[rr][cc] = op_phs_fast(a.col_kind[cc], a.col_k[cc], m, ux, uy, sx, sy, q);
                }
            }
        } else if (e < nsys) {
            int ea, eb;
            mono_exponents(e - n, ea, eb);
#pragma unroll
            for (int cc = 0; cc < R; ++cc) {
                if (cc < ncols) {
                    const int s = a.col_sub[cc];
                    rhs[rr][cc] = op_mono(a.col_kind[cc], a.col_k[cc], ea, eb, L.ys[grp][s][0], L.ys[grp][s][1], sx,
                                          sy);
                }
            }
        }
    }

    // ---- rows of [A | B] (src/interpolationmatrix.jl:2-5, src/rbfblock.jl:14-20,
    // src/generate_operator.jl:157): w columns, lambda columns, padding
    double row[ROWS][WIDTH];
#pragma unroll
    for (int rr = 0; rr < ROWS; ++rr) {
        const int e = t + G * rr;
        if (e < n) {
#pragma unroll
            for (int j = 0; j < NSB; ++j) {
                double v = 0.0;
                if (j < n) {
                    const double ddx = dxn[rr] - L.xi[grp][j][0], ddy = dyn[rr] - L.xi[grp][j][1];
                    v = phs_r2(m, sq_sum(ddx, ddy));
                } else if (j < nsys) {
                    v = L.pm[grp][e * P + (j - n)];
                }
                row[rr][j] = v;
            }
        } else if (e < nsys) {
            const int q = e - n;
#pragma unroll
            for (int j = 0; j < NSB; ++j) row[rr][j] = j < n ? L.pm[grp][j * P + q] : 0.0;
        } else {
            // padding: x_e = 0 (also for slots past NSB, which stay dead)
#pragma unroll
            for (int j = 0; j < NSB; ++j) row[rr][j] = j == e ? 1.0 : 0.0;
        }
#pragma unroll
        for (int cc = 0; cc < R; ++cc) row[rr][NSB + cc] = rhs[rr][cc];
    }
    __builtin_amdgcn_wave_barrier();

    // ---- Gauss-Jordan elimination with partial pivoting (the scheme of
    // solve_lu_kernel): rows slide one register per step, so the pivot column
    // is register 0 and every register index is static.
    uint32_t alive[ROWS];
    int kp[ROWS];
    double pval[ROWS];
#pragma unroll
    for (int rr = 0; rr < ROWS; ++rr) {
        alive[rr] = t + G * rr < NSB ? 0xFFFFFFFFu : 0u;
        kp[rr] = NSB;
        pval[rr] = 1.0;
    }
    double *U = L.u[grp];
    uint32_t key;
    double crp;
    auto candidate = [&]() {
        uint32_t kk = 0;
        double cv = row[0][0];
#pragma unroll
        for (int rr = 0; rr < ROWS; ++rr) {
            const uint32_t kr = ((__float_as_uint((float)row[rr][0]) & 0x7FFFFF00u) |
                                 (0x80u | (uint32_t)(127 - (t + G * rr)))) & alive[rr];
            if (rr == 0) {
                kk = kr;
            } else if (kr > kk) {
                kk = kr;
                cv = row[rr][0];
            }
        }
        double r = __builtin_amdgcn_rcp(cv);
        r = __builtin_fma(r, __builtin_fma(-cv, r, 1.0), r);
        r = __builtin_fma(r, __builtin_fma(-cv, r, 1.0), r);
        key = kk;
        crp = r;
    };
    candidate();
    uint32_t kmax = group_max_u32<G>(key);
    int k = 0;
    // 64 lanes (one stencil per wavefront): the pivot lane is wavefront-
    // uniform and every lane reads the pivot row from its registers with
    // v_readlane (as solve_lu_kernel does at 64 lanes); no LDS round trip
    auto step64 = [&](auto wbc) {
        constexpr int WB = decltype(wbc)::value;
        const int v = 127 - (int)((uint32_t)__builtin_amdgcn_readfirstlane((int)kmax) & 127u);
        const int wl = v & 63, psel = v >> 6;
        auto rl = [&](double x) {
            const long long b = __double_as_longlong(x);
            const int lo = __builtin_amdgcn_readlane((int)b, wl);
            const int hi = __builtin_amdgcn_readlane((int)(b >> 32), wl);
            return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
        };
        for_blocks<0, ROWS>([&](auto rc) {
            constexpr int ps = decltype(rc)::value;
            if (psel != ps) return;  // uniform
            const double rp = rl(crp);
            const double p1 = rl(row[ps][1]);
            if (t == wl) {
                alive[ps] = 0u;
                kp[ps] = k;
                pval[ps] = row[ps][0];
                row[ps][0] = 0.0;  // a zero multiplier keeps the pivot row as it is
            }
            double nl[ROWS];
#pragma unroll
            for (int rr = 0; rr < ROWS; ++rr) nl[rr] = -(row[rr][0] * rp);
#pragma unroll
            for (int rr = 0; rr < ROWS; ++rr) row[rr][0] = __builtin_fma(nl[rr], p1, row[rr][1]);
            candidate();
            kmax = group_max_u32<G>(key);
            constexpr int NQ = (WB - 2) / 4;
            for_blocks<0, NQ>([&](auto qc) {
                constexpr int j = 1 + 4 * decltype(qc)::value;
                const double s0 = rl(row[ps][j + 1]), s1 = rl(row[ps][j + 2]);
                const double s2 = rl(row[ps][j + 3]), s3 = rl(row[ps][j + 4]);
#pragma unroll
                for (int rr = 0; rr < ROWS; ++rr)
                    slide4_s(row[rr][j], row[rr][j + 1], row[rr][j + 2], row[rr][j + 3], row[rr][j + 4], nl[rr], s0, s1,
                             s2, s3);
            });
            for_blocks<1 + 4 * NQ, WB - 1>([&](auto jc) {
                constexpr int j = decltype(jc)::value;
                const double sj = rl(row[ps][j + 1]);
#pragma unroll
                for (int rr = 0; rr < ROWS; ++rr) slide1_s(row[rr][j], row[rr][j + 1], nl[rr], sj);
            });
        });
    };
    auto step = [&](auto wbc) {
        constexpr int WB = decltype(wbc)::value;
        if constexpr (G == 64) {
            step64(wbc);
            return;
        }
        const int v = 127 - (int)(kmax & 127u);
        const bool iam = t == v % G;
        const int psel = v / G;
        double *piv = U;  // piv[j] = pivot row, column k + j
        constexpr int W0 = WB < 16 ? WB : 16;
        constexpr int NPV = (WB + 15) / 16;
        uint64_t orp = 0, op1 = 0;
        if constexpr (G == 16) {
            if (iam) orp = (uint64_t)__double_as_longlong(crp);
#pragma unroll
            for (int rr = 0; rr < ROWS; ++rr)
                if (iam && psel == rr) op1 = (uint64_t)__double_as_longlong(row[rr][1]);
        } else {
            if (iam) L.rp[grp] = crp;
        }
#pragma unroll
        for (int rr = 0; rr < ROWS; ++rr) {
            if (iam && psel == rr) {
#pragma unroll
                for (int j = G == 16 ? 2 : 0; j < W0; ++j) piv[j] = row[rr][j];
                alive[rr] = 0u;
                kp[rr] = k;
                pval[rr] = row[rr][0];
                row[rr][0] = 0.0;  // a zero multiplier keeps the pivot row as it is
            }
        }
        __builtin_amdgcn_wave_barrier();
        double pv[NPV];
        pv[0] = piv[lane & 15];
        double rp, p1;
        if constexpr (G == 16) {
            rp = __longlong_as_double((long long)row_or_u64(orp));
            p1 = __longlong_as_double((long long)row_or_u64(op1));
        } else {
            rp = L.rp[grp];
            p1 = piv[1];
        }
        double nl[ROWS];
#pragma unroll
        for (int rr = 0; rr < ROWS; ++rr) nl[rr] = -(row[rr][0] * rp);
#pragma unroll
        for (int rr = 0; rr < ROWS; ++rr) row[rr][0] = fma3(nl[rr], p1, row[rr][1]);
        candidate();
        kmax = group_max_u32<G>(key);
        auto columns = [&](auto j0c, auto j1c) {
            constexpr int J0 = decltype(j0c)::value, J1 = decltype(j1c)::value;
            constexpr int NQ = (J1 - J0) / 4;
            for_blocks<0, NQ>([&](auto qc) {
                constexpr int j = J0 + 4 * decltype(qc)::value;
                double q0, q1, q2, q3;
                row_bcast64x4<(j + 1) % 16, (j + 2) % 16, (j + 3) % 16, (j + 4) % 16>(
                    pv[(j + 1) / 16], pv[(j + 2) / 16], pv[(j + 3) / 16], pv[(j + 4) / 16], q0, q1, q2, q3);
#pragma unroll
                for (int rr = 0; rr < ROWS; ++rr) row[rr][j] = fma3(nl[rr], q0, row[rr][j + 1]);
#pragma unroll
                for (int rr = 0; rr < ROWS; ++rr) row[rr][j + 1] = fma3(nl[rr], q1, row[rr][j + 2]);
#pragma unroll
                for (int rr = 0; rr < ROWS; ++rr) row[rr][j + 2] = fma3(nl[rr], q2, row[rr][j + 3]);
#pragma unroll
                for (int rr = 0; rr < ROWS; ++rr) row[rr][j + 3] = fma3(nl[rr], q3, row[rr][j + 4]);
            });
            constexpr int J2 = J0 + 4 * NQ;
            if constexpr (J1 - J2 >= 2) {
                constexpr int j = J2;
                double q0, q1;
                row_bcast64x2<(j + 1) % 16, (j + 2) % 16>(pv[(j + 1) / 16], pv[(j + 2) / 16], q0, q1);
#pragma unroll
                for (int rr = 0; rr < ROWS; ++rr) row[rr][j] = fma3(nl[rr], q0, row[rr][j + 1]);
#pragma unroll
                for (int rr = 0; rr < ROWS; ++rr) row[rr][j + 1] = fma3(nl[rr], q1, row[rr][j + 2]);
            }
            if constexpr ((J1 - J2) % 2 == 1) {
                constexpr int j = J1 - 1;
                const double pj = row_bcast64<(j + 1) % 16>(pv[(j + 1) / 16]);
#pragma unroll
                for (int rr = 0; rr < ROWS; ++rr) row[rr][j] = fma3(nl[rr], pj, row[rr][j + 1]);
            }
        };
        constexpr int JE = WB - 1;  // columns j = 1 .. WB-2
        constexpr int JA = JE < 15 ? JE : 15;
        columns(std::integral_constant<int, 1>{}, std::integral_constant<int, JA>{});
        if constexpr (NPV > 1) {
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int rr = 0; rr < ROWS; ++rr) {
                if (iam && psel == rr) {
#pragma unroll
                    for (int j = 16; j < WB; ++j) piv[j] = row[rr][j];
                }
            }
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int i = 1; i < NPV; ++i) pv[i] = piv[(lane & 15) + 16 * i];
            columns(std::integral_constant<int, JA>{}, std::integral_constant<int, JE>{});
        }
    };
    for_blocks<0, (NSB + T::SB - 1) / T::SB>([&](auto bc) {
        constexpr int kb = T::SB * decltype(bc)::value;
        const int kend = NSB < kb + T::SB ? NSB : kb + T::SB;
#pragma nounroll
        for (; k < kend; ++k) {
            step(std::integral_constant<int, WIDTH - kb>{});
            __builtin_amdgcn_wave_barrier();
        }
    });
    __builtin_amdgcn_wave_barrier();

    // ---- output: after NSB slides a row's registers 0..R-1 hold its
    // right-hand sides reduced to pivot * x for its pivot column; the row that
    // solved for node u writes u's weights at the rank of u's global index
    uint64_t bad = 0;
#pragma unroll
    for (int rr = 0; rr < ROWS; ++rr) {
        const int u = kp[rr];
        if (u < n && live) {
            const int gu = L.g[grp][u];
            const int rku = L.rank[grp][u];
            const double rd = 1.0 / pval[rr];
#pragma unroll
            for (int cc = 0; cc < R; ++cc) {
                if (cc < ncols) {
                    const int s = a.col_sub[cc], o = a.col_op[cc];
                    const long base = (long)L.yrow[grp][s] * n + rku;
                    const double f = method3 ? 1.0 : op_post_scale(a.col_kind[cc], a.col_k[cc], sx, sy);
                    const double wv = f * (row[rr][cc] * rd);
                    if (!__builtin_isfinite(wv)) bad |= 1ull << s;
                    a.vals[o][base] = wv;
                    if (o == 0 && a.colidx) a.colidx[base] = gu;
                }
            }
        }
    }
    // a zero pivot: the system is exactly singular (padding rows pivot on 1)
    bool singular = false;
#pragma unroll
    for (int rr = 0; rr < ROWS; ++rr) singular |= pval[rr] == 0.0;
    const uint32_t sflag = group_max_u32<G>(singular ? 1u : 0u);
    const uint64_t badm = group_or_u64<G>(bad);
    if (a.status && live) {
        for (int s = t; s < ny; s += G) {
            const int yk = L.yrow[grp][s];
            const int st = sflag ? 1 : (((badm >> s) & 1u) ? 2 : 0);
            if (a.status_accumulate) a.status[yk] = max(a.status[yk], st);
            else a.status[yk] = st;
        }
    }
}

// ---------------------------------------------------------------- n + P > 128
// One wavefront per stencil; element (e, j) of the augmented system
// [A | B] at A[j * ld + e] in the task's slice of the workspace, so a column
// of the lanes' rows is one coalesced access.  Stencils of at most
// kMemMaxN nodes (the neighbour search's limit) and kMemMaxSys equations.
constexpr int kMemMaxN = 128;
constexpr int kMemMaxSys = 512;
constexpr int kMemWaves = 4;

__device__ __forceinline__ double mono_value(int q, double x, double y)
{
    int ea, eb;
    mono_exponents(q, ea, eb);
    double v = 1.0;
    for (int d = 0; d < ea; ++d) v *= x;
    for (int d = 0; d < eb; ++d) v *= y;
    return v;
}

template <int NW>
__global__ __launch_bounds__(64 * NW) void solve_mem_kernel(const SolveArgs a, long t0, long tcount)
{
    constexpr int RS = kMemMaxSys / 64;  // row slots per lane
    __shared__ double xi_s[NW][kMemMaxN][2];
    __shared__ int g_s[NW][kMemMaxN];
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const long local = (long)blockIdx.x * NW + w;
    if (local >= tcount || t0 + local >= a.ntasks) return;  // whole wavefront
    const long task = t0 + local;
    const long yk = a.rows ? (long)a.rows[task] : task;
    const long xs = a.y2x ? (long)a.y2x[yk] : yk;
    const int n = a.n, P = a.P, nsys = n + P, nops = a.nops, m = a.m;
    const int W = nsys + nops;
    const long ld = a.mem_ld;
    const bool method3 = a.method == 3;
    double *A = a.scratch + local * ld * W;

    // ---- frame (src/scalestencil.jl:10-21)
    const int c = max(a.stencils[xs * n], 0);
    double sx = 1.0, sy = 1.0, yx = kEps, yy = kEps;
    {
#pragma clang fp contract(off)
        const double cx = a.X[2 * (long)c], cy = a.X[2 * (long)c + 1];
        double mx = 0.0, my = 0.0;
        for (int i = lane; i < n; i += 64) {
            const int gi = max(a.stencils[xs * n + i], 0);
            const double dx = a.X[2 * (long)gi] - cx, dy = a.X[2 * (long)gi + 1] - cy;
            xi_s[w][i][0] = dx;
            xi_s[w][i][1] = dy;
            g_s[w][i] = gi;
            mx = fmax(mx, __builtin_fabs(dx));
            my = fmax(my, __builtin_fabs(dy));
        }
        if (!method3) {
            sx = 1.0 / wave_max_nonneg_f64(mx);
            sy = 1.0 / wave_max_nonneg_f64(my);
            yx = (a.Y[2 * yk] - cx) * sx;
            yy = (a.Y[2 * yk + 1] - cy) * sy;
        }
        for (int i = lane; i < n; i += 64) {
            if (method3) {
                if (i == 0) { xi_s[w][0][0] = kEps; xi_s[w][0][1] = kEps; }
            } else {
                xi_s[w][i][0] *= sx;
                xi_s[w][i][1] *= sy;
            }
        }
    }
    __builtin_amdgcn_wave_barrier();

    // ---- assembly
    for (int e = lane; e < nsys; e += 64) {
        if (e < n) {
            const double xx = xi_s[w][e][0], xy = xi_s[w][e][1];
            for (int j = 0; j < n; ++j)
                A[j * ld + e] = phs_r2(m, sq_sum(xx - xi_s[w][j][0], xy - xi_s[w][j][1]));
            for (int q = 0; q < P; ++q) A[(n + q) * ld + e] = mono_value(q, xx, xy);
            double ux, uy;
            {
#pragma clang fp contract(off)
                if (method3) {
                    ux = xx;
                    uy = xy;
                } else {
                    ux = yx - xx;
                    uy = yy - xy;
                    if (ux == 0.0) ux = kEps;
                    if (uy == 0.0) uy = kEps;
                }
            }
            const PhsAt q = phs_at(m, ux, uy);
            for (int o = 0; o < nops; ++o)
                A[(nsys + o) * ld + e] = op_phs_fast(a.col_kind[o], a.col_k[o], m, ux, uy, sx, sy, q);
        } else {
            const int q = e - n;
            int ea, eb;
            mono_exponents(q, ea, eb);
            for (int j = 0; j < n; ++j) A[j * ld + e] = mono_value(q, xi_s[w][j][0], xi_s[w][j][1]);
            for (int j = n; j < nsys; ++j) A[j * ld + e] = 0.0;
            for (int o = 0; o < nops; ++o)
                A[(nsys + o) * ld + e] = op_mono(a.col_kind[o], a.col_k[o], ea, eb, yx, yy, sx, sy);
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");

    // ---- Gauss-Jordan with partial pivoting (ties to the lower row)
    bool alive[RS];
    int kp[RS];
    double pv[RS];
#pragma unroll
    for (int r = 0; r < RS; ++r) {
        alive[r] = lane + 64 * r < nsys;
        kp[r] = nsys;
        pv[r] = 1.0;
    }
    bool singular = false;
    for (int k = 0; k < nsys; ++k) {
        const double *col = A + (long)k * ld;
        uint64_t bb = 0;
        int be = 0x7FFFFFFF;
#pragma unroll
        for (int r = 0; r < RS; ++r) {
            const int e = lane + 64 * r;
            if (alive[r]) {
                // live rows rank above used ones even at zero (bit 63)
                const uint64_t kb = (1ull << 63) | (uint64_t)__double_as_longlong(__builtin_fabs(col[e]));
                if (kb > bb) { bb = kb; be = e; }
            }
        }
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
            const uint64_t ob = (uint64_t)__shfl_xor((unsigned long long)bb, o, 64);
            const int oe = __shfl_xor(be, o, 64);
            if (ob > bb || (ob == bb && oe < be)) { bb = ob; be = oe; }
        }
        const int p = __builtin_amdgcn_readfirstlane(be);
        const double pk = col[p];
        if (pk == 0.0) { singular = true; break; }
#pragma unroll
        for (int r = 0; r < RS; ++r) {
            if (lane + 64 * r == p) { alive[r] = false; kp[r] = k; pv[r] = pk; }
        }
        const double rp = 1.0 / pk;
#pragma unroll
        for (int r = 0; r < RS; ++r) {
            const int e = lane + 64 * r;
            if (e < nsys && e != p) {
                const double l = -(col[e] * rp);
                if (l != 0.0)
                    for (int j = k + 1; j < W; ++j) A[(long)j * ld + e] = __builtin_fma(l, A[(long)j * ld + p], A[(long)j * ld + e]);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
    }

    // ---- output (sorted columns: the rank of each node's global index)
    __builtin_amdgcn_wave_barrier();
    uint32_t bad = 0;
    if (!singular) {
#pragma unroll
        for (int r = 0; r < RS; ++r) {
            const int u = kp[r];
            if (lane + 64 * r < nsys && u < n) {
                const int gu = g_s[w][u];
                int rk = 0;
                for (int j = 0; j < n; ++j) rk += g_s[w][j] < gu ? 1 : 0;
                const long base = yk * n + rk;
                if (a.colidx) a.colidx[base] = gu;
                const int e = lane + 64 * r;
                for (int o = 0; o < nops; ++o) {
                    const double f = method3 ? 1.0 : op_post_scale(a.col_kind[o], a.col_k[o], sx, sy);
                    const double wv = f * (A[(long)(nsys + o) * ld + e] / pv[r]);
                    if (!__builtin_isfinite(wv)) bad = 1;
                    a.vals[o][base] = wv;
                }
            }
        }
    }
    bad = wave_or_u32(bad);
    if (a.status && lane == 0) {
        const int st = singular ? 1 : (bad ? 2 : 0);
        if (a.status_accumulate) a.status[yk] = max(a.status[yk], st);
        else a.status[yk] = st;
    }
}

// ------------------------------------------------------------ launchers
// Systems above 128 equations, a.mem_tasks stencils per launch (the workspace
// holds that many).
inline void launch_gen_mem_impl(const SolveArgs &a, hipStream_t s)
{
    const long chunk = a.mem_tasks > 0 ? a.mem_tasks : 1;
    for (long t0 = 0; t0 < a.ntasks; t0 += chunk) {
        const long cnt = a.ntasks - t0 < chunk ? a.ntasks - t0 : chunk;
        const unsigned blocks = (unsigned)((cnt + kMemWaves - 1) / kMemWaves);
        solve_mem_kernel<kMemWaves><<<blocks, 64 * kMemWaves, 0, s>>>(a, t0, cnt);
    }
}

template <int NSB, int R, int G>
void launch_gen_t(const SolveArgs &a, hipStream_t s)
{
    constexpr int per_block = 64 / G;
    const unsigned blocks = (unsigned)((a.ntasks + per_block - 1) / per_block);
    solve_gen_kernel<NSB, R, G, 1><<<blocks, 64, 0, s>>>(a);
}

// Right-hand-side widths compiled per padded size: 3 and 6 (one output row per
// task), then 12, 24, 36, 48, 72 up to RB (several rows); the launch takes the
// narrowest that holds R.
// Up to kGenNarrowRhs columns, sizes with a narrow layout (GN lanes, two row
// slots per lane) take it: twice the stencils per wavefront share each step's
// pivot search and LDS round trip.
template <int NSB, int G, int GN, int RB, int RC>
void launch_gen_width(const SolveArgs &a, int R, hipStream_t s)
{
    if constexpr (RC <= RB) {
        if (R <= RC) {
            if constexpr (GN > 0 && RC <= kGenNarrowRhs) launch_gen_t<NSB, RC, GN>(a, s);
            else launch_gen_t<NSB, RC, G>(a, s);
            return;
        }
        constexpr int NEXT = RC == 3 ? 6 : (RC == 6 ? 12 : (RC == 12 ? 24 : (RC == 24 ? 36 : (RC == 36 ? 48 : 72))));
        launch_gen_width<NSB, G, GN, RB, NEXT>(a, R, s);
    } else {
        throw Error{2, "too many right-hand sides for the generic kernel"};
    }
}

template <int NSB, int G, int GN, int RB>
void launch_gen_bucket(const SolveArgs &a, int R, hipStream_t s)
{
    launch_gen_width<NSB, G, GN, RB, 3>(a, R, s);
}

}  // namespace rbffd

// One translation unit per padded size keeps the build parallel.
#define RBFFD_DEFINE_GEN(nsb)                                                          \
    namespace rbffd {                                                                  \
    void launch_gen_##nsb(const SolveArgs &a, int R, hipStream_t s)                    \
    {                                                                                  \
        launch_gen_bucket<nsb, gen_lanes(nsb), gen_narrow_lanes(nsb), gen_rb(nsb)>(a, R, s);  \
    }                                                                                  \
    }
