// stencil_gen2.hpp -- K2 for systems of 65 to 128 equations: two wavefronts
// per stencil, one row of [A | B] per lane.
//
// Replaces hot loops 1 and 2 of the reference (src/generate_operator.jl:55-166,
// method 3 at :388-467; src/hyperviscosity_operator.jl:66-161) like the
// generic kernel (stencil_gen.hpp), whose padding, tables, right-hand-side
// columns and output it shares.  In one wavefront such a system needs two row
// slots per lane, 2 (NSB + R) doubles of registers: more than a lane has, so
// the one-wavefront kernels spill and run one wavefront per SIMD.  Here the
// 128 lanes of a workgroup hold one row each.  An elimination step costs two
// workgroup barriers: one after each wavefront has posted its best pivot
// candidate, one after the winning lane has published its row (and the
// reciprocal of its pivot) to LDS; each wavefront then spreads the row over
// its 16-lane DPP rows as the other kernels do.  (Publishing both
// wavefronts' best rows and keeping one barrier per step was slower, 749
// against 585 ms for 4M stencils of 85 equations: the row writes, not the
// barriers, bound the step.)
#pragma once
#include "stencil_gen.hpp"

namespace rbffd {

template <int NSB, int R>
struct Gen2Shape {
    static constexpr int WIDTH = NSB + R;
    static constexpr int SB = 4;  // steps per block of equal static width
    static constexpr int UBUF = 16 * ((WIDTH + 15) / 16);
    static constexpr int PM = NSB * NSB / 4;
    static constexpr int DEGB = mono_deg(NSB - 2);
    static constexpr int QB = (DEGB + 1) * (DEGB + 2) / 2;
};

template <int NSB, int R>
struct alignas(16) Gen2LDS {
    using T = Gen2Shape<NSB, R>;
    double u[T::UBUF + 2];  // the current pivot row
    double rp;              // its reciprocal pivot
    uint32_t key[2];        // each wavefront's best pivot candidate
    double xi[NSB][2];
    double pm[T::PM];
    double ys[R][2];
    double red[4];          // cross-wavefront maxima of the frame
    uint32_t flag[2];
    uint64_t bad[2];
    int g[NSB];
    int rank[NSB];
    int yrow[R];
};

template <int NSB, int R>
__global__ __launch_bounds__(128) void solve_gen2_kernel(const SolveArgs a)
{
    using T = Gen2Shape<NSB, R>;
    constexpr int WIDTH = T::WIDTH;
    static_assert(NSB <= 128, "row ids take 7 bits of the pivot key");
    __shared__ Gen2LDS<NSB, R> L;

    const int t = threadIdx.x;  // row and node slot, 0..127
    const int lane = t & 63;
    const int wv = t >> 6;
    long ntasks = a.ntasks;
    if (a.ntasks_dev) ntasks = min(ntasks, (long)*a.ntasks_dev);
    const long task = blockIdx.x;
    if (task >= ntasks) return;  // whole workgroup

    const int n = a.n, P = a.P, nsys = n + P, m = a.m, nops = a.nops;
    const bool method3 = a.method == 3;
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
    const int ncols = ny * nops;

    // ---- stencil frame (src/scalestencil.jl:10-21, src/generate_operator.jl:100-134)
    const int c = max(a.stencils[xs * n], 0);
    const int gid = t < n ? max(a.stencils[xs * n + t], 0) : c;
    double sx = 1.0, sy = 1.0, xix, xiy;
    {
#pragma clang fp contract(off)
        const double cx = a.X[2 * (long)c], cy = a.X[2 * (long)c + 1];
        const double dx = a.X[2 * (long)gid] - cx, dy = a.X[2 * (long)gid + 1] - cy;
        if (!method3) {
            const double mx = group_max_nonneg_f64<64>(__builtin_fabs(dx));
            const double my = group_max_nonneg_f64<64>(__builtin_fabs(dy));
            if (lane == 0) {
                L.red[2 * wv] = mx;
                L.red[2 * wv + 1] = my;
            }
            __syncthreads();
            sx = 1.0 / fmax(L.red[0], L.red[2]);
            sy = 1.0 / fmax(L.red[1], L.red[3]);
            xix = dx * sx;
            xiy = dy * sy;
        } else {
            // src/generate_operator.jl:393-410: centre := (eps, eps)
            xix = t == 0 ? kEps : dx;
            xiy = t == 0 ? kEps : dy;
        }
        if (t < n) {
            L.xi[t][0] = xix;
            L.xi[t][1] = xiy;
            L.g[t] = gid;
            double pxp[T::DEGB + 1], pyp[T::DEGB + 1];
            pxp[0] = 1.0;
            pyp[0] = 1.0;
#pragma unroll
            for (int d = 1; d <= T::DEGB; ++d) { pxp[d] = pxp[d - 1] * xix; pyp[d] = pyp[d - 1] * xiy; }
#pragma unroll
            for (int q = 0; q < T::QB; ++q)
                if (q < P) L.pm[t * P + q] = pxp[mono_ea(q)] * pyp[mono_eb(q)];
        }
        for (int s = t; s < ny; s += 128) {
            const long yk = a.grp_x ? (long)a.ylist[ybase + s] : ybase;
            L.yrow[s] = (int)yk;
            if (method3) {
                L.ys[s][0] = kEps;
                L.ys[s][1] = kEps;
            } else {
                L.ys[s][0] = (a.Y[2 * yk] - cx) * sx;
                L.ys[s][1] = (a.Y[2 * yk + 1] - cy) * sy;
            }
        }
    }
    __syncthreads();
    if (t < n) {
        int rk = 0;
        for (int j = 0; j < n; ++j) rk += L.g[j] < gid ? 1 : 0;
        L.rank[t] = rk;
    }

    // ---- right-hand sides (columns as in the generic kernel)
    double rhs[R];
#pragma unroll
    for (int cc = 0; cc < R; ++cc) rhs[cc] = 0.0;
    if (t < n) {
        double ux = 0.0, uy = 0.0;
        PhsAt q{};
#pragma unroll
        for (int cc = 0; cc < R; ++cc) {
            if (cc < ncols) {
                if (a.col_op[cc] == 0) {
#pragma clang fp contract(off)
                    if (method3) {
                        ux = xix;
                        uy = xiy;
                    } else {
                        ux = L.ys[a.col_sub[cc]][0] - xix;
                        uy = L.ys[a.col_sub[cc]][1] - xiy;
                        if (ux == 0.0) ux = kEps;
                        if (uy == 0.0) uy = kEps;
                    }
                    q = phs_at(m, ux, uy);
                }
                rhs[cc] = op_phs_fast(a.col_kind[cc], a.col_k[cc], m, ux, uy, sx, sy, q);
            }
        }
    } else if (t < nsys) {
        int ea, eb;
        mono_exponents(t - n, ea, eb);
#pragma unroll
        for (int cc = 0; cc < R; ++cc)
            if (cc < ncols)
                rhs[cc] = op_mono(a.col_kind[cc], a.col_k[cc], ea, eb, L.ys[a.col_sub[cc]][0],
                                  L.ys[a.col_sub[cc]][1], sx, sy);
    }

    // ---- row t of [A | B]: w columns, lambda columns, padding
    double row[WIDTH];
    if (t < n) {
#pragma unroll
        for (int j = 0; j < NSB; ++j) {
            double v = 0.0;
            if (j < n) v = phs_r2(m, sq_sum(xix - L.xi[j][0], xiy - L.xi[j][1]));
            else if (j < nsys) v = L.pm[t * P + (j - n)];
            row[j] = v;
        }
    } else if (t < nsys) {
#pragma unroll
        for (int j = 0; j < NSB; ++j) row[j] = j < n ? L.pm[j * P + (t - n)] : 0.0;
    } else {
#pragma unroll
        for (int j = 0; j < NSB; ++j) row[j] = j == t ? 1.0 : 0.0;
    }
#pragma unroll
    for (int cc = 0; cc < R; ++cc) row[NSB + cc] = rhs[cc];

    // ---- Gauss-Jordan with partial pivoting; rows slide one register per step
    uint32_t alive = t < NSB ? 0xFFFFFFFFu : 0u;
    int kp = NSB;
    double pval = 1.0;
    uint32_t key;
    double crp;
    auto candidate = [&]() {
        key = ((__float_as_uint((float)row[0]) & 0x7FFFFF00u) | (0x80u | (uint32_t)(127 - t))) & alive;
        double r = __builtin_amdgcn_rcp(row[0]);
        r = __builtin_fma(r, __builtin_fma(-row[0], r, 1.0), r);
        crp = __builtin_fma(r, __builtin_fma(-row[0], r, 1.0), r);
    };
    candidate();
    uint32_t kloc = group_max_u32<64>(key);
    int k = 0;
    auto step = [&](auto wbc) {
        constexpr int WB = decltype(wbc)::value;
        constexpr int NPV = (WB + 15) / 16;
        // the two wavefronts' best candidates (0, which always loses, from a
        // wavefront without a live row)
        if (lane == 0) L.key[wv] = kloc;
        __syncthreads();
        const uint32_t kmax = max(L.key[0], L.key[1]);
        if (t == 127 - (int)(kmax & 127u)) {
#pragma unroll
            for (int j = 0; j < WB; ++j) L.u[j] = row[j];
            L.rp = crp;
            alive = 0u;
            kp = k;
            pval = row[0];
            row[0] = 0.0;  // a zero multiplier keeps the pivot row as it is
        }
        __syncthreads();
        double pv[NPV];
#pragma unroll
        for (int i = 0; i < NPV; ++i) pv[i] = L.u[(lane & 15) + 16 * i];
        const double rp = L.rp, p1 = L.u[1];
        const double nl = -(row[0] * rp);
        row[0] = fma3(nl, p1, row[1]);
        candidate();
        kloc = group_max_u32<64>(key);
        // columns j = 1 .. WB-2 get column j+1 of the pivot row, in fours
        constexpr int JE = WB - 1;
        constexpr int NQ = (JE - 1) / 4;
        for_blocks<0, NQ>([&](auto qc) {
            constexpr int j = 1 + 4 * decltype(qc)::value;
            double q0, q1, q2, q3;
            row_bcast64x4<(j + 1) % 16, (j + 2) % 16, (j + 3) % 16, (j + 4) % 16>(
                pv[(j + 1) / 16], pv[(j + 2) / 16], pv[(j + 3) / 16], pv[(j + 4) / 16], q0, q1, q2, q3);
            row[j] = fma3(nl, q0, row[j + 1]);
            row[j + 1] = fma3(nl, q1, row[j + 2]);
            row[j + 2] = fma3(nl, q2, row[j + 3]);
            row[j + 3] = fma3(nl, q3, row[j + 4]);
        });
        for_blocks<1 + 4 * NQ, JE>([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            row[j] = fma3(nl, row_bcast64<(j + 1) % 16>(pv[(j + 1) / 16]), row[j + 1]);
        });
    };
    for_blocks<0, (NSB + T::SB - 1) / T::SB>([&](auto bc) {
        constexpr int kb = T::SB * decltype(bc)::value;
        const int kend = NSB < kb + T::SB ? NSB : kb + T::SB;
#pragma nounroll
        for (; k < kend; ++k) step(std::integral_constant<int, WIDTH - kb>{});
    });

    // ---- output (as the generic kernel)
    uint64_t bad = 0;
    if (kp < n) {
        const int gu = L.g[kp];
        const int rku = L.rank[kp];
        const double rd = 1.0 / pval;
#pragma unroll
        for (int cc = 0; cc < R; ++cc) {
            if (cc < ncols) {
                const int s = a.col_sub[cc], o = a.col_op[cc];
                const long base = (long)L.yrow[s] * n + rku;
                const double f = method3 ? 1.0 : op_post_scale(a.col_kind[cc], a.col_k[cc], sx, sy);
                const double wv2 = f * (row[cc] * rd);
                if (!__builtin_isfinite(wv2)) bad |= 1ull << s;
                a.vals[o][base] = wv2;
                if (o == 0 && a.colidx) a.colidx[base] = gu;
            }
        }
    }
    const uint32_t sw = group_max_u32<64>(pval == 0.0 ? 1u : 0u);
    const uint64_t bw = group_or_u64<64>(bad);
    if (lane == 0) {
        L.flag[wv] = sw;
        L.bad[wv] = bw;
    }
    __syncthreads();
    if (a.status) {
        const uint32_t sflag = L.flag[0] | L.flag[1];
        const uint64_t badm = L.bad[0] | L.bad[1];
        for (int s = t; s < ny; s += 128) {
            const int yk = L.yrow[s];
            const int st = sflag ? 1 : (((badm >> s) & 1u) ? 2 : 0);
            if (a.status_accumulate) a.status[yk] = max(a.status[yk], st);
            else a.status[yk] = st;
        }
    }
}

template <int NSB, int R>
void launch_gen2_t(const SolveArgs &a, hipStream_t s)
{
    solve_gen2_kernel<NSB, R><<<(unsigned)a.ntasks, 128, 0, s>>>(a);
}

// Right-hand-side widths 3, 6, 12 and 24 (several output rows per task).
template <int NSB>
void launch_gen2_bucket(const SolveArgs &a, int R, hipStream_t s)
{
    if (R <= 3) launch_gen2_t<NSB, 3>(a, s);
    else if (R <= 6) launch_gen2_t<NSB, 6>(a, s);
    else if (R <= 12) launch_gen2_t<NSB, 12>(a, s);
    else if (R <= 24) launch_gen2_t<NSB, 24>(a, s);
    else throw Error{2, "too many right-hand sides for the two-wavefront kernel"};
}

}  // namespace rbffd
