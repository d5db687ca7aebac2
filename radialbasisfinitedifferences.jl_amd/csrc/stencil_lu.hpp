// stencil_lu.hpp -- K2, FP64 path: batched assembly and solve of the local
// RBF-FD systems by Gauss-Jordan elimination with partial pivoting, G lanes
// per stencil.
//
// Replaces hot loops 1 and 2 of the reference (src/generate_operator.jl:55-65
// and :89-166, method 3 at :388-467; src/hyperviscosity_operator.jl:66-161).
//
// Layout.  A wavefront holds 64/G stencils; each lane of a group holds
// ROWS = ceil(n / G) row slots of its stencil's saddle-point system
//     A = [Phi P; P^T 0],  Phi_ij = |xi_i - xi_j|^m,  P_iq = x^a y^b (xi_i),
// each row extended by its right-hand-side entries, all in registers.  The
// lambda columns go first (highest degree first): the first NP steps run on
// the RBF equations alone, then the monomial equations take the slots of the
// NP lambda pivot rows (which only determine lambda) and of the empty slots.
// Elimination: a DPP max over the group picks the pivot row, the pivot lane
// publishes that row (and its reciprocal pivot) to LDS, and every lane
// eliminates column k from all of its other rows, earlier pivot rows included;
// each lane reads every 16th pivot entry and a DPP broadcast feeds ROWS FMAs
// per column.  With 64 lanes (one stencil per wavefront) the pivot lane is
// wavefront-uniform and the lanes read the pivot row from its registers with
// v_readlane instead.  The pivot search for step k+1 runs as soon as column
// k+1 is updated, overlapping the rest of step k.  Rows never move (the
// permutation stays implicit), and a column leaves the registers once it is
// eliminated.  The layout updates every row slot at every step anyway, so
// eliminating above the pivot (Gauss-Jordan) costs no extra instructions, and
// it leaves each row with pivot * x in its right-hand-side registers: no back
// substitution.  On the local systems the forward error matches an LU
// solve's to a small factor (about 1e-12 at n = 31 and 4e-11 at n = 57
// against a 40-digit solve).
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

// elimination steps per block of equal static row width: shorter blocks waste
// fewer columns per step and cost code size (the 16-lane kernel at n = 31 is
// 56 KB with 2, inside the 64 KB instruction cache)
template <int G>
constexpr int step_block() { return G <= 32 ? 2 : 4; }

// a * b + c into a register distinct from c's.  The compiler's own lowering
// of the sliding update (row[j] = fma(., ., row[j+1]) in a loop) is a copy
// plus the two-address v_fmac; the three-address form needs no copy.  Register
// operands only, no memory access.
__device__ __forceinline__ double fma3(double a, double b, double c)
{
    double d;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));
    return d;
}

// Lane `L` of each 16-lane DPP row, in every lane of that row (64-bit DPP
// move, row_newbcast).  A DPP read needs two wait states after a VALU write
// of its source, and the compiler does not see inside the asm: the s_nop
// provides them whatever wrote the source.
template <int L>
__device__ __forceinline__ double row_bcast64(double v)
{
    double d;
    asm("s_nop 1\n\tv_mov_b64_dpp %0, %1 row_newbcast:%2 row_mask:0xf bank_mask:0xf" : "=v"(d) : "v"(v), "i"(L));
    return d;
}

// Two such broadcasts behind one wait: nothing inside the asm writes the
// sources, and the outputs are early-clobber so neither aliases the second
// source.
template <int L0, int L1>
__device__ __forceinline__ void row_bcast64x2(double v0, double v1, double &d0, double &d1)
{
    asm("s_nop 1\n\tv_mov_b64_dpp %0, %2 row_newbcast:%4 row_mask:0xf bank_mask:0xf\n\t"
        "v_mov_b64_dpp %1, %3 row_newbcast:%5 row_mask:0xf bank_mask:0xf"
        : "=&v"(d0), "=&v"(d1)
        : "v"(v0), "v"(v1), "i"(L0), "i"(L1));
}

template <int L0, int L1, int L2, int L3>
__device__ __forceinline__ void row_bcast64x4(double v0, double v1, double v2, double v3, double &d0, double &d1,
                                              double &d2, double &d3)
{
    asm("s_nop 1\n\tv_mov_b64_dpp %0, %4 row_newbcast:%8 row_mask:0xf bank_mask:0xf\n\t"
        "v_mov_b64_dpp %1, %5 row_newbcast:%9 row_mask:0xf bank_mask:0xf\n\t"
        "v_mov_b64_dpp %2, %6 row_newbcast:%10 row_mask:0xf bank_mask:0xf\n\t"
        "v_mov_b64_dpp %3, %7 row_newbcast:%11 row_mask:0xf bank_mask:0xf"
        : "=&v"(d0), "=&v"(d1), "=&v"(d2), "=&v"(d3)
        : "v"(v0), "v"(v1), "v"(v2), "v"(v3), "i"(L0), "i"(L1), "i"(L2), "i"(L3));
}

// The sliding update of four columns in place with scalar pivot entries:
// r0 <- a s0 + r1, r1 <- a s1 + r2, r2 <- a s2 + r3, r3 <- a s3 + r4, in
// that order (each FMA reads its source before the next one overwrites it).
// The scalar operands come from v_readlane, a VALU write of a scalar
// register, which a VALU read may use only two wait states later; the
// compiler does not see inside the asm, so the s_nop provides them.
__device__ __forceinline__ void slide4_s(double &r0, double &r1, double &r2, double &r3, double r4, double a,
                                         double s0, double s1, double s2, double s3)
{
    asm("s_nop 1\n\t"
        "v_fma_f64 %0, %4, %5, %1\n\t"
        "v_fma_f64 %1, %4, %6, %2\n\t"
        "v_fma_f64 %2, %4, %7, %3\n\t"
        "v_fma_f64 %3, %4, %8, %9"
        : "+v"(r0), "+v"(r1), "+v"(r2), "+v"(r3)
        : "v"(a), "s"(s0), "s"(s1), "s"(s2), "s"(s3), "v"(r4));
}

__device__ __forceinline__ void slide1_s(double &r0, double r1, double a, double s0)
{
    asm("s_nop 1\n\tv_fma_f64 %0, %1, %2, %3" : "+v"(r0) : "v"(a), "s"(s0), "v"(r1));
}

template <int B, int NB, class F>
__device__ __forceinline__ void for_blocks(F &&f)
{
    if constexpr (B < NB) {
        f(std::integral_constant<int, B>{});
        for_blocks<B + 1, NB>(f);
    }
}

// Maximum over each aligned group of G lanes; every lane gets its group's value.
// DPP inside each 16-lane row, then lane swaps across rows and halves
// (no LDS or scalar round trip).
template <int G>
__device__ __forceinline__ uint32_t group_max_u32(uint32_t v)
{
    static_assert(G == 16 || G == 32 || G == 64, "group size");
    v = max(v, dpp_u32<0xB1>(v));   // quad_perm [1,0,3,2]
    v = max(v, dpp_u32<0x4E>(v));   // quad_perm [2,3,0,1]
    v = max(v, dpp_u32<0x141>(v));  // row_half_mirror
    v = max(v, dpp_u32<0x140>(v));  // row_mirror
    if constexpr (G >= 32) {
        // v_permlane16_swap exchanges the odd 16-lane rows of one copy with the
        // even rows of the other: the two results hold rows (0, 0, 2, 2) and
        // (1, 1, 3, 3), so their maximum is each 32-lane half's maximum
        const auto sw = __builtin_amdgcn_permlane16_swap(v, v, false, false);
        v = max((uint32_t)sw[0], (uint32_t)sw[1]);
    }
    if constexpr (G == 64) {
        // v_permlane32_swap likewise gives halves (low, low) and (high, high)
        const auto sw = __builtin_amdgcn_permlane32_swap(v, v, false, false);
        v = max((uint32_t)sw[0], (uint32_t)sw[1]);
    }
    return v;
}

// Bitwise OR over each 16-lane DPP row, in every lane of the row: with one
// lane holding a value and the others zero, a broadcast of that lane's bits.
__device__ __forceinline__ uint64_t row_or_u64(uint64_t b)
{
    uint32_t lo = (uint32_t)b, hi = (uint32_t)(b >> 32);
    lo |= dpp_u32<0xB1>(lo);   // quad_perm [1,0,3,2]
    hi |= dpp_u32<0xB1>(hi);
    lo |= dpp_u32<0x4E>(lo);   // quad_perm [2,3,0,1]
    hi |= dpp_u32<0x4E>(hi);
    lo |= dpp_u32<0x141>(lo);  // row_half_mirror
    hi |= dpp_u32<0x141>(hi);
    lo |= dpp_u32<0x140>(lo);  // row_mirror
    hi |= dpp_u32<0x140>(hi);
    return ((uint64_t)hi << 32) | lo;
}

// Bitwise OR over each aligned group of G lanes, in every lane of the group:
// DPP inside 16-lane rows, then the lane swaps across rows and halves.
template <int G>
__device__ __forceinline__ uint64_t group_or_u64(uint64_t b)
{
    b = row_or_u64(b);
    uint32_t lo = (uint32_t)b, hi = (uint32_t)(b >> 32);
    if constexpr (G >= 32) {
        const auto sl = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
        const auto sh = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
        lo = (uint32_t)sl[0] | (uint32_t)sl[1];
        hi = (uint32_t)sh[0] | (uint32_t)sh[1];
    }
    if constexpr (G == 64) {
        const auto sl = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
        const auto sh = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
        lo = (uint32_t)sl[0] | (uint32_t)sl[1];
        hi = (uint32_t)sh[0] | (uint32_t)sh[1];
    }
    return ((uint64_t)hi << 32) | lo;
}

// Maximum of non-negative doubles over each aligned group of G lanes (their
// bit patterns order like unsigned integers): DPP inside 16-lane rows, then
// lane swaps across rows and halves, as group_max_u32.
__device__ __forceinline__ uint64_t max_u64(uint64_t a, uint64_t b) { return a > b ? a : b; }

template <int G>
__device__ __forceinline__ double group_max_nonneg_f64(double v)
{
    uint64_t b = (uint64_t)__double_as_longlong(v);
    auto dpp_step = [&](auto ctrl) {
        constexpr int C = decltype(ctrl)::value;
        const uint32_t lo = dpp_u32<C>((uint32_t)b), hi = dpp_u32<C>((uint32_t)(b >> 32));
        b = max_u64(b, ((uint64_t)hi << 32) | lo);
    };
    dpp_step(std::integral_constant<int, 0xB1>{});   // quad_perm [1,0,3,2]
    dpp_step(std::integral_constant<int, 0x4E>{});   // quad_perm [2,3,0,1]
    dpp_step(std::integral_constant<int, 0x141>{});  // row_half_mirror
    dpp_step(std::integral_constant<int, 0x140>{});  // row_mirror
    if constexpr (G >= 32) {
        const auto lo = __builtin_amdgcn_permlane16_swap((uint32_t)b, (uint32_t)b, false, false);
        const auto hi = __builtin_amdgcn_permlane16_swap((uint32_t)(b >> 32), (uint32_t)(b >> 32), false, false);
        b = max_u64(((uint64_t)(uint32_t)hi[0] << 32) | (uint32_t)lo[0], ((uint64_t)(uint32_t)hi[1] << 32) | (uint32_t)lo[1]);
    }
    if constexpr (G == 64) {
        const auto lo = __builtin_amdgcn_permlane32_swap((uint32_t)b, (uint32_t)b, false, false);
        const auto hi = __builtin_amdgcn_permlane32_swap((uint32_t)(b >> 32), (uint32_t)(b >> 32), false, false);
        b = max_u64(((uint64_t)(uint32_t)hi[0] << 32) | (uint32_t)lo[0], ((uint64_t)(uint32_t)hi[1] << 32) | (uint32_t)lo[1]);
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

// a^2 + b^2 without contraction, so the matrix entries and the identity
// operator's right-hand side round alike
__device__ __forceinline__ double sq_sum(double a, double b)
{
#pragma clang fp contract(off)
    return a * a + b * b;
}

// |v|^m for the matrix block: odd m takes one square root
__device__ __forceinline__ double phs_r2(int m, double r2)
{
    double p = upow(r2, m >> 1);
    return (m & 1) ? p * __builtin_sqrt(r2) : p;
}

// Powers of r = |v| shared by the closed-form operators at one point v.
struct PhsAt {
    double r2, rm2, rm4;  // r^2, r^(m-2), r^(m-4)
};

__device__ __forceinline__ PhsAt phs_at(int m, double vx, double vy)
{
    PhsAt q;
    q.r2 = sq_sum(vx, vy);
    q.rm2 = (m >= 2) ? phs_r2(m - 2, q.r2) : phs_r2(m, q.r2) / q.r2;
    q.rm4 = q.rm2 / q.r2;
    return q;
}

// Operator applied to the PHS at v: closed forms for the first- and
// second-order operators, the general series for the others.
__device__ __forceinline__ double op_phs_fast(int op, int k, int m, double vx, double vy, double sx, double sy,
                                              const PhsAt &q)
{
    // no contraction: the value does not depend on how the caller's code
    // around it lets the compiler fuse multiplies and adds (rows solved alone
    // and rows solved in a group get the same bits)
#pragma clang fp contract(off)
    if (op > OP_DXY && op != OP_LAP) return op_phs(op, k, m, vx, vy, sx, sy);
    // E uses the matrix entries' own formula: for collocated rows b = A e_c holds
    // bit for bit, so the solve returns the identity row to rounding
    if (op == OP_E) return phs_r2(m, q.r2);
    const double dm = (double)m;
    switch (op) {
    case OP_DX:  return dm * vx * q.rm2;
    case OP_DY:  return dm * vy * q.rm2;
    case OP_DXX: return dm * q.rm2 + dm * (dm - 2.0) * vx * vx * q.rm4;
    case OP_DYY: return dm * q.rm2 + dm * (dm - 2.0) * vy * vy * q.rm4;
    case OP_DXY: return dm * (dm - 2.0) * vx * vy * q.rm4;
    default: {  // OP_LAP with per-axis scales
        const double fxx = dm * q.rm2 + dm * (dm - 2.0) * vx * vx * q.rm4;
        const double fyy = dm * q.rm2 + dm * (dm - 2.0) * vy * vy * q.rm4;
        return sx * sx * fxx + sy * sy * fyy;
    }
    }
}

__device__ __forceinline__ double op_phs_fast(int op, int k, int m, double vx, double vy, double sx, double sy)
{
    return op_phs_fast(op, k, m, vx, vy, sx, sy, phs_at(m, vx, vy));
}

template <int NN, int NP, int R, int G>
struct LuShape {
    static constexpr int NS = NN + NP;
    // Register row slots per lane: the RBF equations (slot rr of lane t holds
    // equation t + G rr, this lane's node rr).  The lambda columns are
    // eliminated first, highest degree first, and only the RBF equations are
    // nonzero there: the first NP steps run on them alone.  A lambda pivot row
    // is not needed after those steps (it only gives lambda), so the monomial
    // equations then take the slots of the NP lambda pivots (and of empty
    // slots), and the remaining NN steps run on the same ROWS slots.
    static constexpr int ROWS = (NN + G - 1) / G;
    static constexpr int WIDTH = NS + R;
    static constexpr int S = 64 / G;               // stencils per wavefront
    static constexpr int NODES = ROWS;             // stencil nodes per lane
    static constexpr int QR = (NP + G - 1) / G;    // monomial equations per lane during assembly
    static constexpr int SB = step_block<G>();
    // the current pivot row: one buffer, reused by every step (LDS serves a
    // wavefront's operations in order, so a step's reads finish before the
    // next step's writes).  A step writes up to SB-1 stale values past its
    // row and reads the row in whole runs of 16.
    static constexpr int UBUF = 16 * ((WIDTH + 15) / 16);
    // during assembly: Phi as a packed upper triangle and the scaled node
    // coordinates (x, y) pairs.  Once the rows are built, P (node-major) and
    // the monomial equations' right-hand sides (NP x R) go over the dead
    // triangle past the pivot-row buffer, to be read when the monomial
    // equations are loaded.
    static constexpr int PHI0 = 0;
    static constexpr int XI0 = PHI0 + NN * (NN + 1) / 2;
    static constexpr int PB0 = UBUF;
    static constexpr int MR0 = PB0 + NN * NP;
    static constexpr int ASM = cmax_<XI0 + 2 * NN, MR0 + NP * R>::v;
    // Phi(a, b), a <= b, at PHI0 + tri(a) + b
    static constexpr __host__ __device__ int tri(int a) { return a * (2 * NN - a + 1) / 2 - a; }
    // even, so every group's row buffer is 16-byte aligned; +2 keeps the
    // groups' slots off one LDS bank
    static constexpr int USTRIDE = ((cmax_<UBUF, ASM>::v + 1) & ~1) + 2;
};

template <int NN, int NP, int R, int G>
struct alignas(16) LuLDS {
    using T = LuShape<NN, NP, R, G>;
    double u[T::S][T::USTRIDE];  // Phi, xi and P during assembly, then the pivot row (and P)
    double rp[T::S];              // reciprocal of the current pivot
    int g[T::S][NN];
    int rank[T::S][NN];  // rank of node i's global index inside the stencil
};

// Grouped tasks: the Y points of a task's output rows in the stencil frame,
// and the rows themselves.
template <int R, int S>
struct LuRows {
    double ys[S][R][2];
    int yrow[S][R];
};

// GRP: a task is one stencil with the Y rows that use it (a.grp_x,
// a.grp_start, a.grp_count, a.ylist); right-hand-side column cc belongs to
// output row a.col_sub[cc] of the task and operator a.col_op[cc], and every
// column sees the same operations wherever it sits, so a row's weights are
// those of a task of its own, bit for bit.  Else one output row per task.
template <int NN, int NP, int R, int G, int WPB, int MINW, bool GRP = false>
__global__ __launch_bounds__(64 * WPB, MINW) void solve_lu_kernel(const SolveArgs a)
{
    using T = LuShape<NN, NP, R, G>;
    constexpr int NS = T::NS, ROWS = T::ROWS, WIDTH = T::WIDTH, S = T::S, NODES = T::NODES;
    static_assert(NS <= 128, "system size above 128");
    static_assert(!GRP || R <= 64, "one status bit per output row of a task");
    __shared__ LuLDS<NN, NP, R, G> lds_all[WPB];
    __shared__ LuRows<R, S> rows_all[GRP ? WPB : 1];

    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int grp = lane / G;
    const int t = lane % G;
    auto &L = lds_all[w];
    long ntasks = a.ntasks;
    if (GRP && a.ntasks_dev) ntasks = min(ntasks, (long)*a.ntasks_dev);
    // every lane of the wavefront stays until the end (shuffles span the wave);
    // a group without a task works on a copy of task 0 and stores nothing
    const long task0 = ((long)blockIdx.x * WPB + w) * S;
    if (task0 >= ntasks) return;  // whole wavefront
    long task = task0 + grp;
    const bool live = task < ntasks;
    if (!live) task = task0;

    long yk, xs, ybase = 0;
    int ny = 1;
    if constexpr (GRP) {
        xs = a.grp_x[task];
        ybase = a.grp_start[task];
        ny = a.grp_count[task];
        yk = a.ylist[ybase];
    } else {
        yk = a.rows ? (long)a.rows[task] : task;
        xs = a.y2x ? (long)a.y2x[yk] : yk;
    }
    // right-hand-side columns in use (at most R: the launcher's choice)
    const int ncols = GRP ? ny * a.nops : a.nops;
    const int m = a.m;

    // ---- stencil frame (src/scalestencil.jl:10-21, src/generate_operator.jl:100-134)
    int gid[NODES];
    double dxn[NODES], dyn[NODES];
    // a failed neighbour search leaves -1 entries; clamp so every read stays in X
    const int c = max(a.stencils[xs * NN], 0);
    double sx = 1.0, sy = 1.0, yx, yy;
    {
#pragma clang fp contract(off)
        const double cx = a.X[2 * (long)c], cy = a.X[2 * (long)c + 1];
        double mx = 0.0, my = 0.0;
#pragma unroll
        for (int r = 0; r < NODES; ++r) {
            const int i = t + G * r;
            gid[r] = i < NN ? max(a.stencils[xs * NN + i], 0) : c;
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
        if constexpr (GRP) {
            // the task's Y points in the stencil frame (src/generate_operator.jl:112-124);
            // method 3 evaluates the polynomial right-hand side at the centre (:429-430)
            for (int q = t; q < ny; q += G) {
                const long yq = a.ylist[ybase + q];
                rows_all[w].yrow[grp][q] = (int)yq;
                rows_all[w].ys[grp][q][0] = a.method == 3 ? kEps : (a.Y[2 * yq] - cx) * sx;
                rows_all[w].ys[grp][q][1] = a.method == 3 ? kEps : (a.Y[2 * yq + 1] - cy) * sy;
            }
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
                L.u[grp][T::XI0 + 2 * i] = xi_x;
                L.u[grp][T::XI0 + 2 * i + 1] = xi_y;
                L.g[grp][i] = gid[r];
            }
        }
        // each node's output column: the rank of its global index inside the
        // stencil (CSR rows come out with sorted columns, as sparse() sorts)
        __builtin_amdgcn_wave_barrier();
        int rk[NODES];
#pragma unroll
        for (int r = 0; r < NODES; ++r) rk[r] = 0;
#pragma unroll
        for (int j = 0; j < NN; ++j) {
            const int gj = L.g[grp][j];
#pragma unroll
            for (int r = 0; r < NODES; ++r) rk[r] += gj < gid[r] ? 1 : 0;
        }
#pragma unroll
        for (int r = 0; r < NODES; ++r)
            if (t + G * r < NN) L.rank[grp][t + G * r] = rk[r];
    }
    constexpr int DEG = mono_deg(NP - 1);
    // Phi_ab = |xi_a - xi_b|^m once per pair: node i takes the pairs (i, i+d mod n)
    // for d = 1..n/2, which covers each pair once (twice for d = n/2, n even)
    {
        double *phi = L.u[grp] + T::PHI0;
        // MC > 0: the power as a constant (the usual odd powers), so the
        // square-and-multiply folds; the branch on m is uniform
        auto pairs = [&](auto mc) {
            constexpr int MC = decltype(mc)::value;
            const int mm = MC > 0 ? MC : m;
            const double diag = phs_r2(mm, 0.0);
#pragma unroll
            for (int r = 0; r < NODES; ++r) {
                const int i = t + G * r;
                if (i < NN) {
                    phi[T::tri(i) + i] = diag;
#pragma unroll 3
                    for (int d = 1; d <= NN / 2; ++d) {
                        int j = i + d;
                        if (j >= NN) j -= NN;
                        const double ddx = dxn[r] - L.u[grp][T::XI0 + 2 * j], ddy = dyn[r] - L.u[grp][T::XI0 + 2 * j + 1];
                        const int lo = i < j ? i : j, hi = i < j ? j : i;
                        phi[T::tri(lo) + hi] = phs_r2(mm, sq_sum(ddx, ddy));
                    }
                }
            }
        };
        switch (m) {
        case 3: pairs(std::integral_constant<int, 3>{}); break;
        case 5: pairs(std::integral_constant<int, 5>{}); break;
        case 7: pairs(std::integral_constant<int, 7>{}); break;
        case 9: pairs(std::integral_constant<int, 9>{}); break;
        default: pairs(std::integral_constant<int, 0>{}); break;
        }
    }
    __builtin_amdgcn_wave_barrier();

    // ---- equations by row slot.  Lane t's slot rr starts with RBF equation
    // i = t + G rr (src/rbfblock.jl:14-20 with the P block); eq_of gives i, or
    // NS for an empty slot.  The monomial equations (the P^T rows) enter the
    // slots later, when the lambda columns are done.
    auto eq_of = [&](int rr) -> int {
        const int i = t + G * rr;
        return i < NN ? i : NS;
    };

    // ---- right-hand sides first, while few registers are live: RBF rows at
    // u = Yscaled - xi with exact zeros -> eps (src/generate_operator.jl:125-154),
    // polynomial rows at Yscaled (:68-79); lane t computes monomial equation
    // t + G r for its slots' later tenants
    double rhs[ROWS][R];
#pragma unroll
    for (int rr = 0; rr < ROWS; ++rr) {
        const int e = eq_of(rr);
#pragma unroll
        for (int cc = 0; cc < R; ++cc) rhs[rr][cc] = 0.0;
        if (e < NN) {
            const int i = e;
            const double xix = L.u[grp][T::XI0 + 2 * i], xiy = L.u[grp][T::XI0 + 2 * i + 1];
            // u = Y - xi for the Y point at (px, py)
            auto u_at = [&](double px, double py, double &ux, double &uy) {
#pragma clang fp contract(off)
                if (a.method == 3) {
                    ux = xix;
                    uy = xiy;
                } else {
                    ux = px - xix;
                    uy = py - xiy;
                    if (ux == 0.0) ux = kEps;
                    if (uy == 0.0) uy = kEps;
                }
            };
            if constexpr (GRP) {
                double ux = 0.0, uy = 0.0;
                PhsAt q{};
#pragma unroll
                for (int cc = 0; cc < R; ++cc) {
                    if (cc < ncols) {
                        if (a.col_op[cc] == 0) {  // the first operator of the next output row
                            const int sub = a.col_sub[cc];
                            u_at(rows_all[w].ys[grp][sub][0], rows_all[w].ys[grp][sub][1], ux, uy);
                            q = phs_at(m, ux, uy);
                        }
                        rhs[rr][cc] = op_phs_fast(a.col_kind[cc], a.col_k[cc], m, ux, uy, sx, sy, q);
                    }
                }
            } else {
                double ux, uy;
                u_at(yx, yy, ux, uy);
                const PhsAt q = phs_at(m, ux, uy);
#pragma unroll
                for (int cc = 0; cc < R; ++cc)
                    if (cc < a.nops) rhs[rr][cc] = op_phs_fast(a.op_kind[cc], a.op_k[cc], m, ux, uy, sx, sy, q);
            }
        }
    }
    double mrhs[T::QR][R];
#pragma unroll
    for (int r = 0; r < T::QR; ++r) {
        const int q = t + G * r;
#pragma unroll
        for (int cc = 0; cc < R; ++cc) mrhs[r][cc] = 0.0;
        if (q < NP) {
            int tt = 0;
            while ((tt + 1) * (tt + 2) / 2 <= q) ++tt;
            const int ea = q - tt * (tt + 1) / 2, eb = tt - ea;
#pragma unroll
            for (int cc = 0; cc < R; ++cc) {
                if constexpr (GRP) {
                    if (cc < ncols) {
                        const int sub = a.col_sub[cc];
                        mrhs[r][cc] = op_mono(a.col_kind[cc], a.col_k[cc], ea, eb, rows_all[w].ys[grp][sub][0],
                                              rows_all[w].ys[grp][sub][1], sx, sy);
                    }
                } else {
                    if (cc < a.nops) mrhs[r][cc] = op_mono(a.op_kind[cc], a.op_k[cc], ea, eb, yx, yy, sx, sy);
                }
            }
        }
    }

    // ---- rows of [A | B] (src/interpolationmatrix.jl:2-5, src/generate_operator.jl:157).
    // Column order: lambda (highest degree first), then w, then the
    // right-hand sides at NS.
    double row[ROWS][WIDTH];
    const double *phi = L.u[grp] + T::PHI0;
    const double *pblk = L.u[grp] + T::PB0;  // P(node j, monomial q) at j * NP + q
#pragma unroll
    for (int rr = 0; rr < ROWS; ++rr) {
        const int e = eq_of(rr);
        if (e < NN) {
            const int rbase = T::tri(e);  // Phi(e, j) for j >= e
#pragma unroll
            for (int j = 0; j < NN; ++j) row[rr][NP + j] = phi[j >= e ? rbase + j : T::tri(j) + e];
            // node e = t + G rr is this lane's node rr: its monomials, highest first
            double pxp[DEG + 1], pyp[DEG + 1];
            pxp[0] = 1.0;
            pyp[0] = 1.0;
#pragma unroll
            for (int d = 1; d <= DEG; ++d) { pxp[d] = pxp[d - 1] * dxn[rr]; pyp[d] = pyp[d - 1] * dyn[rr]; }
#pragma unroll
            for (int q = 0; q < NP; ++q) row[rr][q] = pxp[mono_ea(NP - 1 - q)] * pyp[mono_eb(NP - 1 - q)];
        } else {
#pragma unroll
            for (int j = 0; j < NS; ++j) row[rr][j] = 0.0;
        }
#pragma unroll
        for (int cc = 0; cc < R; ++cc) row[rr][NS + cc] = rhs[rr][cc];
    }
    // P and the monomial equations' right-hand sides, over the triangle just read
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int rr = 0; rr < ROWS; ++rr) {
        const int e = eq_of(rr);
        if (e < NN) {
#pragma unroll
            for (int q = 0; q < NP; ++q) L.u[grp][T::PB0 + e * NP + (NP - 1 - q)] = row[rr][q];
        }
    }
#pragma unroll
    for (int r = 0; r < T::QR; ++r) {
        const int q = t + G * r;
        if (q < NP) {
#pragma unroll
            for (int cc = 0; cc < R; ++cc) L.u[grp][T::MR0 + q * R + cc] = mrhs[r][cc];
        }
    }
    __builtin_amdgcn_wave_barrier();

    // ---- Gauss-Jordan elimination with partial pivoting; the right-hand-side
    // columns ride along.  Rows slide one place per step (row[j] <- row[j+1] - l u[j+1]), so the
    // pivot column is always register 0 and every register index is static.
    // Steps run as a real loop in blocks of SB steps with a static width that
    // shrinks per block: compact code, at most SB-1 stale columns of
    // extra work per step.
    // alive[rr]: all ones while row slot rr holds an equation not yet used as a
    // pivot, zero afterwards and for empty slots
    uint32_t alive[ROWS];
    // the column each row was the pivot for, and that pivot
    int kp[ROWS];
    double pval[ROWS];
#pragma unroll
    for (int rr = 0; rr < ROWS; ++rr) {
        alive[rr] = eq_of(rr) < NS ? 0xFFFFFFFFu : 0u;
        kp[rr] = NS;
        pval[rr] = 1.0;
    }
    double *U = L.u[grp];
    // This lane's pivot candidate among its first RC slots: the largest
    // |row[.][0]| among live rows, as a key whose low 7 bits name the row slot
    // (bit 7 set, so a live zero still beats a used row), and the reciprocal
    // of that entry, so the winning lane publishes 1/pivot with the pivot row.
    uint32_t key;
    double crp;
    auto candidate = [&](auto rcc) {
        constexpr int RC = decltype(rcc)::value;
        uint32_t kk = 0;
        double cv = row[0][0];
#pragma unroll
        for (int rr = 0; rr < RC; ++rr) {
            // |f| and the low byte cleared in one mask
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
    using RowsAll = std::integral_constant<int, ROWS>;
    candidate(RowsAll{});
    uint32_t kmax = group_max_u32<G>(key);
    int k = 0;
    // One step at static width WB >= WIDTH - k: slots below RF are updated
    // (the winner is among them), slots below RC enter the next pivot search.
    // 64 lanes (one stencil per wavefront): the pivot lane is the same for the
    // whole wavefront, so every lane reads the pivot row straight from that
    // lane's registers (v_readlane into scalar registers, which the FMAs take
    // as operands) and the step makes no LDS round trip.
    auto step64 = [&](auto wbc, auto rfc, auto rcc) {
        constexpr int WB = decltype(wbc)::value;
        constexpr int RF = decltype(rfc)::value;
        const int v = 127 - (int)((uint32_t)__builtin_amdgcn_readfirstlane((int)kmax) & 127u);
        const int wl = v & 63, psel = v >> 6;
        // lane wl's x, in every lane
        auto rl = [&](double x) {
            const long long b = __double_as_longlong(x);
            const int lo = __builtin_amdgcn_readlane((int)b, wl);
            const int hi = __builtin_amdgcn_readlane((int)(b >> 32), wl);
            return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
        };
        for_blocks<0, RF>([&](auto rc) {
            constexpr int rp_slot = decltype(rc)::value;
            if (psel != rp_slot) return;  // uniform
            const double rp = rl(crp);
            const double p1 = rl(row[rp_slot][1]);
            if (t == wl) {
                alive[rp_slot] = 0u;
                kp[rp_slot] = k;
                pval[rp_slot] = row[rp_slot][0];
                // a zero multiplier keeps the pivot row as it is
                row[rp_slot][0] = 0.0;
            }
            double nl[ROWS];
#pragma unroll
            for (int rr = 0; rr < RF; ++rr) nl[rr] = -(row[rr][0] * rp);
#pragma unroll
            for (int rr = 0; rr < RF; ++rr) row[rr][0] = __builtin_fma(nl[rr], p1, row[rr][1]);
            candidate(rcc);
            kmax = group_max_u32<G>(key);
            // columns j = 1 .. WB-2 take column j+1 of the pivot row, in fours
            constexpr int NQ = (WB - 2) / 4;
            for_blocks<0, NQ>([&](auto qc) {
                constexpr int j = 1 + 4 * decltype(qc)::value;
                const double s0 = rl(row[rp_slot][j + 1]), s1 = rl(row[rp_slot][j + 2]);
                const double s2 = rl(row[rp_slot][j + 3]), s3 = rl(row[rp_slot][j + 4]);
#pragma unroll
                for (int rr = 0; rr < RF; ++rr)
                    slide4_s(row[rr][j], row[rr][j + 1], row[rr][j + 2], row[rr][j + 3], row[rr][j + 4], nl[rr], s0, s1,
                             s2, s3);
            });
            for_blocks<1 + 4 * NQ, WB - 1>([&](auto jc) {
                constexpr int j = decltype(jc)::value;
                const double sj = rl(row[rp_slot][j + 1]);
#pragma unroll
                for (int rr = 0; rr < RF; ++rr) slide1_s(row[rr][j], row[rr][j + 1], nl[rr], sj);
            });
        });
    };
    auto step = [&](auto wbc, auto rfc, auto rcc) {
        constexpr int WB = decltype(wbc)::value;
        constexpr int RF = decltype(rfc)::value;
        if constexpr (G == 64) {
            step64(wbc, rfc, rcc);
            return;
        }
        const int v = 127 - (int)(kmax & 127u);
        const bool iam = t == v % G;
        const int psel = v / G;
        double *piv = U;  // piv[j] = pivot row, column k + j
        // The row goes out in two parts: columns 0..15 and the reciprocal
        // pivot, read back at once; the rest after the update of the
        // columns the first part feeds.  LDS serves a wavefront's
        // operations in order, so those FMAs wait for the first part only.
        constexpr int W0 = WB < 16 ? WB : 16;
        constexpr int NPV = (WB + 15) / 16;
        // 16-lane groups: the winner's reciprocal pivot and column k+1 go to
        // the group by an OR over the DPP row (the other lanes contribute zero
        // bits), so the next pivot search does not wait on LDS
        uint64_t orp = 0, op1 = 0;
        if constexpr (G == 16) {
            if (iam) orp = (uint64_t)__double_as_longlong(crp);
#pragma unroll
            for (int rr = 0; rr < RF; ++rr)
                if (iam && psel == rr) op1 = (uint64_t)__double_as_longlong(row[rr][1]);
        } else {
            if (iam) L.rp[grp] = crp;
        }
#pragma unroll
        for (int rr = 0; rr < RF; ++rr) {
            if (iam && psel == rr) {
#pragma unroll
                for (int j = G == 16 ? 2 : 0; j < W0; ++j) piv[j] = row[rr][j];  // 16 lanes: 0, 1 go by DPP
                alive[rr] = 0u;
                kp[rr] = k;
                pval[rr] = row[rr][0];
                // a zero multiplier keeps the pivot row as it is
                row[rr][0] = 0.0;
            }
        }
        // the other lanes read what the pivot lane just wrote
        __builtin_amdgcn_wave_barrier();
        // the pivot row spread over each 16-lane DPP row: lane q holds
        // columns q, q+16, ...; a DPP broadcast hands column j to the FMAs
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
        // Gauss-Jordan: rows that were pivots earlier are eliminated too, so
        // no back substitution is needed (rows past the system are zero)
#pragma unroll
        for (int rr = 0; rr < RF; ++rr) nl[rr] = -(row[rr][0] * rp);
        // column k+1 first: the next step's pivot search overlaps the rest
#pragma unroll
        for (int rr = 0; rr < RF; ++rr) row[rr][0] = fma3(nl[rr], p1, row[rr][1]);
        candidate(rcc);
        kmax = group_max_u32<G>(key);
        // columns j = J0 .. J1-1 get column j+1 of the pivot row: in fours
        // (one wait for four broadcasts), then a pair and a single
        auto columns = [&](auto j0c, auto j1c) {
            constexpr int J0 = decltype(j0c)::value, J1 = decltype(j1c)::value;
            constexpr int NQ = (J1 - J0) / 4;
            for_blocks<0, NQ>([&](auto qc) {
                constexpr int j = J0 + 4 * decltype(qc)::value;
                double p0, p1, p2, p3;
                row_bcast64x4<(j + 1) % 16, (j + 2) % 16, (j + 3) % 16, (j + 4) % 16>(
                    pv[(j + 1) / 16], pv[(j + 2) / 16], pv[(j + 3) / 16], pv[(j + 4) / 16], p0, p1, p2, p3);
#pragma unroll
                for (int rr = 0; rr < RF; ++rr) row[rr][j] = fma3(nl[rr], p0, row[rr][j + 1]);
#pragma unroll
                for (int rr = 0; rr < RF; ++rr) row[rr][j + 1] = fma3(nl[rr], p1, row[rr][j + 2]);
#pragma unroll
                for (int rr = 0; rr < RF; ++rr) row[rr][j + 2] = fma3(nl[rr], p2, row[rr][j + 3]);
#pragma unroll
                for (int rr = 0; rr < RF; ++rr) row[rr][j + 3] = fma3(nl[rr], p3, row[rr][j + 4]);
            });
            constexpr int J2 = J0 + 4 * NQ;
            if constexpr (J1 - J2 >= 2) {
                constexpr int j = J2;
                double p0, p1;
                row_bcast64x2<(j + 1) % 16, (j + 2) % 16>(pv[(j + 1) / 16], pv[(j + 2) / 16], p0, p1);
#pragma unroll
                for (int rr = 0; rr < RF; ++rr) row[rr][j] = fma3(nl[rr], p0, row[rr][j + 1]);
#pragma unroll
                for (int rr = 0; rr < RF; ++rr) row[rr][j + 1] = fma3(nl[rr], p1, row[rr][j + 2]);
            }
            if constexpr ((J1 - J2) % 2 == 1) {
                constexpr int j = J1 - 1;
                const double pj = row_bcast64<(j + 1) % 16>(pv[(j + 1) / 16]);
#pragma unroll
                for (int rr = 0; rr < RF; ++rr) row[rr][j] = fma3(nl[rr], pj, row[rr][j + 1]);
            }
        };
        // columns up to 15 come from the first part; the rest of the row
        // goes out only now, so the wait for the first part does not
        // cover it.  The pivot lane's registers from 16 on are not yet
        // updated (the update reaches register j from j+1 = 16 next).
        constexpr int JE = WB - 1;  // columns j = 1 .. WB-2
        constexpr int JA = JE < 15 ? JE : 15;
        columns(std::integral_constant<int, 1>{}, std::integral_constant<int, JA>{});
        if constexpr (NPV > 1) {
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int rr = 0; rr < RF; ++rr) {
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
    // steps K0..K1-1 in blocks of SB steps, each block a real loop at the
    // static width of its first step
    auto run_steps = [&](auto k0c, auto k1c, auto rfc, auto rcc) {
        constexpr int K0 = decltype(k0c)::value, K1 = decltype(k1c)::value;
        for_blocks<0, (K1 - K0 + T::SB - 1) / T::SB>([&](auto bc) {
            constexpr int kb = K0 + T::SB * decltype(bc)::value;
            const int kend = K1 < kb + T::SB ? K1 : kb + T::SB;
#pragma nounroll
            for (; k < kend; ++k) {
                step(std::integral_constant<int, WIDTH - kb>{}, rfc, rcc);
                // the next step's pivot lane overwrites the stale tail read here
                __builtin_amdgcn_wave_barrier();
            }
        });
    };
    // the lambda columns, on the RBF equations
    run_steps(std::integral_constant<int, 0>{}, std::integral_constant<int, NP>{}, RowsAll{}, RowsAll{});
    // The monomial equations, as they stand after NP slides (column NP + j in
    // register j: they are zero in the lambda columns, so those steps would
    // not have changed them), into the free slots: the NP lambda pivot rows
    // and the empty slots, at least NP in all.  The group's free slots are
    // ranked slot by slot, lane by lane, and the slot of rank q takes monomial
    // equation q.
    {
        int base = 0;
#pragma unroll
        for (int rr = 0; rr < ROWS; ++rr) {
            const bool fr = alive[rr] == 0u;
            const uint64_t bal = __builtin_amdgcn_ballot_w64(fr);
            const uint64_t gm = G == 64 ? bal : (bal >> (grp * G)) & ((1ull << (G & 63)) - 1);
            const int q = base + __builtin_popcountll(gm & ((1ull << t) - 1));
            base += __builtin_popcountll(gm);
            if (fr && q < NP) {
#pragma unroll
                for (int j = 0; j < NN; ++j) row[rr][j] = pblk[j * NP + q];
#pragma unroll
                for (int cc = 0; cc < R; ++cc) row[rr][NN + cc] = U[T::MR0 + q * R + cc];
                alive[rr] = 0xFFFFFFFFu;
                kp[rr] = NS;
                pval[rr] = 1.0;
            }
        }
    }
    candidate(RowsAll{});
    kmax = group_max_u32<G>(key);
    // the w columns, on the remaining RBF equations and the monomial equations
    run_steps(std::integral_constant<int, NP>{}, std::integral_constant<int, NS>{}, RowsAll{}, RowsAll{});
    __builtin_amdgcn_wave_barrier();

    // ---- solution: after NS slides a row's registers 0..R-1 hold its
    // right-hand sides, reduced to pivot * x for the row's pivot column
    double xv[ROWS][R];
#pragma unroll
    for (int rr = 0; rr < ROWS; ++rr) {
        const double rd = 1.0 / pval[rr];
#pragma unroll
        for (int cc = 0; cc < R; ++cc) xv[rr][cc] = row[rr][cc] * rd;
    }

    // ---- CSR output: the row that solved for node u writes it, at the rank
    // of u's global index inside the stencil (sorted columns)
    if constexpr (GRP) {
        // column cc to output row col_sub[cc], operator col_op[cc]; one
        // non-finite bit per output row
        uint64_t badm = 0;
#pragma unroll
        for (int rr = 0; rr < ROWS; ++rr) {
            const int u = kp[rr] - NP;
            if (u >= 0 && u < NN && live) {
                const int gu = L.g[grp][u];
                const int rku = L.rank[grp][u];
#pragma unroll
                for (int cc = 0; cc < R; ++cc) {
                    if (cc < ncols) {
                        const int sub = a.col_sub[cc], o = a.col_op[cc];
                        const long base = (long)rows_all[w].yrow[grp][sub] * NN + rku;
                        const double f = a.method == 3 ? 1.0 : op_post_scale(a.col_kind[cc], a.col_k[cc], sx, sy);
                        const double wv = f * xv[rr][cc];
                        if (!__builtin_isfinite(wv)) badm |= 1ull << sub;
                        a.vals[o][base] = wv;
                        if (o == 0 && a.colidx) a.colidx[base] = gu;
                    }
                }
            }
        }
        bool singular = false;
#pragma unroll
        for (int rr = 0; rr < ROWS; ++rr) singular |= pval[rr] == 0.0;
        const uint32_t sflag = group_max_u32<G>(singular ? 1u : 0u);
        badm = group_or_u64<G>(badm);
        if (a.status && live) {
            for (int q = t; q < ny; q += G) {
                const int yq = rows_all[w].yrow[grp][q];
                const int st = sflag ? 1 : (((badm >> q) & 1u) ? 2 : 0);
                if (a.status_accumulate) a.status[yq] = max(a.status[yq], st);
                else a.status[yq] = st;
            }
        }
        return;
    }
    uint32_t bad = 0;
#pragma unroll
    for (int rr = 0; rr < ROWS; ++rr) {
        const int u = kp[rr] - NP;  // the w column it solved for
        if (u >= 0 && u < NN && live) {
            const int gu = L.g[grp][u];
            const long base = yk * NN + L.rank[grp][u];
            if (a.colidx) a.colidx[base] = gu;
#pragma unroll
            for (int cc = 0; cc < R; ++cc) {
                if (cc < a.nops) {
                    const double f = a.method == 3 ? 1.0 : op_post_scale(a.op_kind[cc], a.op_k[cc], sx, sy);
                    const double wv = f * xv[rr][cc];
                    if (!__builtin_isfinite(wv)) bad = 1;
                    a.vals[cc][base] = wv;
                }
            }
        }
    }
    // a zero pivot: the system is exactly singular (rows past the system keep 1)
    bool singular = false;
#pragma unroll
    for (int rr = 0; rr < ROWS; ++rr) singular |= pval[rr] == 0.0;
    const uint32_t sflag = group_max_u32<G>((singular ? 2u : 0u) | bad);
    if (a.status && t == 0 && live) {
        const int st = (sflag & 2u) ? 1 : ((sflag & 1u) ? 2 : 0);
        if (a.status_accumulate) a.status[yk] = max(a.status[yk], st);
        else a.status[yk] = st;
    }
}

}  // namespace rbffd
