// device_common.hpp -- wavefront helpers and closed-form basis functions shared
// by the gfx950 kernels of librbffd.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rbffd {

// Julia eps() for Float64 = 2^-52; the reference substitutes it for exact zero
// components of the stencil coordinates (src/generate_operator.jl:127-132).
constexpr double kEps = 2.220446049250313e-16;

enum OpKind : int {
    OP_E = 0, OP_DX = 1, OP_DY = 2, OP_DXX = 3, OP_DYY = 4, OP_DXY = 5,
    OP_XK = 6, OP_YK = 7, OP_LAP = 8, OP_LAPK = 9
};

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// ---- wave64 reductions --------------------------------------------------
// DPP ladder: quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_half_mirror,
// row_mirror give every lane its 16-lane row maximum; row_bcast:15 and
// row_bcast:31 fold the four rows into lane 63.
template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROW_MASK, 0xF, false);
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v)
{
    v = max(v, dpp_u32<0xB1>(v));
    v = max(v, dpp_u32<0x4E>(v));
    v = max(v, dpp_u32<0x141>(v));
    v = max(v, dpp_u32<0x140>(v));
    v = max(v, dpp_u32<0x142, 0xA>(v));
    v = max(v, dpp_u32<0x143, 0xC>(v));
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

__device__ __forceinline__ uint32_t wave_or_u32(uint32_t v)
{
    v |= dpp_u32<0xB1>(v);
    v |= dpp_u32<0x4E>(v);
    v |= dpp_u32<0x141>(v);
    v |= dpp_u32<0x140>(v);
    v |= dpp_u32<0x142, 0xA>(v);
    v |= dpp_u32<0x143, 0xC>(v);
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

// Maximum of non-negative doubles across the wave (bit patterns of
// non-negative IEEE doubles order like unsigned integers).
__device__ __forceinline__ double wave_max_nonneg_f64(double v)
{
    uint64_t b = (uint64_t)__double_as_longlong(v);
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        uint64_t ob = (uint64_t)__shfl_xor((unsigned long long)b, o, 64);
        b = ob > b ? ob : b;
    }
    return __longlong_as_double((long long)b);
}

__device__ __forceinline__ double bcast_f64(double v, int src)
{
    int2 p = *reinterpret_cast<int2 *>(&v);
    p.x = __builtin_amdgcn_readlane(p.x, src);
    p.y = __builtin_amdgcn_readlane(p.y, src);
    return *reinterpret_cast<double *>(&p);
}

// ---- sums over lane groups (SpMV rows) ------------------------------------
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v)
{
    int2 p = *reinterpret_cast<int2 *>(&v);
    p.x = __builtin_amdgcn_update_dpp(0, p.x, CTRL, 0xF, 0xF, false);
    p.y = __builtin_amdgcn_update_dpp(0, p.y, CTRL, 0xF, 0xF, false);
    return *reinterpret_cast<double *>(&p);
}

// Sum over each aligned group of G lanes, in every lane of the group: DPP
// inside a 16-lane row (quad_perm [1,0,3,2], quad_perm [2,3,0,1],
// row_half_mirror, row_mirror), a shuffle across rows.  Every lane adds the
// same two partial sums, so all lanes hold the same bits.
template <int G>
__device__ __forceinline__ double group_sum_f64(double v)
{
    static_assert(G == 1 || G == 2 || G == 4 || G == 8 || G == 16 || G == 32, "group size");
    if constexpr (G >= 2) v += dpp_f64<0xB1>(v);
    if constexpr (G >= 4) v += dpp_f64<0x4E>(v);
    if constexpr (G >= 8) v += dpp_f64<0x141>(v);
    if constexpr (G >= 16) v += dpp_f64<0x140>(v);
    if constexpr (G >= 32) v += __shfl_xor(v, 16, 32);
    return v;
}

// ---- closed-form polyharmonic spline derivatives ---------------------------
// phi(v) = |v|^m (src/rbfbasis.jl:9).  d^a/dx^a d^b/dy^b of g^s, g = vx^2+vy^2,
// s = m/2:
//   sum_{i<=a/2, j<=b/2} a!/(i!(a-2i)!) b!/(j!(b-2j)!) s(s-1)..(s-q+1)
//        (2vx)^(a-2i) (2vy)^(b-2j) r^(m-2q),     q = a+b-i-j.
__device__ __forceinline__ double ipow_d(double x, int e)
{
    double r = 1.0;
    int n = e < 0 ? -e : e;
    for (int t = 0; t < n; ++t) r *= x;
    return e < 0 ? 1.0 / r : r;
}

__device__ __forceinline__ double fact_d(int n)
{
    double r = 1.0;
    for (int i = 2; i <= n; ++i) r *= (double)i;
    return r;
}

// r^m with r = sqrt(r2): odd m uses one square root, even m none.
__device__ __forceinline__ double phs_value(int m, double r2)
{
    if (m & 1) return ipow_d(r2, m >> 1) * __builtin_sqrt(r2);
    return ipow_d(r2, m >> 1);
}

__device__ inline double phs_deriv(int m, int a, int b, double vx, double vy)
{
    double r2 = vx * vx + vy * vy;
    if (a == 0 && b == 0) return phs_value(m, r2);
    double r = __builtin_sqrt(r2);
    double s = 0.5 * (double)m;
    double sum = 0.0;
    for (int i = 0; 2 * i <= a; ++i) {
        for (int j = 0; 2 * j <= b; ++j) {
            int q = a + b - i - j;
            double c = (fact_d(a) / (fact_d(i) * fact_d(a - 2 * i)))
                     * (fact_d(b) / (fact_d(j) * fact_d(b - 2 * j)));
            double ff = 1.0;
            for (int t = 0; t < q; ++t) ff *= (s - (double)t);
            sum += c * ff * ipow_d(2.0 * vx, a - 2 * i) * ipow_d(2.0 * vy, b - 2 * j)
                 * ipow_d(r, m - 2 * q);
        }
    }
    return sum;
}

__device__ inline double mono_deriv(int e, int f, int a, int b, double x, double y)
{
    if (a > e || b > f) return 0.0;
    double c = 1.0;
    for (int t = 0; t < a; ++t) c *= (double)(e - t);
    for (int t = 0; t < b; ++t) c *= (double)(f - t);
    return c * ipow_d(x, e - a) * ipow_d(y, f - b);
}

__device__ __forceinline__ double binom_d(int n, int k)
{
    return fact_d(n) / (fact_d(k) * fact_d(n - k));
}

// Operator applied to the PHS at v (Laplacian-type operators carry the
// per-axis scales inside; single-derivative operators are rescaled after the
// solve, src/generate_operator.jl:161-166).
__device__ __attribute__((noinline)) double op_phs(int op, int k, int m, double vx, double vy, double sx, double sy)
{
    switch (op) {
    case OP_E:   return phs_deriv(m, 0, 0, vx, vy);
    case OP_DX:  return phs_deriv(m, 1, 0, vx, vy);
    case OP_DY:  return phs_deriv(m, 0, 1, vx, vy);
    case OP_DXX: return phs_deriv(m, 2, 0, vx, vy);
    case OP_DYY: return phs_deriv(m, 0, 2, vx, vy);
    case OP_DXY: return phs_deriv(m, 1, 1, vx, vy);
    case OP_XK:  return phs_deriv(m, k, 0, vx, vy);
    case OP_YK:  return phs_deriv(m, 0, k, vx, vy);
    case OP_LAP:
        return sx * sx * phs_deriv(m, 2, 0, vx, vy) + sy * sy * phs_deriv(m, 0, 2, vx, vy);
    case OP_LAPK: {
        double sum = 0.0;
        for (int a = 0; a <= k; ++a)
            sum += binom_d(k, a) * ipow_d(sx * sx, a) * ipow_d(sy * sy, k - a)
                 * phs_deriv(m, 2 * a, 2 * (k - a), vx, vy);
        return sum;
    }
    default: return 0.0;
    }
}

__device__ __attribute__((noinline)) double op_mono(int op, int k, int e, int f, double x, double y, double sx, double sy)
{
    switch (op) {
    case OP_E:   return mono_deriv(e, f, 0, 0, x, y);
    case OP_DX:  return mono_deriv(e, f, 1, 0, x, y);
    case OP_DY:  return mono_deriv(e, f, 0, 1, x, y);
    case OP_DXX: return mono_deriv(e, f, 2, 0, x, y);
    case OP_DYY: return mono_deriv(e, f, 0, 2, x, y);
    case OP_DXY: return mono_deriv(e, f, 1, 1, x, y);
    case OP_XK:  return mono_deriv(e, f, k, 0, x, y);
    case OP_YK:  return mono_deriv(e, f, 0, k, x, y);
    case OP_LAP:
        return sx * sx * mono_deriv(e, f, 2, 0, x, y) + sy * sy * mono_deriv(e, f, 0, 2, x, y);
    case OP_LAPK: {
        double sum = 0.0;
        for (int a = 0; a <= k; ++a)
            sum += binom_d(k, a) * ipow_d(sx * sx, a) * ipow_d(sy * sy, k - a)
                 * mono_deriv(e, f, 2 * a, 2 * (k - a), x, y);
        return sum;
    }
    default: return 0.0;
    }
}

// Multiplier applied to the solved weights (src/generate_operator.jl:161-166,
// src/hyperviscosity_operator.jl:159-160).
__device__ __forceinline__ double op_post_scale(int op, int k, double sx, double sy)
{
    switch (op) {
    case OP_DX:  return sx;
    case OP_DY:  return sy;
    case OP_DXX: return sx * sx;
    case OP_DYY: return sy * sy;
    case OP_DXY: return sx * sy;
    case OP_XK:  return ipow_d(sx, k);
    case OP_YK:  return ipow_d(sy, k);
    default:     return 1.0;
    }
}

// Monomial exponents of basis entry q: graded by total degree, increasing
// power of x inside a degree (1, y, x, y^2, xy, x^2, ...).
__host__ __device__ constexpr int mono_deg(int q)
{
    int t = 0;
    while ((t + 1) * (t + 2) / 2 <= q) ++t;
    return t;
}
__host__ __device__ constexpr int mono_ea(int q) { return q - mono_deg(q) * (mono_deg(q) + 1) / 2; }
__host__ __device__ constexpr int mono_eb(int q) { return mono_deg(q) - mono_ea(q); }

}  // namespace rbffd
