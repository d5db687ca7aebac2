/*
 * oracle_core.h -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * CPU restatement of the per-stencil RBF-FD algorithm of
 * RadialBasisFiniteDifferences.jl, generic over the real type REAL.
 * rbffd_oracle.c includes this file twice: once with REAL=double (the FP64
 * oracle that mirrors the reference's Float64 arithmetic) and once with
 * REAL=__float128 (the high-precision oracle).  Nothing in the product path
 * links or calls this code; only tests, the smoke check and the CPU-baseline
 * leg of the benchmark do.
 *
 * Macros expected from the includer:
 *   REAL      arithmetic type
 *   RS(name)  name mangling with a type suffix
 *   RSQRT     square root for REAL
 *   RFABS     absolute value for REAL
 */

/* ---------------------------------------------------------------------------
 * Polyharmonic spline derivatives.
 *
 * phi(v) = (sqrt(vx^2 + vy^2))^m            (src/rbfbasis.jl:9, rbfbasis_k.jl:9)
 * The reference differentiates phi symbolically (src/rbfbasis.jl:19-30,
 * src/rbfbasis_k.jl:14-18).  Here every mixed derivative d^a/dx^a d^b/dy^b is
 * evaluated from the closed form for derivatives of g(v)^s, g = vx^2 + vy^2,
 * s = m/2:
 *
 *   sum_{i<=a/2} sum_{j<=b/2} a!/(i!(a-2i)!) * b!/(j!(b-2j)!)
 *        * s(s-1)...(s-q+1) * (2vx)^(a-2i) * (2vy)^(b-2j) * r^(m-2q),
 *   q = a + b - i - j.
 * -------------------------------------------------------------------------*/
static REAL RS(ipow)(REAL x, int e)
{
    REAL r = 1;
    int neg = e < 0;
    if (neg) e = -e;
    for (int t = 0; t < e; ++t) r *= x;
    return neg ? (REAL)1 / r : r;
}

static REAL RS(phs_deriv)(int m, int a, int b, REAL vx, REAL vy)
{
    REAL r2 = vx * vx + vy * vy;
    REAL r = RSQRT(r2);
    if (a == 0 && b == 0) return RS(ipow)(r, m);
    REAL s = (REAL)m / 2;
    REAL sum = 0;
    for (int i = 0; 2 * i <= a; ++i) {
        for (int j = 0; 2 * j <= b; ++j) {
            int q = a + b - i - j;
            REAL c = (REAL)(fact(a) / (fact(i) * fact(a - 2 * i)))
                   * (REAL)(fact(b) / (fact(j) * fact(b - 2 * j)));
            REAL ff = 1;
            for (int t = 0; t < q; ++t) ff *= (s - t);
            sum += c * ff * RS(ipow)(2 * vx, a - 2 * i) * RS(ipow)(2 * vy, b - 2 * j)
                 * RS(ipow)(r, m - 2 * q);
        }
    }
    return sum;
}

/* d^a/dx^a d^b/dy^b of x^e y^f at (x, y)  (src/polynomialbasis.jl:14-30,
 * src/polynomialbasis_k.jl:14-19 differentiate the monomials symbolically). */
static REAL RS(mono_deriv)(int e, int f, int a, int b, REAL x, REAL y)
{
    if (a > e || b > f) return 0;
    REAL c = 1;
    for (int t = 0; t < a; ++t) c *= (REAL)(e - t);
    for (int t = 0; t < b; ++t) c *= (REAL)(f - t);
    return c * RS(ipow)(x, e - a) * RS(ipow)(y, f - b);
}

/* Apply operator `op` (order k) to the PHS at v, with per-axis scale (sx, sy)
 * folded in only for the Laplacian-type operators (the reference rescales the
 * single-derivative operators after the solve, src/generate_operator.jl:161-166). */
static REAL RS(op_phs)(int op, int k, int m, REAL vx, REAL vy, REAL sx, REAL sy)
{
    switch (op) {
    case ORC_OP_E:   return RS(phs_deriv)(m, 0, 0, vx, vy);
    case ORC_OP_DX:  return RS(phs_deriv)(m, 1, 0, vx, vy);
    case ORC_OP_DY:  return RS(phs_deriv)(m, 0, 1, vx, vy);
    case ORC_OP_DXX: return RS(phs_deriv)(m, 2, 0, vx, vy);
    case ORC_OP_DYY: return RS(phs_deriv)(m, 0, 2, vx, vy);
    case ORC_OP_DXY: return RS(phs_deriv)(m, 1, 1, vx, vy);
    case ORC_OP_XK:  return RS(phs_deriv)(m, k, 0, vx, vy);
    case ORC_OP_YK:  return RS(phs_deriv)(m, 0, k, vx, vy);
    case ORC_OP_LAP:
        return sx * sx * RS(phs_deriv)(m, 2, 0, vx, vy) + sy * sy * RS(phs_deriv)(m, 0, 2, vx, vy);
    case ORC_OP_LAPK: {
        REAL sum = 0;
        for (int a = 0; a <= k; ++a) {
            REAL c = (REAL)binom(k, a) * RS(ipow)(sx * sx, a) * RS(ipow)(sy * sy, k - a);
            sum += c * RS(phs_deriv)(m, 2 * a, 2 * (k - a), vx, vy);
        }
        return sum;
    }
    }
    return 0;
}

static REAL RS(op_mono)(int op, int k, int e, int f, REAL x, REAL y, REAL sx, REAL sy)
{
    switch (op) {
    case ORC_OP_E:   return RS(mono_deriv)(e, f, 0, 0, x, y);
    case ORC_OP_DX:  return RS(mono_deriv)(e, f, 1, 0, x, y);
    case ORC_OP_DY:  return RS(mono_deriv)(e, f, 0, 1, x, y);
    case ORC_OP_DXX: return RS(mono_deriv)(e, f, 2, 0, x, y);
    case ORC_OP_DYY: return RS(mono_deriv)(e, f, 0, 2, x, y);
    case ORC_OP_DXY: return RS(mono_deriv)(e, f, 1, 1, x, y);
    case ORC_OP_XK:  return RS(mono_deriv)(e, f, k, 0, x, y);
    case ORC_OP_YK:  return RS(mono_deriv)(e, f, 0, k, x, y);
    case ORC_OP_LAP:
        return sx * sx * RS(mono_deriv)(e, f, 2, 0, x, y) + sy * sy * RS(mono_deriv)(e, f, 0, 2, x, y);
    case ORC_OP_LAPK: {
        REAL sum = 0;
        for (int a = 0; a <= k; ++a) {
            REAL c = (REAL)binom(k, a) * RS(ipow)(sx * sx, a) * RS(ipow)(sy * sy, k - a);
            sum += c * RS(mono_deriv)(e, f, 2 * a, 2 * (k - a), x, y);
        }
        return sum;
    }
    }
    return 0;
}

/* Weight multiplier applied after the solve (src/generate_operator.jl:161-166,
 * src/hyperviscosity_operator.jl:159-160). */
static REAL RS(op_post_scale)(int op, int k, REAL sx, REAL sy)
{
    switch (op) {
    case ORC_OP_DX:  return sx;
    case ORC_OP_DY:  return sy;
    case ORC_OP_DXX: return sx * sx;
    case ORC_OP_DYY: return sy * sy;
    case ORC_OP_DXY: return sx * sy;
    case ORC_OP_XK:  return RS(ipow)(sx, k);
    case ORC_OP_YK:  return RS(ipow)(sy, k);
    default:         return 1;
    }
}

/* In-place LU with partial pivoting, row-major ns x ns, LAPACK getf2 order:
 * first maximal |a| in the column, full row swap, multipliers formed with the
 * reciprocal of the pivot.  Returns 0, or k+1 for an exactly zero pivot at k. */
static int RS(lu)(REAL *A, int ns, int *piv)
{
    int info = 0;
    for (int k = 0; k < ns; ++k) {
        int p = k;
        REAL best = RFABS(A[k * ns + k]);
        for (int i = k + 1; i < ns; ++i) {
            REAL v = RFABS(A[i * ns + k]);
            if (v > best) { best = v; p = i; }
        }
        piv[k] = p;
        if (p != k)
            for (int j = 0; j < ns; ++j) {
                REAL t = A[k * ns + j]; A[k * ns + j] = A[p * ns + j]; A[p * ns + j] = t;
            }
        REAL d = A[k * ns + k];
        if (d == 0) { if (!info) info = k + 1; continue; }
        REAL rd = (REAL)1 / d;
        for (int i = k + 1; i < ns; ++i) {
            REAL l = A[i * ns + k] * rd;
            A[i * ns + k] = l;
            if (l != 0)
                for (int j = k + 1; j < ns; ++j) A[i * ns + j] -= l * A[k * ns + j];
        }
    }
    return info;
}

/* Solve LU x = P b for nr right-hand sides stored row-major ns x nr in B. */
static void RS(lu_solve)(const REAL *LU, const int *piv, int ns, REAL *B, int nr)
{
    for (int k = 0; k < ns; ++k)
        if (piv[k] != k)
            for (int c = 0; c < nr; ++c) {
                REAL t = B[k * nr + c]; B[k * nr + c] = B[piv[k] * nr + c]; B[piv[k] * nr + c] = t;
            }
    for (int i = 0; i < ns; ++i)
        for (int j = 0; j < i; ++j) {
            REAL l = LU[i * ns + j];
            for (int c = 0; c < nr; ++c) B[i * nr + c] -= l * B[j * nr + c];
        }
    for (int i = ns - 1; i >= 0; --i) {
        for (int j = i + 1; j < ns; ++j) {
            REAL u = LU[i * ns + j];
            for (int c = 0; c < nr; ++c) B[i * nr + c] -= u * B[j * nr + c];
        }
        REAL d = LU[i * ns + i];
        for (int c = 0; c < nr; ++c) B[i * nr + c] /= d;
    }
}

/*
 * One output row: the weights of Y node `yk` on the stencil of its nearest X
 * node.  Follows src/generate_operator.jl:55-65 (factor), :68-79 (polynomial
 * RHS at the shifted and scaled Y point), :89-166 (RBF RHS, solve, rescale) for
 * methods 1 and 2, and :388-467 for method 3 (unscaled, centre replaced by
 * (eps, eps), RBF RHS at +xi).
 *
 * work must hold 2*ns*ns + 2*ns*nops + 4*n REAL values; ipiv ns ints.
 * W receives nops blocks of n weights in stencil order.
 */
static int RS(stencil_weights)(const double *X, const double *Y, long yk, const int *stencil,
                               int n, int m, int P, const int *ea, const int *eb, int method,
                               int nops, const int *ops, const int *opk, int solve_kind,
                               REAL *work, int *ipiv, double *W, double *scale_out)
{
    const int ns = n + P;
    REAL *A = work;                 /* ns*ns */
    REAL *Ainv = A + ns * ns;       /* ns*ns */
    REAL *B = Ainv + ns * ns;       /* ns*nops */
    REAL *S = B + ns * nops;        /* ns*nops */
    REAL *xi = S + ns * nops;       /* 2n */
    REAL *u = xi + 2 * n;           /* 2n */
    const double eps = 2.220446049250313e-16;   /* Julia eps() = 2^-52 */

    const long c = stencil[0];
    const double cx = X[2 * c], cy = X[2 * c + 1];
    REAL sx = 1, sy = 1;
    REAL yx, yy;                    /* Y point in the stencil frame */

    if (method == 3) {
        /* src/generate_operator.jl:393-410: shift only, centre := (eps, eps) */
        for (int j = 0; j < n; ++j) {
            long g = stencil[j];
            xi[2 * j] = (REAL)(X[2 * g] - cx);
            xi[2 * j + 1] = (REAL)(X[2 * g + 1] - cy);
        }
        xi[0] = (REAL)eps; xi[1] = (REAL)eps;
        yx = xi[0]; yy = xi[1];
        for (int j = 0; j < n; ++j) { u[2 * j] = xi[2 * j]; u[2 * j + 1] = xi[2 * j + 1]; }
    } else {
        /* src/scalestencil.jl:10-21 (shift by the first neighbour, per-axis scale) */
        double dxs[512], dys[512];
        double mx = 0, my = 0;
        for (int j = 0; j < n; ++j) {
            long g = stencil[j];
            dxs[j] = X[2 * g] - cx;
            dys[j] = X[2 * g + 1] - cy;
            if (fabs(dxs[j]) > mx) mx = fabs(dxs[j]);
            if (fabs(dys[j]) > my) my = fabs(dys[j]);
        }
        double sxd = 1.0 / mx, syd = 1.0 / my;
        sx = (REAL)sxd; sy = (REAL)syd;
        for (int j = 0; j < n; ++j) {
            xi[2 * j] = (REAL)(dxs[j] * sxd);
            xi[2 * j + 1] = (REAL)(dys[j] * syd);
        }
        /* src/generate_operator.jl:112-124: Yscaled = (Y - X_c) * s */
        double yxd = (Y[2 * yk] - cx) * sxd;
        double yyd = (Y[2 * yk + 1] - cy) * syd;
        yx = (REAL)yxd; yy = (REAL)yyd;
        /* :125-134: u_j = Yscaled - xi_j, exact zero components -> eps() */
        for (int j = 0; j < n; ++j) {
            double ux = yxd - dxs[j] * sxd;
            double uy = yyd - dys[j] * syd;
            if (ux == 0) ux = eps;
            if (uy == 0) uy = eps;
            u[2 * j] = (REAL)ux; u[2 * j + 1] = (REAL)uy;
        }
        if (scale_out) { scale_out[0] = sxd; scale_out[1] = syd; }
    }
    if (method == 3 && scale_out) { scale_out[0] = 1; scale_out[1] = 1; }

    /* A = [Phi P; P^T 0]  (src/interpolationmatrix.jl:2-5, src/rbfblock.jl:14-19,
     * src/polynomialblock.jl:23-27) */
    for (int i = 0; i < ns; ++i)
        for (int j = 0; j < ns; ++j) A[i * ns + j] = 0;
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < n; ++j)
            A[i * ns + j] = RS(phs_deriv)(m, 0, 0, xi[2 * i] - xi[2 * j], xi[2 * i + 1] - xi[2 * j + 1]);
        for (int q = 0; q < P; ++q) {
            REAL v = RS(mono_deriv)(ea[q], eb[q], 0, 0, xi[2 * i], xi[2 * i + 1]);
            A[i * ns + n + q] = v;
            A[(n + q) * ns + i] = v;
        }
    }

    /* RHS columns [RBF rows; polynomial rows] per operator (:157) */
    for (int o = 0; o < nops; ++o) {
        for (int j = 0; j < n; ++j)
            B[j * nops + o] = RS(op_phs)(ops[o], opk[o], m, u[2 * j], u[2 * j + 1], sx, sy);
        for (int q = 0; q < P; ++q)
            B[(n + q) * nops + o] = RS(op_mono)(ops[o], opk[o], ea[q], eb[q], yx, yy, sx, sy);
    }

    int info = RS(lu)(A, ns, ipiv);
    if (info) return 1;

    if (solve_kind == 0) {
        /* inv(A) (src/interpolationmatrix.jl:8), then stenc = inv(A) * RHS (:158) */
        for (int i = 0; i < ns; ++i)
            for (int j = 0; j < ns; ++j) Ainv[i * ns + j] = (i == j);
        RS(lu_solve)(A, ipiv, ns, Ainv, ns);
        for (int i = 0; i < n; ++i)
            for (int o = 0; o < nops; ++o) {
                REAL acc = 0;
                for (int j = 0; j < ns; ++j) acc += Ainv[i * ns + j] * B[j * nops + o];
                S[i * nops + o] = acc;
            }
    } else {
        for (int i = 0; i < ns * nops; ++i) S[i] = B[i];
        RS(lu_solve)(A, ipiv, ns, S, nops);
    }

    int bad = 0;
    for (int o = 0; o < nops; ++o) {
        REAL f = (method == 3) ? (REAL)1 : RS(op_post_scale)(ops[o], opk[o], sx, sy);
        for (int i = 0; i < n; ++i) {
            double w = (double)(f * S[i * nops + o]);
            if (!isfinite(w)) bad = 1;
            W[(long)o * n + i] = w;
        }
    }
    return bad ? 2 : 0;
}
