"""Pins the oracle's closed-form basis functions against sympy (an independent
symbolic engine, as the reference uses Symbolics.jl: src/rbfbasis.jl:19-30,
src/rbfbasis_k.jl:14-18) and against the 2-D Laplacian closed form."""
import math

import numpy as np
import pytest
import sympy as sy


@pytest.mark.parametrize("m", [3, 5, 7, 9])
def test_phs_derivatives_match_sympy(O, m):
    x, y = sy.symbols("x y", real=True)
    phi = sy.sqrt(x**2 + y**2) ** m
    rng = np.random.default_rng(m)
    pts = rng.uniform(-1.5, 1.5, size=(6, 2))
    pts[0] = (0.3, 0.0)      # an axis-aligned point (one coordinate exactly zero)
    pts[1] = (2.2e-16, -0.7)  # the eps substitution regime of the reference
    for a in range(0, 7):
        for b in range(0, 7 - a):
            if a + b > m + 3:
                continue
            expr = sy.diff(phi, x, a, y, b) if (a or b) else phi
            f = sy.lambdify((x, y), expr, "mpmath")
            for vx, vy in pts:
                ref = float(f(sy.Float(vx, 40), sy.Float(vy, 40)))
                got = O.phs_deriv(m, a, b, vx, vy)
                gq = O.phs_deriv(m, a, b, vx, vy, quad=True)
                # the derivative is homogeneous of degree m - a - b; its terms
                # are of size (m + a + b)^(a + b) r^(m - a - b), which bounds the
                # cancellation error of any FP64 evaluation
                r = math.hypot(vx, vy)
                terms = (m + a + b) ** (a + b) * r ** (m - a - b)
                assert abs(got - ref) <= 1e-14 * terms + 1e-13 * abs(ref), (m, a, b, vx, vy, got, ref)
                assert abs(gq - ref) <= 1e-30 * terms + 2.3e-16 * abs(ref), (m, a, b, vx, vy, gq, ref)


@pytest.mark.parametrize("m,k", [(5, 1), (7, 1), (7, 2), (9, 3), (9, 2), (7, 3)])
def test_laplacian_powers_closed_form(O, m, k):
    # Delta^k r^m = prod_{j<k} (m - 2j)^2 r^(m - 2k) in two dimensions
    c = 1.0
    for j in range(k):
        c *= (m - 2 * j) ** 2
    rng = np.random.default_rng(100 + m + k)
    for vx, vy in rng.uniform(-1, 1, size=(5, 2)):
        r = math.hypot(vx, vy)
        ref = c * r ** (m - 2 * k)
        got = O.op_phs(O.OP_LAPK, k, m, vx, vy)
        assert abs(got - ref) <= 1e-12 * abs(ref)
        if k == 1:
            assert abs(O.op_phs(O.OP_LAP, 0, m, vx, vy) - ref) <= 1e-12 * abs(ref)


def test_scaled_laplacian_is_the_physical_one(O):
    # with per-axis scales the operator is sx^2 d_xixi + sy^2 d_etaeta: check it
    # against sympy on phi(sx*X, sy*Y) differentiated in the physical frame
    X, Y = sy.symbols("X Y", real=True)
    sx, syy, m = 3.0, 0.5, 7
    phi = sy.sqrt((sx * X) ** 2 + (syy * Y) ** 2) ** m
    lap2 = sy.diff(phi, X, 4) + 2 * sy.diff(phi, X, 2, Y, 2) + sy.diff(phi, Y, 4)
    f = sy.lambdify((X, Y), lap2, "mpmath")
    for xi, eta in [(0.3, -0.2), (-0.7, 0.55)]:
        ref = float(f(sy.Float(xi / sx, 40), sy.Float(eta / syy, 40)))
        got = O.op_phs(O.OP_LAPK, 2, m, xi, eta, sx, syy)
        assert abs(got - ref) <= 1e-11 * abs(ref)


@pytest.mark.parametrize("d", range(0, 8))
def test_monomial_basis_complete(O, d):
    ea, eb = O.monomials(d)
    assert len(ea) == (d + 1) * (d + 2) // 2
    pairs = set(zip(ea.tolist(), eb.tolist()))
    assert pairs == {(a, t - a) for t in range(d + 1) for a in range(t + 1)}
    assert (ea[0], eb[0]) == (0, 0)
