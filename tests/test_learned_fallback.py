"""CPU checks of the Jacobi fallback's decision logic (learned.py) with a stub
context: a polynomial solve that ends with a NaN true residual must not be
reported as converged, and Jacobi then restarts from X = 0."""
import math

import numpy as np

from cgxx_amd import learned


class StubCtx:
    """Records the calls of solve_poly_with_fallback; the polynomial solve
    returns `poly_res`, a Jacobi solve converges to 1e-10 of its start."""

    def __init__(self, poly_res):
        self.poly_res = poly_res
        self.calls = []

    def fill(self, vec, value):
        self.calls.append(("fill", vec, value))

    def set_poly(self, coeffs, b):
        self.calls.append(("set_poly",))

    def solve_mixed(self, precond, max_iter, tol, **kw):
        self.calls.append(("solve", precond, tol))
        res = self.poly_res if precond == "poly" else 1e-10
        return {"iterations": 7, "updates": 1, "true_residual": res}


def test_nan_residual_falls_back_to_jacobi_from_zero():
    ctx = StubCtx(float("nan"))
    r = learned.solve_poly_with_fallback(ctx, np.ones(9), 2.0, tol=1e-9)
    assert r["method"] == "poly+jacobi"
    # X cleared before the Jacobi run, which targets the full tolerance
    solves = [c for c in ctx.calls if c[0] == "solve"]
    assert [s[1] for s in solves] == ["poly", "jacobi"]
    assert solves[1][2] == 1e-9
    i_j = ctx.calls.index(solves[1])
    assert ("fill", "X", 0.0) in ctx.calls[1:i_j]
    assert math.isfinite(r["true_residual"]) and r["true_residual"] <= 1e-9


def test_converged_polynomial_keeps_its_result():
    ctx = StubCtx(5e-10)
    r = learned.solve_poly_with_fallback(ctx, np.ones(9), 2.0, tol=1e-9)
    assert r["method"] == "poly" and r["jacobi_iterations"] == 0
    assert r["true_residual"] == 5e-10


def test_unconverged_polynomial_continues_from_its_x():
    ctx = StubCtx(1e-6)
    r = learned.solve_poly_with_fallback(ctx, np.ones(9), 2.0, tol=1e-9)
    assert r["method"] == "poly+jacobi"
    solves = [c for c in ctx.calls if c[0] == "solve"]
    assert solves[1][2] == 1e-9 / 1e-6  # relative to where the polynomial stopped
    assert r["true_residual"] == 1e-6 * 1e-10
