"""Jacobi fallback of the polynomial path, judged by the FP64 true residual,
and the learned solve end to end on held-out family members.

Failure is forced from the host: a polynomial that is not positive on
[0, b] (the library refuses it), and an SpMV budget too small for the
polynomial run to converge.  Every kernel still exits normally."""
import os

import numpy as np
import pytest

from cgxx_amd import learned, problems, spectrum

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from cgxx_amd.cghip import Context

    c = Context(0)
    yield c
    c.close()


def _upload(ctx, m):
    ctx.set_size(m.n)
    ctx.upload_crs(m.ptr, m.index, m.value)
    ctx.set_format("CRS")
    ctx.upload_jacobi(1.0 / m.diagonal())
    ctx.prepare_fp32()
    ctx.upload("K", m.ones_rhs())


def _true_rel(m, x):
    k = m.ones_rhs()
    return np.linalg.norm(k - m.matvec(x)) / np.linalg.norm(k)


@pytest.fixture(scope="module")
def system():
    return problems.anisotropic_family(100007, sides=(48,))


def test_rejected_polynomial_falls_back_to_jacobi(ctx, system):
    _upload(ctx, system)
    bad = np.array([0.5, 1.0, 0.0])  # p(0) = 0.5 - 1.0 < 0
    r = learned.solve_poly_with_fallback(ctx, bad, 2.0)
    assert r["method"] == "jacobi" and "positive" in r["rejected"]
    assert r["poly_iterations"] == 0 and r["jacobi_iterations"] > 0
    x = ctx.download("X")
    assert r["true_residual"] <= 1e-9 and _true_rel(system, x) <= 1.05e-9


def test_exhausted_budget_continues_with_jacobi(ctx, system):
    _upload(ctx, system)
    b = spectrum.gershgorin_bound(system.ptr, system.index, system.value)
    c = spectrum.cheb_residual_coeffs(b / 100, b, learned.DEGREE)
    r = learned.solve_poly_with_fallback(ctx, c, b, budget=3 * (learned.DEGREE + 1))
    assert r["method"] == "poly+jacobi"
    assert r["poly_iterations"] == 3 and r["jacobi_iterations"] > 0
    x = ctx.download("X")
    rel = _true_rel(system, x)
    assert rel <= 1.05e-9 and abs(rel - r["true_residual"]) <= 1e-3 * rel
    assert np.abs(x - 1.0).max() <= 1e-5


def test_polynomial_alone_within_budget(ctx, system):
    _upload(ctx, system)
    b = spectrum.gershgorin_bound(system.ptr, system.index, system.value)
    c = spectrum.cheb_residual_coeffs(b / 100, b, learned.DEGREE)
    r = learned.solve_poly_with_fallback(ctx, c, b)
    assert r["method"] == "poly" and r["jacobi_iterations"] == 0
    assert _true_rel(system, ctx.download("X")) <= 1.05e-9


@pytest.mark.skipif(not os.path.exists(learned.MODEL_PATH), reason="no trained model file")
def test_learned_solve_end_to_end(ctx):
    """The committed network on held-out seeds: converges to the FP64 true
    residual, and when it chooses the polynomial the library accepts it
    (no fallback to Jacobi)."""
    from cgxx_amd.cghip import Mlp

    model = learned.load_model()
    mlp = Mlp(tuple(int(v) for v in model["dims"]), 1)
    mlp.set_params(model["params"])
    for seed in (100001, 100002, 100003):
        m = problems.anisotropic_family(seed, sides=(64,))
        _upload(ctx, m)
        r = learned.solve_learned(ctx, mlp, model, m.n)
        assert r["method"] in ("poly", "jacobi-selected"), r
        assert _true_rel(m, ctx.download("X")) <= 1.05e-9
    mlp.close()
