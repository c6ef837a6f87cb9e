"""Symmetric FP16 storage of the Clenshaw steps' operator values.

When the scaled operator's DIA diagonals are the three mirrored pairs of a
3D 7-point stencil and every FP16 value equals its transpose's bit for
bit, the library keeps only the upper diagonals (one stream of halves
each) and each row reads its lower values from the rows i - o; the
products and their order are those of the FP16 row records, so the solves stay bit for bit equal to the oracle's
mixed restatement.  One perturbed entry makes the values asymmetric: the
library must then keep the row records, again bit for bit, as it must for
a 2D Laplacian (two pairs, offsets +-1, +-300).  Covered: the plain DIA
walk (laplace3d(40): x-y plane 1600 rows, no strip geometry) and the strip
walk (diffusion3d(64): plane 4096 rows = 16 chunks).
"""
import numpy as np
import pytest

from cgxx_amd import problems, spectrum
from oracle import oracle
from tests.cases import assert_bitwise

pytestmark = pytest.mark.gpu

ITERS = 10


@pytest.fixture(scope="module")
def ctx():
    from cgxx_amd.cghip import Context

    c = Context(0)
    yield c
    c.close()


def _perturbed(m, row):
    """m with the value of the entry (row, row + 1) scaled by 1.01 (its
    transpose unchanged): FP16 values no longer symmetric."""
    value = m.value.copy()
    e = next(e for e in range(m.ptr[row], m.ptr[row + 1]) if m.index[e] == row + 1)
    value[e] *= 1.01
    return value


def _solve_and_compare(ctx, m, value, want_layout, iters=ITERS, degree=6):
    k = m.ones_rhs()
    ctx.set_size(m.n)
    ctx.upload_crs(m.ptr, m.index, value)
    ctx.set_format("CRS")
    ctx.upload_jacobi(1.0 / m.diagonal())
    ctx.prepare_fp32()
    ctx.upload("K", k)
    assert ctx.fp32_layout() == "DIA"
    assert ctx.poly_values_layout() == want_layout
    b = spectrum.gershgorin_bound(m.ptr, m.index, value)
    poly = (spectrum.cheb_residual_coeffs(b / 100.0, b, degree), b)
    ctx.set_poly(*poly)
    ctx.fill("X", 0.0)
    print(f"\npolynomial, {want_layout} values, n = {m.n}", flush=True)
    info = ctx.solve_mixed(precond="poly", max_iter=iters, tol=1e-300)
    ox, oinfo, oh = oracle.mixed_pcg(n=m.n, crs=(m.ptr, m.index, value), k=k, precond="poly",
                                     poly=poly, max_iter=iters, tol=1e-300)
    assert info["iterations"] == oinfo["iterations"] == iters
    assert_bitwise(ctx.history("r2", iters + 1), oh["r2"], "inner r2")
    assert_bitwise(ctx.history("rho", iters + 1), oh["rho"], "rho = <r, z>")
    assert_bitwise(ctx.history("alpha", iters), oh["alpha"], "alpha")
    assert_bitwise(ctx.download("X"), ox, "x")


CASES = {
    "laplace3d_40": lambda: problems.laplace3d(40),
    "diffusion3d_64": lambda: problems.diffusion3d(64, seed=7),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_symmetric_storage_bitwise(ctx, name):
    m = CASES[name]()
    _solve_and_compare(ctx, m, m.value, "FP16-SYM")


def test_two_diagonal_pairs_keep_the_records(ctx):
    m = problems.laplace2d(300)
    _solve_and_compare(ctx, m, m.value, "FP16")


def test_symmetric_storage_two_row_strip_walk(ctx):
    """176^3 = 5,451,776 rows: the strip walk two rows per trip (5.2 chunks
    per block), r and B_{j+2} cached."""
    m = problems.laplace3d(176)
    _solve_and_compare(ctx, m, m.value, "FP16-SYM", iters=3, degree=4)


def test_symmetric_storage_non_temporal(ctx):
    """Above 2^28 / 22 rows: non-temporal r and B_{j+2}, four rows per
    strip-walk trip, as at 256^3 (272^3 = 20,123,648 rows: plane 73,984
    rows = 289 chunks)."""
    m = problems.laplace3d(272)
    assert 22 * m.n > 1 << 28
    _solve_and_compare(ctx, m, m.value, "FP16-SYM", iters=3, degree=4)


@pytest.mark.parametrize("name", sorted(CASES))
def test_asymmetric_values_keep_the_records(ctx, name):
    m = CASES[name]()
    _solve_and_compare(ctx, m, _perturbed(m, m.n // 2), "FP16")


def test_fp32_values_and_byte_model(ctx):
    """With FP32 values the steps read the CG operator's values; the
    iteration's byte model counts 2 bytes per stored pair with symmetric
    storage and 4 with the records."""
    m = problems.laplace3d(40)
    _solve_and_compare(ctx, m, m.value, "FP16-SYM")
    sym = ctx.mixed_iteration_bytes("poly")
    ctx.set_poly_values(32)
    try:
        assert ctx.poly_values_layout() == "FP32"
    finally:
        ctx.set_poly_values(16)
    _solve_and_compare(ctx, m, _perturbed(m, 7), "FP16")
    rec = ctx.mixed_iteration_bytes("poly")
    # degree 6, 6 diagonals: 6 steps of 12 n record bytes against 6 n
    assert np.isclose(rec - sym, 6 * 6 * m.n)
