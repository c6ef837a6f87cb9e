"""The vector kernels' in-flight paths, bit for bit against the oracle.

Every vector kernel walks element pairs (FP64) or quads (FP32) with several
in flight per thread and takes what a thread has left one at a time.  The
in-flight path runs only when a thread owns several pairs or quads: above
4 * 2 * 2048 * 256 = 4,194,304 rows for the FP64 kernels (kUnroll = 4
pairs, grids of at most 2048 blocks) and 2 * 4 * 2048 * 256 for the FP32
ones.  The other parity tests stay below that, so these run 30 iterations
on 2D Laplacians of 2200^2 = 4,840,000 rows and 2201^2 = 4,844,401 rows
(odd: the last pair and quad are partial): the fused FP64 loop in CRS and
ELL against oracle order 1, and the mixed solver (Jacobi, and the degree-4
polynomial whose Clenshaw steps feed the y/p update) with a reliable
update every few iterations (delta = 0.5: the restart, residual and
accumulate kernels run at this size too) against the oracle's mixed
restatement.
Every r2 and alpha and the final x are compared bitwise.
"""
import numpy as np
import pytest

from cgxx_amd import problems, spectrum
from oracle import oracle
from tests.cases import assert_bitwise

pytestmark = pytest.mark.gpu

ITERS = 30


@pytest.fixture(scope="module")
def ctx():
    from cgxx_amd.cghip import Context

    c = Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("side", [2200, 2201])
def test_fused_fp64_loop_in_flight_path_bitwise(ctx, side):
    m = problems.laplace2d(side)
    assert m.n > 4 * 2 * 2048 * 256
    k, c = m.ones_rhs(), 1.0 / m.diagonal()
    width, length, eidx, edat = m.to_ell()
    ctx.set_size(m.n)
    ctx.upload_crs(m.ptr, m.index, m.value)
    ctx.upload_ell(width, length, eidx, edat)
    ctx.upload_jacobi(c)
    ctx.upload("K", k)
    for fmt in ("CRS", "ELL"):
        print(f"\nfused FP64 loop, {fmt}, n = {m.n}", flush=True)
        ctx.set_format(fmt)
        ctx.fill("X", 0.0)
        info = ctx.solve(precond="jacobi", max_iter=ITERS, tol=1e-300)
        assert info["iterations"] == ITERS
        kw = dict(crs=(m.ptr, m.index, m.value)) if fmt == "CRS" else \
            dict(ell=(width, length, eidx, edat))
        ox, oit, _, oh = oracle.pcg(n=m.n, fmt=fmt, precond="jacobi", order=1, k=k, C_jac=c,
                                    max_iter=ITERS, tol=1e-300, **kw)
        assert oit == ITERS
        assert_bitwise(ctx.history("r2", ITERS + 1), oh["r2"], f"r2 ({fmt})")
        assert_bitwise(ctx.history("alpha", ITERS), oh["alpha"], f"alpha ({fmt})")
        assert_bitwise(ctx.download("X"), ox, f"x ({fmt})")


@pytest.mark.parametrize("precond", ["jacobi", "poly"])
@pytest.mark.parametrize("side", [2200, 2201])
def test_mixed_in_flight_path_and_updates_bitwise(ctx, side, precond):
    m = problems.laplace2d(side)
    k = m.ones_rhs()
    ctx.set_size(m.n)
    ctx.upload_crs(m.ptr, m.index, m.value)
    ctx.set_format("CRS")
    ctx.upload_jacobi(1.0 / m.diagonal())
    ctx.prepare_fp32()
    ctx.upload("K", k)
    poly = None
    if precond == "poly":  # degree 4 on [b / 100, b]: the Clenshaw steps and their y/p update
        b = spectrum.gershgorin_bound(m.ptr, m.index, m.value)
        poly = (spectrum.cheb_residual_coeffs(b / 100.0, b, 4), b)
        ctx.set_poly(*poly)
    ctx.fill("X", 0.0)
    print(f"\nmixed {precond}, delta 0.5, n = {m.n}, layout {ctx.fp32_layout()}", flush=True)
    info = ctx.solve_mixed(precond=precond, max_iter=ITERS, tol=1e-300, delta=0.5)
    ox, oinfo, oh = oracle.mixed_pcg(n=m.n, crs=(m.ptr, m.index, m.value), k=k, precond=precond,
                                     poly=poly, max_iter=ITERS, tol=1e-300, delta=0.5)
    assert info["iterations"] == oinfo["iterations"] == ITERS
    assert info["updates"] == oinfo["updates"] >= 2
    assert_bitwise(ctx.history("r2", ITERS + 1), oh["r2"], "inner r2")
    if precond == "poly":
        assert_bitwise(ctx.history("rho", ITERS + 1), oh["rho"], "rho = <r, z>")
    assert_bitwise(ctx.history("alpha", ITERS), oh["alpha"], "alpha")
    assert_bitwise(ctx.download("X"), ox, "x")


@pytest.mark.parametrize("bits", [16, 32])
def test_poly_four_row_strip_walk_bitwise(ctx, bits):
    """The Clenshaw steps' four-rows-per-trip strip walk, which the library
    picks only when every block walks at least 8 chunks of 256 rows: a 3D
    Laplacian of 208^3 = 8,998,912 rows (35,152 chunks over 4096 blocks; its
    x-y plane of 169 chunks gives the strip geometry), the degree-4
    polynomial on FP16 and on FP32 operator values, 6 iterations bit for
    bit against the oracle."""
    m = problems.laplace3d(208)
    assert m.n // 256 >= 8 * 4096
    k = m.ones_rhs()
    ctx.set_size(m.n)
    ctx.upload_crs(m.ptr, m.index, m.value)
    ctx.set_format("CRS")
    ctx.upload_jacobi(1.0 / m.diagonal())
    ctx.prepare_fp32()
    ctx.upload("K", k)
    assert ctx.fp32_layout() == "DIA"
    b = spectrum.gershgorin_bound(m.ptr, m.index, m.value)
    poly = (spectrum.cheb_residual_coeffs(b / 100.0, b, 4), b)
    ctx.set_poly_values(bits)
    ctx.set_poly(*poly)
    ctx.fill("X", 0.0)
    print(f"\npolynomial, FP{bits} values, n = {m.n}", flush=True)
    its = 6
    info = ctx.solve_mixed(precond="poly", max_iter=its, tol=1e-300)
    ox, oinfo, oh = oracle.mixed_pcg(n=m.n, crs=(m.ptr, m.index, m.value), k=k, precond="poly",
                                     poly=poly, max_iter=its, tol=1e-300,
                                     poly_fp32_values=bits == 32)
    assert info["iterations"] == oinfo["iterations"] == its
    assert info["updates"] == oinfo["updates"]
    assert_bitwise(ctx.history("r2", its + 1), oh["r2"], "inner r2")
    assert_bitwise(ctx.history("rho", its + 1), oh["rho"], "rho = <r, z>")
    assert_bitwise(ctx.history("alpha", its), oh["alpha"], "alpha")
    assert_bitwise(ctx.download("X"), ox, "x")
    ctx.set_poly_values(16)
