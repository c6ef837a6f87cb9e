"""The polynomial preconditioner of the mixed solver on the GPU.

Bit-exact against the oracle's restatement (same FP32 Clenshaw expression
order, same reduction trees), for degrees that exercise every rotation of
the three work vectors; argument checks of cghip_set_poly; the FP64 true
residual of the result recomputed on the host.
"""
import numpy as np
import pytest

from cgxx_amd import problems, spectrum
from oracle import oracle
from tests.cases import Case, assert_bitwise
from tests.conftest import load_golden

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from cgxx_amd.cghip import Context

    c = Context(0)
    yield c
    c.close()


def _setup(ctx, ptr, idx, val, k, C):
    ctx.set_size(ptr.size - 1)
    ctx.upload_crs(ptr, idx, val)
    ctx.set_format("CRS")
    ctx.upload_jacobi(C)
    ctx.prepare_fp32()
    ctx.upload("K", k)
    ctx.upload("X", np.zeros(ptr.size - 1))


def _coeffs(ptr, idx, val, degree, ratio=100.0):
    b = spectrum.gershgorin_bound(ptr, idx, val)
    return spectrum.cheb_residual_coeffs(b / ratio, b, degree), b


@pytest.mark.parametrize("degree", [1, 2, 3, 4, 8, 16])
@pytest.mark.parametrize("name", ["lap2d_64", "diff3d_24", "rand_gen_300", "zipf_3000"])
def test_poly_bitwise_vs_oracle(ctx, name, degree, tmp_path):
    fx = [f for f in load_golden() if f["case"] == name][0]
    case = Case(fx, str(tmp_path))
    ptr, idx, val = case.crs
    _setup(ctx, ptr, idx, val, case.k, case.C)
    c, b = _coeffs(ptr, idx, val, degree)
    ctx.set_poly(c, b)
    info = ctx.solve_mixed(precond="poly", max_iter=20000, tol=1e-9)
    x = ctx.download("X")
    ox, oinfo, oh = oracle.mixed_pcg(n=case.n, crs=case.crs, k=case.k, precond="poly", C_jac=case.C,
                                     poly=(c, b), max_iter=20000, tol=1e-9)
    assert (info["iterations"], info["updates"]) == (oinfo["iterations"], oinfo["updates"])
    it = info["iterations"]
    assert_bitwise(ctx.history("r2", it + 1), oh["r2"], "inner r2")
    assert_bitwise(ctx.history("rho", it + 1), oh["rho"], "rho = <r, z>")
    assert_bitwise(ctx.history("alpha", it), oh["alpha"], "alpha")
    assert_bitwise(x, ox, "x")
    r = case.k - case.gen.matvec(x)
    assert np.linalg.norm(r) / np.linalg.norm(case.k) <= 1.05e-9
    assert np.abs(x - 1.0).max() <= 1e-5


@pytest.mark.parametrize("degree", [2, 8])
@pytest.mark.parametrize("name", ["lap2d_64", "diff3d_24"])
def test_poly_fp32_values_bitwise_vs_oracle(ctx, name, degree, tmp_path):
    """cghip_set_poly_values(32): the Clenshaw steps read the DIA operator's
    FP32 values instead of its FP16 copy; bit-exact with the oracle's
    restatement of that variant, and a different preconditioner from the
    default FP16 one.  Changing the precision drops the polynomial."""
    from cgxx_amd.cghip import CGHIPError

    fx = [f for f in load_golden() if f["case"] == name][0]
    case = Case(fx, str(tmp_path))
    ptr, idx, val = case.crs
    _setup(ctx, ptr, idx, val, case.k, case.C)
    assert ctx.fp32_layout() == "DIA"
    c, b = _coeffs(ptr, idx, val, degree)
    xs = {}
    try:
        for bits in (32, 16):
            ctx.set_poly_values(bits)
            with pytest.raises(CGHIPError, match="set_poly"):
                ctx.solve_mixed(precond="poly", max_iter=20000, tol=1e-9)
            ctx.set_poly(c, b)
            ctx.fill("X", 0.0)
            info = ctx.solve_mixed(precond="poly", max_iter=20000, tol=1e-9)
            xs[bits] = ctx.download("X")
            ox, oinfo, oh = oracle.mixed_pcg(n=case.n, crs=case.crs, k=case.k, precond="poly",
                                             poly=(c, b), max_iter=20000, tol=1e-9,
                                             poly_fp32_values=bits == 32)
            assert (info["iterations"], info["updates"]) == (oinfo["iterations"], oinfo["updates"])
            it = info["iterations"]
            assert_bitwise(ctx.history("rho", it + 1), oh["rho"], f"rho ({bits}-bit values)")
            assert_bitwise(xs[bits], ox, f"x ({bits}-bit values)")
            assert info["true_residual"] <= 1e-9
    finally:
        ctx.set_poly_values(16)
    # the 2D Laplacian's scaled off-diagonal values (-1/4) are exact in FP16:
    # the same preconditioner; the variable coefficients are not
    assert np.array_equal(xs[16], xs[32]) == (name == "lap2d_64")


def test_poly_profile_and_batch_agree(ctx, tmp_path):
    fx = [f for f in load_golden() if f["case"] == "diff3d_24"][0]
    case = Case(fx, str(tmp_path))
    ref = None
    for batch, profile in [(0, False), (1, False), (7, False), (0, True)]:
        _setup(ctx, *case.crs, case.k, case.C)
        ctx.set_poly(*_coeffs(*case.crs, 8))
        info = ctx.solve_mixed(precond="poly", batch=batch, profile=profile)
        x = ctx.download("X")
        if ref is None:
            ref = (info["iterations"], x)
        assert info["iterations"] == ref[0]
        assert_bitwise(x, ref[1], f"batch={batch} profile={profile}")
        if profile:
            assert info["n_matvec"] == info["iterations"] and info["n_refresh"] == info["updates"] + 1


def test_set_poly_rejects(ctx, tmp_path):
    from cgxx_amd.cghip import CGHIPError

    fx = [f for f in load_golden() if f["case"] == "lap2d_16"][0]
    case = Case(fx, str(tmp_path))
    _setup(ctx, *case.crs, case.k, case.C)
    with pytest.raises(CGHIPError, match="cghip_set_poly"):
        ctx.solve_mixed(precond="poly")
    bad = [
        ([1.0], 2.0),                      # degree 0
        ([1.0] * 18, 2.0),                 # degree 17
        ([1.0, 0.5], 0.0),                 # b <= 0
        ([1.0, 0.5], float("nan")),
        ([0.5, 1.0], 2.0),                 # p(0) = 0.5 - 1.0 < 0
        ([1.0, float("inf")], 2.0),
    ]
    for c, b in bad:
        with pytest.raises(CGHIPError):
            ctx.set_poly(c, b)
    # a matrix upload drops the polynomial with the FP32 data
    ctx.set_poly(*_coeffs(*case.crs, 4))
    ctx.upload_crs(*case.crs)
    ctx.prepare_fp32()
    with pytest.raises(CGHIPError):
        ctx.solve_mixed(precond="poly")


@pytest.mark.parametrize("side", [96, 128])
def test_poly_strip_walk_bitwise_vs_oracle(ctx, side):
    """Grids whose x-y plane is a whole number of 256-row chunks: the
    Clenshaw steps without a dot walk the DIA operator strip by strip
    (2 strips of 18 chunks at 96^3, of 32 at 128^3); every row is computed
    alone, so 12 iterations stay bitwise equal to the oracle."""
    m = problems.anisotropic_family(100011, sides=(side,))
    k, C = m.ones_rhs(), 1.0 / m.diagonal()
    _setup(ctx, m.ptr, m.index, m.value, k, C)
    assert ctx.fp32_layout() == "DIA"
    c, b = _coeffs(m.ptr, m.index, m.value, 8, ratio=300.0)
    ctx.set_poly(c, b)
    info = ctx.solve_mixed(precond="poly", max_iter=12, tol=1e-300)
    ox, oinfo, oh = oracle.mixed_pcg(n=m.n, crs=(m.ptr, m.index, m.value), k=k, precond="poly",
                                     C_jac=C, poly=(c, b), max_iter=12, tol=1e-300)
    assert (info["iterations"], info["updates"]) == (oinfo["iterations"], oinfo["updates"])
    assert info["iterations"] == 12
    assert_bitwise(ctx.history("rho", 13), oh["rho"], "rho = <r, z>")
    assert_bitwise(ctx.download("X"), ox, "x")


def test_poly_full_size_c2(ctx):
    """128^3 Laplacian: degree-8 polynomial reaches the FP64 true residual."""
    m = problems.laplace3d(128)
    k = m.ones_rhs()
    C = 1.0 / m.diagonal()
    _setup(ctx, m.ptr, m.index, m.value, k, C)
    ctx.set_poly(*_coeffs(m.ptr, m.index, m.value, 8, ratio=300.0))
    info = ctx.solve_mixed(precond="poly")
    nr, nb = ctx.true_residual("X", "K")
    assert nr / nb <= 1e-9
    x = ctx.download("X")
    assert np.abs(x - 1.0).max() <= 1e-5
    ctx.fill("X", 0.0)
    infoj = ctx.solve_mixed(precond="jacobi")
    assert info["iterations"] < infoj["iterations"] / 3


def test_time_poly_step(ctx, tmp_path):
    from cgxx_amd.cghip import CGHIPError

    fx = [f for f in load_golden() if f["case"] == "diff3d_24"][0]
    case = Case(fx, str(tmp_path))
    _setup(ctx, *case.crs, case.k, case.C)
    with pytest.raises(CGHIPError):
        ctx.time_poly_step()
    ctx.set_poly(*_coeffs(*case.crs, 8))
    assert ctx.time_poly_step(reps=3) > 0
    # the solver still runs bit-exactly after the timing overwrote its buffers
    info = ctx.solve_mixed(precond="poly")
    ox, oinfo, _ = oracle.mixed_pcg(n=case.n, crs=case.crs, k=case.k, precond="poly",
                                    poly=_coeffs(*case.crs, 8))
    assert info["iterations"] == oinfo["iterations"]
    assert_bitwise(ctx.download("X"), ox, "x after timing")
