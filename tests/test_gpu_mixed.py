"""The FP32-inner / FP64 reliable-update solver on the GPU (Jacobi as the
scaled operator D^-1/2 A D^-1/2, no preconditioner on the FP32 copy of A).

No reference counterpart (the reference is FP64 throughout), so it is pinned
by (1) bit-exactness against the oracle's kernel-by-kernel restatement with
the same casts and summation orders, and (2) the FP64 true residual of the
returned x recomputed independently (numpy), and agreement with the FP64
solution.
"""
import numpy as np
import pytest

from cgxx_amd import problems
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


# layout of the scaled operator: the stencils lie on 4 or 6 diagonals (DIA),
# the random matrices have short rows on many diagonals (ELL), zipf_3000 has
# rows of up to 202 entries (CRS row blocks): every layout and dot tree
LAYOUT = {"lap2d_16": "DIA", "lap2d_64": "DIA", "rand_sym_40": "ELL", "rand_gen_300": "ELL",
          "diff3d_24": "DIA", "lap3d_16": "DIA", "lap2d_256": "DIA", "zipf_3000": "CRS"}
CASES = list(LAYOUT)


@pytest.mark.parametrize("precond", ["jacobi", "none"])
@pytest.mark.parametrize("name", CASES)
def test_mixed_bitwise_vs_oracle(ctx, name, precond, tmp_path):
    fx = [f for f in load_golden() if f["case"] == name][0]
    case = Case(fx, str(tmp_path))
    _setup(ctx, *case.crs, case.k, case.C)
    assert ctx.fp32_layout() == LAYOUT[name]
    info = ctx.solve_mixed(precond=precond, max_iter=20000, tol=1e-9)
    x = ctx.download("X")
    ox, oinfo, oh = oracle.mixed_pcg(n=case.n, crs=case.crs, k=case.k, precond=precond,
                                     C_jac=case.C, max_iter=20000, tol=1e-9)
    assert info["iterations"] == oinfo["iterations"]
    assert info["updates"] == oinfo["updates"]
    it = info["iterations"]
    assert_bitwise(ctx.history("r2", it + 1), oh["r2"], "inner r2")
    assert_bitwise(ctx.history("alpha", it), oh["alpha"], "alpha")
    assert_bitwise(x, ox, "x")
    assert info["true_residual"] == oinfo["true_residual"]
    # independent FP64 check of the returned solution
    m = case.gen
    r = case.k - m.matvec(x)
    assert np.linalg.norm(r) / np.linalg.norm(case.k) <= 1.05e-9
    assert np.abs(x - 1.0).max() <= 1e-5


def test_mixed_max_iter_and_delta(ctx, tmp_path):
    fx = [f for f in load_golden() if f["case"] == "lap2d_64"][0]
    case = Case(fx, str(tmp_path))
    for max_iter, delta in [(37, 0.1), (1000, 0.5), (1000, 0.01)]:
        _setup(ctx, *case.crs, case.k, case.C)
        info = ctx.solve_mixed(max_iter=max_iter, delta=delta)
        ox, oinfo, _ = oracle.mixed_pcg(n=case.n, crs=case.crs, k=case.k, C_jac=case.C,
                                        max_iter=max_iter, delta=delta)
        assert (info["iterations"], info["updates"]) == (oinfo["iterations"], oinfo["updates"])
        assert_bitwise(ctx.download("X"), ox, f"x max_iter={max_iter} delta={delta}")
    assert info["iterations"] <= 1000


def test_mixed_profile_and_batch_agree(ctx, tmp_path):
    fx = [f for f in load_golden() if f["case"] == "diff3d_24"][0]
    case = Case(fx, str(tmp_path))
    ref = None
    for batch, profile in [(0, False), (1, False), (5, False), (0, True)]:
        _setup(ctx, *case.crs, case.k, case.C)
        info = ctx.solve_mixed(batch=batch, profile=profile)
        x = ctx.download("X")
        if ref is None:
            ref = (info["iterations"], x)
        assert info["iterations"] == ref[0]
        assert_bitwise(x, ref[1], f"batch={batch} profile={profile}")
        if profile:
            assert info["n_matvec"] == info["iterations"] and info["ms_matvec"] > 0
            assert info["n_refresh"] == info["updates"] + 1


def test_mixed_full_size_c2(ctx):
    """128^3 Laplacian (C2): FP64 true residual <= 1e-9, x close to the FP64 solve."""
    m = problems.laplace3d(128)
    k = m.ones_rhs()
    C = 1.0 / m.diagonal()
    _setup(ctx, m.ptr, m.index, m.value, k, C)
    info = ctx.solve_mixed()
    x_mixed = ctx.download("X")
    nr, nb = ctx.true_residual("X", "K")
    assert nr / nb <= 1e-9 and abs(nr / nb - info["true_residual"]) <= 1e-12
    ctx.fill("X", 0.0)
    info64 = ctx.solve(precond="jacobi")
    x64 = ctx.download("X")
    assert info["iterations"] <= 1.3 * info64["iterations"]
    assert np.abs(x_mixed - x64).max() <= 1e-5
    assert np.abs(x_mixed - 1.0).max() <= 1e-5


@pytest.mark.parametrize("side", [84, 104])
@pytest.mark.parametrize("precond", ["jacobi", "poly"])
def test_mixed_dia_grid_above_2048_vs_oracle(ctx, precond, side):
    """Variable-coefficient diffusion on 84^3 (592,704 rows) and 104^3
    (1,124,864 rows): the DIA kernels run 2316 blocks, more than the 2048 of
    the ELL and CRS kernels, and 4096 blocks (the DIA cap; 4394 chunks), so
    the <p,q> and <r,z> partials use the DIA grid, below and at its cap.
    Bitwise against the oracle."""
    from cgxx_amd import spectrum

    print(f"DIA grid parity: {side}^3 {precond}", flush=True)
    m = problems.diffusion3d(side, sigma=1.0, corr=8.0, seed=7)
    k = m.ones_rhs()
    C = 1.0 / m.diagonal()
    _setup(ctx, m.ptr, m.index, m.value, k, C)
    assert ctx.fp32_layout() == "DIA"
    poly = None
    if precond == "poly":
        b = spectrum.gershgorin_bound(m.ptr, m.index, m.value)
        poly = (spectrum.cheb_residual_coeffs(b / 100.0, b, 4), b)
        ctx.set_poly(*poly)
    info = ctx.solve_mixed(precond=precond, max_iter=20000, tol=1e-9)
    x = ctx.download("X")
    ox, oinfo, oh = oracle.mixed_pcg(n=m.n, crs=(m.ptr, m.index, m.value), k=k, precond=precond,
                                     C_jac=C, poly=poly, max_iter=20000, tol=1e-9)
    assert (info["iterations"], info["updates"]) == (oinfo["iterations"], oinfo["updates"])
    it = info["iterations"]
    assert_bitwise(ctx.history("r2", it + 1), oh["r2"], "inner r2")
    assert_bitwise(ctx.history("alpha", it), oh["alpha"], "alpha")
    assert_bitwise(x, ox, "x")
    r = k - m.matvec(x)
    assert np.linalg.norm(r) / np.linalg.norm(k) <= 1.05e-9


def test_dia_refused_when_fp16_would_overflow(ctx):
    """A tridiagonal matrix whose scaled off-diagonal entries reach 1e6
    (a_ii = 1e-6, a_i,i+-1 = -1: positive diagonal, not SPD) cannot keep an
    FP16 copy of its values, so the scaled operator goes to ELL; the same
    pattern with a_ii = 4 is DIA."""
    n = 1000
    for diag, layout in ((1e-6, "ELL"), (4.0, "DIA")):
        rows, cols, vals = [], [], []
        for i in range(n):
            for j, v in ((i - 1, -1.0), (i, diag), (i + 1, -1.0)):
                if 0 <= j < n:
                    rows.append(i), cols.append(j), vals.append(v)
        ptr = np.searchsorted(np.array(rows), np.arange(n + 1)).astype(np.int32)
        ctx.set_size(n)
        ctx.upload_crs(ptr, np.array(cols, np.int32), np.array(vals))
        ctx.set_format("CRS")
        ctx.prepare_fp32()
        assert ctx.fp32_layout() == layout
