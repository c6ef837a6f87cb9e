"""Parity of the HIP path with the oracle and with the reference's own output.

Tiers (all through the C ABI, include/cghip.h):
  * hooks: every kernel bit-exact against the oracle's serial kernels, and
    the tree dot bit-exact against the oracle's restated tree order;
  * the reference loop driven hook by hook with left-to-right dots: bit-exact
    against the reference's golden histories (r2, rho, dot_pq, a, b, x[:10]);
  * the same loop with tree dots: bit-exact against oracle order 2;
  * the fused device loop: bit-exact against oracle order 1, and within
    1e-10 relative per iteration of the reference's history while the
    residual is above 1e-6 -- the tolerance the FP64 path states -- and down
    to the stop test within 1e-10 or the measured worst of its tree-order
    sums (GPU_R2_WORST; C1 1.5e-12, C2 7.7e-10), iterations +-1, the
    CG_CHECK_TOLERANCE check passing.
"""
import numpy as np
import pytest

from oracle import oracle
from tests.cases import Case, assert_bitwise, golden_sequences, max_rel, run_env
from tests.conftest import golden_runs, load_golden, run_id

pytestmark = pytest.mark.gpu

# COO runs execute on the device CRS (the stable CRS of the COO entries adds
# each row in COO order, as SerialCG's COO SpMV does), so the reference's COO
# histories are targets of the device runs too.
SMALL = golden_runs(max_n=70_000)
DEVICE_RUNS = SMALL
ALL_DEVICE_RUNS = golden_runs()


@pytest.fixture(scope="module")
def ctx():
    from cgxx_amd.cghip import Context

    c = Context(0)
    yield c
    c.close()


def device_fmt(fmt: str) -> str:
    """The device format a run's CG_MATRIX_FORMAT executes in."""
    return "ELL" if fmt == "ELL" else "CRS"


def _upload(ctx, case: Case, fmt: str, precond: str):
    if fmt == "ELL":
        ctx.upload_ell(*case.ell)
    else:  # CRS, and COO on the device CRS
        ctx.upload_crs(*case.crs)
    ctx.set_format("ELL" if fmt == "ELL" else "CRS")
    if precond == "jacobi":
        ctx.upload_jacobi(case.C)
    ctx.upload("K", case.k)
    ctx.upload("X", np.zeros(case.n))


def _fresh(ctx, case, fmt, precond):
    ctx.set_size(case.n)
    _upload(ctx, case, fmt, precond)


@pytest.mark.parametrize("name", ["lap2d_16", "rand_sym_40", "rand_gen_300", "diff3d_24", "zipf_3000"])
def test_hooks_bitwise(ctx, name, tmp_path):
    fx = [f for f in load_golden() if f["case"] == name][0]
    case = Case(fx, str(tmp_path))
    n = case.n
    rng = np.random.default_rng(0)
    x = rng.standard_normal(n)
    y = rng.standard_normal(n)
    L = oracle.lib()
    for fmt in ("CRS", "ELL"):
        _fresh(ctx, case, fmt, "jacobi")
        ctx.upload("P", x)
        ctx.matvec("P", "Q")
        ref = np.zeros(n)
        if fmt == "CRS":
            L.orc_spmv_crs(n, *case.crs, x, ref)
        else:
            L.orc_spmv_ell(n, case.ell[1], case.ell[2], case.ell[3], x, ref)
        assert_bitwise(ctx.download("Q"), ref, f"{fmt} matvec")
    ctx.upload("P", x)
    ctx.upload("Q", y)
    ctx.axpy(0.37, "P", "Q")
    ref = y.copy()
    L.orc_axpy(n, 0.37, x, ref)
    assert_bitwise(ctx.download("Q"), ref, "axpy")
    ctx.upload("Q", y)
    ctx.xpay("P", -1.3, "Q")
    ref = y.copy()
    L.orc_xpay(n, x, -1.3, ref)
    assert_bitwise(ctx.download("Q"), ref, "xpay")
    ctx.apply_preconditioner("P", "Z")
    ref = np.zeros(n)
    L.orc_jacobi_apply(n, case.C, x, ref)
    assert_bitwise(ctx.download("Z"), ref, "jacobi")
    ctx.set_dot_mode("serial")
    assert ctx.dot("P", "Q") == L.orc_dot_serial(n, x, ctx.download("Q"))
    ctx.set_dot_mode("tree")
    assert ctx.dot("P", "Q") == L.orc_dot_tree(n, x, ctx.download("Q"))
    ctx.cpy("R", "P")
    assert_bitwise(ctx.download("R"), x, "cpy")


@pytest.mark.parametrize("n", [1, 63, 64, 65, 255, 257, 4096, 524288, 524289, 2_000_003])
def test_dot_tree_geometry(ctx, n):
    rng = np.random.default_rng(n)
    a = rng.standard_normal(n)
    b = rng.standard_normal(n)
    ctx.set_size(n)
    ctx.upload("P", a)
    ctx.upload("Q", b)
    ctx.set_dot_mode("tree")
    assert ctx.dot("P", "Q") == oracle.lib().orc_dot_tree(n, a, b)
    if n <= 65536:
        ctx.set_dot_mode("serial")
        assert ctx.dot("P", "Q") == oracle.lib().orc_dot_serial(n, a, b)
        ctx.set_dot_mode("tree")


def test_long_rows_and_empty_rows(ctx):
    """Rows above the LDS capacity (chunked path) and rows with no entry."""
    rng = np.random.default_rng(5)
    n = 3000
    lens = rng.integers(0, 6, n)
    lens[7] = 2900  # longer than one LDS chunk
    lens[2500] = 4500 % n
    lens[100] = 0
    cols = [np.sort(rng.choice(n, size=l, replace=False)) for l in lens]
    ptr = np.zeros(n + 1, np.int64)
    ptr[1:] = np.cumsum(lens)
    idx = np.concatenate(cols).astype(np.int32)
    val = rng.standard_normal(idx.size)
    ptr = ptr.astype(np.int32)
    x = rng.standard_normal(n)
    ctx.set_size(n)
    ctx.upload_crs(ptr, idx, val)
    ctx.set_format("CRS")
    ctx.upload("P", x)
    ctx.matvec("P", "Q")
    ref = np.zeros(n)
    oracle.lib().orc_spmv_crs(n, ptr, idx, val, x, ref)
    assert_bitwise(ctx.download("Q"), ref, "long-row CRS")
    from tests.cases import to_ell

    w, length, eidx, edat = to_ell(ptr, idx, val)
    ctx.upload_ell(w, length, eidx, edat)
    ctx.set_format("ELL")
    ctx.matvec("P", "Q")
    assert_bitwise(ctx.download("Q"), ref, "long-row ELL")


def test_invalid_input_is_rejected(ctx):
    from cgxx_amd.cghip import CGHIPError

    ptr = np.array([0, 1, 2], np.int32)
    with pytest.raises(CGHIPError, match="out of range"):
        ctx.upload_crs(ptr, np.array([0, 2], np.int32), np.ones(2))
    with pytest.raises(CGHIPError, match="ptr"):
        ctx.upload_crs(np.array([0, 2, 1], np.int32), np.array([0, 1], np.int32), np.ones(2))
    with pytest.raises(CGHIPError, match="length"):
        ctx.upload_ell(1, np.array([2, 0], np.int32), np.zeros(2, np.int32), np.ones(2))


@pytest.mark.parametrize("pair", DEVICE_RUNS, ids=[run_id(p) for p in DEVICE_RUNS])
def test_reference_loop_serial_dots_bitwise(ctx, pair, tmp_path):
    """The reference's loop, hook by hook on the GPU, left-to-right dots."""
    from cgxx_amd.cghip import reference_loop

    fx, run = pair
    case = Case(fx, str(tmp_path))
    fmt, pre, maxit, tol = run_env(run)
    g = golden_sequences(run)
    _fresh(ctx, case, fmt, pre)
    ctx.set_dot_mode("serial")
    try:
        it, res, h = reference_loop(ctx, pre, maxit, tol)
    finally:
        ctx.set_dot_mode("tree")
    assert it == run["iterations"]
    assert_bitwise(h["r2"], g["r2"], "r2")
    assert_bitwise(h["rho"][0], g["rho0"], "rho0")
    assert_bitwise(h["pq"], g["pq"], "dot_pq")
    assert_bitwise(h["alpha"], g["alpha"], "a")
    assert_bitwise(h["beta"], g["beta"], "b")
    x = ctx.download("X")
    assert_bitwise(x[: g["x_head"].size], g["x_head"], "x[:10]")
    assert f"{res:e}" == run["residual_printed"]


@pytest.mark.parametrize("pair", DEVICE_RUNS, ids=[run_id(p) for p in DEVICE_RUNS])
def test_reference_loop_tree_dots_match_oracle(ctx, pair, tmp_path):
    from cgxx_amd.cghip import reference_loop

    fx, run = pair
    case = Case(fx, str(tmp_path))
    fmt, pre, maxit, tol = run_env(run)
    _fresh(ctx, case, fmt, pre)
    it, res, h = reference_loop(ctx, pre, maxit, tol)
    ox, oit, ores, oh = oracle.pcg(n=case.n, fmt=device_fmt(fmt), precond=pre, order=2, k=case.k, tol=tol,
                                   max_iter=maxit, crs=case.crs, ell=case.ell, C_jac=case.C)
    assert it == oit
    assert_bitwise(h["r2"], oh["r2"], "r2")
    assert_bitwise(h["pq"], oh["pq"], "dot_pq")
    assert_bitwise(ctx.download("X"), ox, "x")


# Worst relative r2 deviation of the fused (tree-order) loop from the
# reference's serial-order history, over every iteration with residual above
# 1e-12, where it exceeds 1e-10 (measured; the loop is bit-exact with the
# oracle's order 1, so these are the oracle's figures): lap3d_128 (C2, CRS
# and ELL) 7.67e-10 (4.91e-11 above residual 1e-6), diff3d_24 1.41e-10,
# lap2d_16 1.32e-10, rand_sym_40 4.56e-9.  C1 (lap2d_256): 1.51e-12.
GPU_R2_WORST = {"lap3d_128": 8e-10, "diff3d_24": 1.5e-10, "lap2d_16": 1.4e-10, "rand_sym_40": 5e-9}


@pytest.mark.parametrize("pair", ALL_DEVICE_RUNS, ids=[run_id(p) for p in ALL_DEVICE_RUNS])
def test_fused_solve(ctx, pair, tmp_path):
    """Fused device loop: bit-exact vs oracle order 1; 1e-10 vs the reference."""
    fx, run = pair
    case = Case(fx, str(tmp_path))
    fmt, pre, maxit, tol = run_env(run)
    g = golden_sequences(run)
    _fresh(ctx, case, fmt, pre)
    info = ctx.solve(precond=pre, max_iter=maxit, tol=tol)
    it = info["iterations"]
    h = {k: ctx.history(k, it + 1) for k in ("r2", "rho", "pq", "alpha", "beta")}
    x = ctx.download("X")
    if case.n <= 300_000:
        ox, oit, ores, oh = oracle.pcg(n=case.n, fmt=device_fmt(fmt), precond=pre, order=1, k=case.k, tol=tol,
                                       max_iter=maxit, crs=case.crs, ell=case.ell, C_jac=case.C)
        assert it == oit
        assert_bitwise(h["r2"], oh["r2"], "r2")
        assert_bitwise(h["pq"], oh["pq"], "dot_pq")
        assert_bitwise(h["alpha"], oh["alpha"], "alpha")
        assert_bitwise(x, ox, "x")
        assert info["residual"] == ores
    # Against the reference's own run.  The fused loop sums its dots in a
    # tree, SerialCG left to right, so the two recurrences round differently
    # and drift apart as r2 falls.  Bounds, relative per iteration, over
    # every iteration down to the stop test (CG.cpp:344-355): 1e-10 while
    # the residual is above 1e-6 on every case, and through the last
    # iteration 1e-10 wherever the measured worst allows it, else that
    # measured worst rounded up (GPU_R2_WORST); below a residual of 1e-12
    # (zipf_3000's single step lands at 1e-13) r2 is rounding noise of the
    # reference's own sums and is not compared.
    assert abs(it - run["iterations"]) <= 1
    m = min(it, run["iterations"])
    a, b = h["r2"][: m + 1], g["r2"][: m + 1]
    res = np.sqrt(b / b[0])
    rel = np.abs(a - b) / np.abs(b)
    early = res > 1e-6
    assert rel[early].max() <= 1e-10, f"r2 deviates by {rel[early].max():.3e} above residual 1e-6"
    late = res > 1e-12
    bound = GPU_R2_WORST.get(fx["case"], 1e-10)
    assert rel[late].max() <= bound, f"r2 deviates by {rel[late].max():.3e} (bound {bound:.1e})"
    errors = int(np.sum(np.abs(x - 1.0) > 1e-5))
    expect_ok = run["check"] == "Solution is correct!"
    assert (errors == 0) == expect_ok


def test_fused_solve_is_deterministic(ctx, tmp_path):
    fx = [f for f in load_golden() if f["case"] == "lap2d_256"][0]
    case = Case(fx, str(tmp_path))
    outs = []
    for _ in range(2):
        _fresh(ctx, case, "CRS", "jacobi")
        info = ctx.solve(precond="jacobi")
        outs.append((info["iterations"], ctx.download("X"), ctx.history("r2", info["iterations"] + 1)))
    assert outs[0][0] == outs[1][0]
    assert_bitwise(outs[0][1], outs[1][1], "x")
    assert_bitwise(outs[0][2], outs[1][2], "r2")


def test_huge_max_iter_and_history_after_convergence(ctx, tmp_path):
    """CG_MAX_ITER = INT_MAX (accepted by CG.cpp:56): the history buffer stays
    bounded and the solves match a run with a small limit bit for bit; after
    the converging iteration rho and beta are NaN (the reference computes
    them only when the loop goes on, CG.cpp:357-370)."""
    fx = [f for f in load_golden() if f["case"] == "lap2d_64"][0]
    case = Case(fx, str(tmp_path))
    out = []
    for maxit in (1000, 2**31 - 1):
        _fresh(ctx, case, "CRS", "jacobi")
        info = ctx.solve(precond="jacobi", max_iter=maxit)
        it = info["iterations"]
        out.append((it, ctx.download("X"), ctx.history("r2", it + 1), ctx.history("rho", it + 1),
                    ctx.history("beta", it)))
    assert out[0][0] == out[1][0] < 1000
    for a, b in zip(out[0][1:], out[1][1:]):
        assert_bitwise(a, b, "huge max_iter")
    it, _, r2, rho, beta = out[1]
    assert r2.size == it + 1 and np.all(np.isfinite(r2))
    assert np.isnan(rho[it]) and np.all(np.isfinite(rho[:it]))
    assert np.isnan(beta[it - 1]) and np.all(np.isfinite(beta[: it - 1]))
    # the mixed solver with the same limit
    ctx.prepare_fp32()
    ctx.fill("X", 0.0)
    m = ctx.solve_mixed(precond="jacobi", max_iter=2**31 - 1, tol=1e-9)
    assert m["true_residual"] <= 1e-9
    assert ctx.history("r2", m["iterations"] + 1).size == m["iterations"] + 1


def test_fused_profile_and_batches_agree(ctx, tmp_path):
    fx = [f for f in load_golden() if f["case"] == "lap2d_64"][0]
    case = Case(fx, str(tmp_path))
    ref = None
    for batch, profile in [(0, 0), (1, 0), (7, 0), (0, 1)]:
        _fresh(ctx, case, "CRS", "jacobi")
        info = ctx.solve(precond="jacobi", batch=batch, profile=profile)
        x = ctx.download("X")
        if ref is None:
            ref = (info["iterations"], x)
        assert info["iterations"] == ref[0]
        assert_bitwise(x, ref[1], f"x batch={batch} profile={profile}")
        if profile:
            assert info["n_matvec"] == info["iterations"]
            assert info["n_update"] == info["iterations"]
            assert info["n_xpay"] == info["iterations"] - 1
            assert info["ms_matvec"] > 0 and info["ms_update"] > 0 and info["ms_xpay"] > 0
            assert info["n_init"] == 1 and info["ms_init"] > 0


def test_device_sqrt_and_division_are_correctly_rounded(ctx, tmp_path):
    """residual = sqrt(r2)/sqrt(r2_0) and alpha = rho/pq as the host computes them."""
    fx = [f for f in load_golden() if f["case"] == "diff3d_24"][0]
    case = Case(fx, str(tmp_path))
    _fresh(ctx, case, "CRS", "jacobi")
    info = ctx.solve(precond="jacobi", max_iter=5000)
    it = info["iterations"]
    r2 = ctx.history("r2", it + 1)
    rho = ctx.history("rho", it + 1)
    pq = ctx.history("pq", it)
    alpha = ctx.history("alpha", it)
    assert info["residual"] == float(np.sqrt(r2[-1]) / np.sqrt(r2[0]))
    assert_bitwise(alpha, rho[:it] / pq, "alpha")


@pytest.mark.parametrize("fmt", ["CRS", "ELL"])
def test_time_hook(ctx, fmt, tmp_path):
    """cghip_time_hook runs every hook's device work and returns a positive
    average; the hooks it times leave finite values and the same results as
    the hook calls themselves."""
    case = Case(load_golden()[0], str(tmp_path))
    _fresh(ctx, case, fmt, "jacobi")
    n = case.n
    x = np.linspace(0.5, 1.5, n)
    for v in ("P", "R", "Q", "Z"):
        ctx.upload(v, x)
    for hook in ("matvec", "axpy", "xpay", "dot", "jacobi", "cpy"):
        ms = ctx.time_hook(hook, reps=3)
        assert ms > 0, hook
    ctx.upload("R", x)
    ctx.time_hook("jacobi", reps=1)
    assert np.array_equal(ctx.download("Z"), case.C * x)
    ctx.time_hook("matvec", reps=1)
    assert np.all(np.isfinite(ctx.download("Q")))
