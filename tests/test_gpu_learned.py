"""The learned part on the GPU: network kernels against the CPU oracle,
matrix statistics against the brute-force oracle, and the argument checks."""
import numpy as np
import pytest

from cgxx_amd import learned, problems
from oracle import features_oracle, mlp_oracle as mo

pytestmark = pytest.mark.gpu


def _mlp(dims, rows):
    from cgxx_amd.cghip import Mlp

    return Mlp(dims, rows)


def _data(dims, rows, seed):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((rows, dims[0])).astype(np.float32)
    T = rng.standard_normal((rows, dims[3])).astype(np.float32)
    return X, T


@pytest.mark.parametrize("dims,rows", [((42, 128, 128, 9), 300), ((5, 7, 3, 2), 9), ((1, 1, 1, 1), 1),
                                       ((256, 256, 256, 64), 130)])
def test_forward_and_gradient_vs_oracle(dims, rows):
    m = _mlp(dims, rows)
    p = learned.init_params(dims, 1)
    p += (0.05 * np.random.default_rng(2).standard_normal(p.size)).astype(np.float32)
    m.set_params(p)
    np.testing.assert_array_equal(m.get_params(), p)
    X, T = _data(dims, rows, 3)
    Y = m.forward(X)
    _, _, Yr = mo.forward(p, dims, X)
    np.testing.assert_allclose(Y, Yr, rtol=2e-5, atol=2e-5 * np.abs(Yr).max())
    g, L = m.gradient(X, T)
    gr, Lr = mo.gradient(p, dims, X, T)
    assert L == pytest.approx(Lr, rel=1e-5)
    np.testing.assert_allclose(g, gr, rtol=1e-3, atol=1e-4 * np.abs(gr).max())
    m.close()


def test_adam_step_vs_oracle_and_determinism():
    dims, rows = (42, 128, 128, 9), 257
    X, T = _data(dims, rows, 4)
    p0 = learned.init_params(dims, 5)
    m = _mlp(dims, rows)
    m.set_params(p0)
    g, _ = m.gradient(X, T)
    hist = m.train(X, T, 1, lr=1e-3)
    p1 = m.get_params()
    pr, _, _ = mo.adam_step(p0, np.zeros_like(p0), np.zeros_like(p0), g, 1, lr=1e-3)
    np.testing.assert_allclose(p1, pr, rtol=1e-6, atol=1e-9)
    # several steps: Adam state carried across calls equals one longer call
    m.set_params(p0)
    m.train(X, T, 3)
    m.train(X, T, 4)
    pa = m.get_params()
    m.set_params(p0)
    h7 = m.train(X, T, 7)
    np.testing.assert_array_equal(m.get_params(), pa)
    assert h7[0] == hist[0]
    m.close()


def test_training_fits_a_smooth_target():
    dims, rows = (4, 32, 32, 2), 512
    rng = np.random.default_rng(6)
    X = rng.uniform(-1, 1, (rows, 4)).astype(np.float32)
    T = np.stack([np.sin(X[:, 0]) + X[:, 1] * X[:, 2], X[:, 3] ** 2], 1).astype(np.float32)
    m = _mlp(dims, rows)
    m.set_params(learned.init_params(dims, 7))
    hist = m.train(X, T, 2000, lr=3e-3)
    assert hist[-1] < hist[0] / 50
    assert mo.loss(m.get_params(), dims, X, T) == pytest.approx(hist[-1], rel=0.2)
    m.close()


def test_mlp_rejects_bad_arguments():
    from cgxx_amd.cghip import CGHIPError, Mlp

    with pytest.raises(CGHIPError):
        Mlp((0, 4, 4, 2), 4)
    with pytest.raises(CGHIPError):
        Mlp((4, 4, 257, 2), 4)
    m = Mlp((3, 4, 4, 2), 4)
    with pytest.raises(CGHIPError):
        m.forward(np.zeros((5, 3), np.float32))  # rows > max_rows
    bad = np.zeros(m.total, np.float32)
    bad[3] = np.nan
    with pytest.raises(CGHIPError):
        m.set_params(bad)
    X, T = _data((3, 4, 4, 2), 4, 0)
    with pytest.raises(CGHIPError):
        m.train(X, T, 1, lr=-1.0)
    m.close()


@pytest.fixture(scope="module")
def ctx():
    from cgxx_amd.cghip import Context

    c = Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("name", ["lap2d", "diff3d", "aniso", "spd", "zipf"])
def test_matrix_stats_vs_oracle(ctx, name):
    m = {
        "lap2d": lambda: problems.laplace2d(33),
        "diff3d": lambda: problems.diffusion3d(20, 1.0, 4.0, 3),
        "aniso": lambda: problems.anisotropic_family(11, sides=(24,)),
        "spd": lambda: problems.random_spd_dense(70, 0.4, 2),
        "zipf": lambda: problems.zipf_spd(20000, seed=5),
    }[name]()
    ctx.set_size(m.n)
    ctx.upload_crs(m.ptr, m.index, m.value)
    s = ctx.matrix_stats()
    r = features_oracle.matrix_stats(m.ptr, m.index, m.value)
    std = np.zeros(s.size, bool)
    std[1:20:4] = True
    # the kernel's one-pass variance (sum q^2 / n - mean^2) loses about
    # 1e-16 * mean^2, i.e. ~1e-8 * |mean| in the standard deviation
    np.testing.assert_allclose(s[~std], r[~std], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(s[std], r[std], rtol=1e-6, atol=1e-7 * (1 + np.abs(r[0:20:4]).max()))
    assert s[20:].sum() == pytest.approx(1.0)


def test_matrix_stats_rejects_missing_diagonal(ctx):
    from cgxx_amd.cghip import CGHIPError

    ptr = np.array([0, 2, 3], np.int32)
    idx = np.array([0, 1, 0], np.int32)  # row 1 has no diagonal entry
    val = np.array([2.0, -1.0, -1.0])
    ctx.set_size(2)
    ctx.upload_crs(ptr, idx, val)
    with pytest.raises(CGHIPError, match="diagonal"):
        ctx.matrix_stats()


@pytest.mark.parametrize("layout_case", ["dia", "ell"])
def test_iteration_bytes_rank_like_device_time(ctx, layout_case):
    """The byte model that chooses between the polynomial and Jacobi
    (cghip_mixed_iteration_bytes) against the device: the ratio of the
    polynomial's to Jacobi's bytes per iteration is within 30% of the ratio
    of their kernel-timed iterations on a 7M-row system (DIA) and on a
    randomly renumbered 2M-row grid stored in ELL.  (Smaller systems are
    closer to the kernels' launch latency, which the byte model leaves out
    and which favours the polynomial's iteration of many short kernels.)"""
    from cgxx_amd import spectrum

    if layout_case == "dia":
        m = problems.diffusion3d(192, 1.0, 8.0, 5)
    else:  # the same grid renumbered at random: 7 entries per row on many diagonals
        import scipy.sparse as sp

        d = problems.diffusion3d(128, 1.0, 8.0, 5)
        perm = np.random.default_rng(1).permutation(d.n)
        B = sp.csr_matrix((d.value, d.index, d.ptr), shape=(d.n, d.n))[perm][:, perm]
        B.sort_indices()
        m = problems.CRS(d.n, B.indptr.astype(np.int32), B.indices.astype(np.int32),
                         B.data.astype(np.float64))
    ctx.set_size(m.n)
    ctx.upload_crs(m.ptr, m.index, m.value)
    ctx.set_format("CRS")
    ctx.prepare_fp32()
    ctx.upload("K", m.ones_rhs())
    assert ctx.fp32_layout() == layout_case.upper()
    b = spectrum.gershgorin_bound(m.ptr, m.index, m.value)
    ctx.set_poly(spectrum.cheb_residual_coeffs(b / 100.0, b, 8), b)
    bj, bp = ctx.mixed_iteration_bytes("jacobi"), ctx.mixed_iteration_bytes("poly")
    assert bp > bj > 0
    ms = {}
    for pre, its in (("jacobi", 20), ("poly", 5)):
        ctx.fill("X", 0.0)
        info = ctx.solve_mixed(precond=pre, max_iter=its, tol=1e-300, profile=True)
        ms[pre] = (info["ms_matvec"] + info["ms_update"] + info["ms_poly"] + info["ms_xpay"]) / info["n_matvec"]
    ratio_t, ratio_b = ms["poly"] / ms["jacobi"], bp / bj
    assert abs(ratio_t / ratio_b - 1.0) <= 0.3, (ratio_t, ratio_b)
