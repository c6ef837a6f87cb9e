"""Polynomial preconditioner and spectral estimates (host logic + oracle).

Pins (none of these has a reference counterpart):
  - Ritz values from PCG alpha/beta against the closed-form spectrum of the
    1D Dirichlet Laplacian, lambda_k = 2 - 2 cos(k pi / (n + 1)), Jacobi-
    scaled by its constant diagonal 2; inside [lambda_min, lambda_max] and
    monotone in the number of steps (Cauchy interlacing).
  - Chebyshev residual polynomial: max |1 - l p(l)| on [a, b] equals the
    closed form 1 / T_{k+1}((b + a) / (b - a)); p > 0 on [0, b].
  - The oracle's FP32 Clenshaw application against the dense FP64 operator
    p(A_hat), A_hat = D^-1/2 A D^-1/2 (the preconditioner p(D^-1 A) D^-1 in
    scaled form), built by the three-term recurrence
    T_{j+1} = 2 Y T_j - T_{j-1} (brute force), and its symmetry
    <B u, v> = <u, B v>.
  - The oracle's mixed solver with the polynomial reaches the FP64 true
    residual and x = 1.
  - FP16 rounding of the DIA operator's values (the Clenshaw steps read
    them): against numpy's float16 conversion (round to nearest even, with
    subnormals and overflow), and the Clenshaw application on a DIA matrix
    against the dense FP64 operator built from the FP16-rounded A_hat.
"""
import numpy as np
import pytest

from cgxx_amd import problems, spectrum
from cgxx_amd.problems import CRS
from oracle import oracle


def laplace1d(n: int) -> CRS:
    ptr, idx, val = [0], [], []
    for i in range(n):
        for j, v in ((i - 1, -1.0), (i, 2.0), (i + 1, -1.0)):
            if 0 <= j < n:
                idx.append(j)
                val.append(v)
        ptr.append(len(idx))
    return CRS(n, np.array(ptr, np.int32), np.array(idx, np.int32), np.array(val))


def dense(m: CRS) -> np.ndarray:
    A = np.zeros((m.n, m.n))
    for i in range(m.n):
        A[i, m.index[m.ptr[i]:m.ptr[i + 1]]] = m.value[m.ptr[i]:m.ptr[i + 1]]
    return A


def dense_poly_operator(A, coeffs, b):
    """p(A_hat), A_hat = D^-1/2 A D^-1/2, by the Chebyshev three-term
    recurrence T_{j+1} = 2 Y T_j - T_{j-1} with Y = 2 A_hat / b - I, FP64.
    (p(A_hat) = D^1/2 p(D^-1 A) D^-1 D^1/2: the preconditioner in scaled form.)"""
    n = A.shape[0]
    s = 1.0 / np.sqrt(np.diag(A))
    Ah = s[:, None] * A * s[None, :]
    Y = 2.0 * Ah / b - np.eye(n)
    T0, T1 = np.eye(n), Y
    P = coeffs[0] * T0 + coeffs[1] * T1
    for j in range(2, len(coeffs)):
        T0, T1 = T1, 2.0 * Y @ T1 - T0
        P = P + coeffs[j] * T1
    return P


def test_ritz_closed_form_laplace1d():
    n = 40
    m = laplace1d(n)
    C = np.full(n, 0.5)
    lam = np.sort(1.0 - np.cos(np.arange(1, n + 1) * np.pi / (n + 1)))  # D^-1 A, D = 2
    k = np.random.default_rng(7).uniform(-1, 1, n)
    _, it, _, h = oracle.pcg(n=n, fmt="CRS", precond="jacobi", order=0, k=k, tol=1e-300,
                             max_iter=n, crs=(m.ptr, m.index, m.value), C_jac=C)
    lo_prev, hi_prev = np.inf, -np.inf
    for steps in range(1, 16):
        th = spectrum.ritz_values(h["alpha"][:steps], h["beta"][:steps - 1])
        assert th[0] >= lam[0] * (1 - 1e-10) and th[-1] <= lam[-1] * (1 + 1e-10)
        assert th[0] <= lo_prev * (1 + 1e-12) and th[-1] >= hi_prev * (1 - 1e-12)
        lo_prev, hi_prev = th[0], th[-1]
    # after n steps in exact arithmetic the Ritz values are the eigenvalues;
    # the extreme ones survive the loss of orthogonality
    th = spectrum.ritz_values(h["alpha"][:n], h["beta"][:n - 1])
    assert abs(th[-1] - lam[-1]) <= 1e-9 and abs(th[0] - lam[0]) <= 1e-9


def test_ritz_rejects_bad_input():
    with pytest.raises(ValueError):
        spectrum.ritz_values([], [])
    with pytest.raises(ValueError):
        spectrum.ritz_values([1.0, 2.0, 3.0], [0.5])


@pytest.mark.parametrize("degree", [1, 2, 3, 8, 16])
def test_chebyshev_residual_closed_form(degree):
    a, b = 0.003, 2.0
    c = spectrum.cheb_residual_coeffs(a, b, degree)
    assert c.size == degree + 1
    lam = np.linspace(a, b, 20001)
    res = 1.0 - lam * spectrum.poly_eval(c, b, lam)
    bound = 1.0 / np.cosh((degree + 1) * np.arccosh((b + a) / (b - a)))
    assert abs(np.abs(res).max() - bound) <= 1e-6 * max(bound, 1e-3)
    assert spectrum.poly_eval(c, b, np.linspace(0, b, 4096)).min() > 0


def test_gershgorin_bound():
    m = problems.laplace2d(8)
    assert spectrum.gershgorin_bound(m.ptr, m.index, m.value) == 2.0
    m = problems.random_spd_dense(30, 0.5, 3)
    lam = np.linalg.eigvals(np.diag(1 / m.diagonal()) @ dense(m)).real
    assert lam.max() <= spectrum.gershgorin_bound(m.ptr, m.index, m.value) * (1 + 1e-12)


@pytest.mark.parametrize("degree", [1, 2, 3, 4, 8, 16])
@pytest.mark.parametrize("name", ["lap2d", "spd"])
def test_oracle_poly_apply_brute_force(name, degree):
    m = problems.laplace2d(7) if name == "lap2d" else problems.random_spd_dense(40, 0.3, 11)
    b = spectrum.gershgorin_bound(m.ptr, m.index, m.value)
    c = spectrum.cheb_residual_coeffs(b / 50, b, degree)
    c32 = c.astype(np.float32).astype(np.float64)
    b32 = 2.0 / float(np.float32(2.0 / b))
    B = dense_poly_operator(dense(m), c32, b32)
    crs = (m.ptr, m.index, m.value)
    rng = np.random.default_rng(degree)
    for _ in range(3):
        r = rng.standard_normal(m.n).astype(np.float32)
        z = oracle.poly_apply(crs, r, c, b)
        zr = B @ r.astype(np.float64)
        # FP32 Clenshaw: error grows with the degree and with sum |c_j|
        tol = 2e-6 * (degree + 1) * np.abs(c32).sum() * np.abs(r).max()
        assert np.abs(z - zr).max() <= tol
    # symmetry of the operator (FP32 application)
    u = rng.standard_normal(m.n).astype(np.float32)
    v = rng.standard_normal(m.n).astype(np.float32)
    bu = oracle.poly_apply(crs, u, c, b).astype(np.float64)
    bv = oracle.poly_apply(crs, v, c, b).astype(np.float64)
    assert abs(bu @ v - u @ bv) <= 1e-5 * np.abs(B).max() * m.n
    np.testing.assert_allclose(B, B.T, rtol=0, atol=1e-10 * np.abs(B).max())


def test_oracle_poly_constant_is_identity_step():
    """degree 1 with c = (1, 0): z = r exactly (the Jacobi step in scaled form)."""
    m = problems.laplace2d(6)
    r = np.random.default_rng(0).standard_normal(m.n).astype(np.float32)
    z = oracle.poly_apply((m.ptr, m.index, m.value), r, [1.0, 0.0], 2.0)
    np.testing.assert_array_equal(z, r)


def test_oracle_scaled_operator_rejects_missing_diagonal():
    ptr = np.array([0, 2, 3], np.int32)
    idx = np.array([0, 1, 0], np.int32)
    val = np.array([2.0, -1.0, -1.0])
    with pytest.raises(ValueError, match="-3"):
        oracle.mixed_pcg(n=2, crs=(ptr, idx, val), k=np.ones(2), precond="jacobi")


@pytest.mark.parametrize("degree", [2, 8])
def test_oracle_mixed_poly_converges(degree):
    m = problems.diffusion3d(12, 1.0, 3.0, 5)
    k = m.ones_rhs()
    C = 1.0 / m.diagonal()
    b = spectrum.gershgorin_bound(m.ptr, m.index, m.value)
    c = spectrum.cheb_residual_coeffs(b / 200, b, degree)
    crs = (m.ptr, m.index, m.value)
    x, info, h = oracle.mixed_pcg(n=m.n, crs=crs, k=k, precond="poly", C_jac=C, poly=(c, b))
    xj, infoj, _ = oracle.mixed_pcg(n=m.n, crs=crs, k=k, precond="jacobi", C_jac=C)
    r = k - m.matvec(x)
    assert np.linalg.norm(r) / np.linalg.norm(k) <= 1.05e-9
    assert info["true_residual"] <= 1e-9
    assert np.abs(x - 1).max() <= 1e-5
    assert info["iterations"] < infoj["iterations"]


def test_oracle_round_half_matches_numpy():
    import ctypes as C
    L = oracle.lib()
    L.orc_round_half.restype = C.c_float
    L.orc_round_half.argtypes = [C.c_float]
    rng = np.random.default_rng(5)
    vals = np.concatenate([
        rng.standard_normal(4000), rng.uniform(-1, 1, 4000) * 1e-5,  # normal and subnormal halves
        rng.uniform(-1, 1, 1000) * 1e-8, rng.uniform(-7e4, 7e4, 1000),  # underflow, overflow
        # exact ties between neighbouring halves (round to even) and their neighbours
        (np.arange(1, 2049) + 0.5) / 2048.0, (np.arange(1, 200) + 0.5) * 2.0 ** -24,
        [0.0, -0.0, 65504.0, 65519.99, 65520.0, 2.0 ** -14, 2.0 ** -24, 2.0 ** -25, 3 * 2.0 ** -26],
    ]).astype(np.float32)
    got = np.array([L.orc_round_half(float(v)) for v in vals], dtype=np.float32)
    want = vals.astype(np.float16).astype(np.float32)
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))


def _fp16_poly_operator(m, coeffs, b):
    """p(A16) with A16 the library's DIA values: FP32 (a_ij s_i) s_j rounded
    to FP16, unit diagonal; FP64 three-term recurrence (brute force)."""
    n = m.n
    s = 1.0 / np.sqrt(m.diagonal())
    A = np.eye(n)
    for i in range(n):
        for e in range(m.ptr[i], m.ptr[i + 1]):
            j = m.index[e]
            if j != i:
                v32 = np.float32((m.value[e] * s[i]) * s[j])
                A[i, j] = float(np.float16(v32))
    Y = 2.0 * A / b - np.eye(n)
    T0, T1 = np.eye(n), Y
    P = coeffs[0] * T0 + coeffs[1] * T1
    for j in range(2, len(coeffs)):
        T0, T1 = T1, 2.0 * Y @ T1 - T0
        P = P + coeffs[j] * T1
    return P


@pytest.mark.parametrize("degree", [2, 8])
def test_oracle_poly_apply_fp16_dia(degree):
    """A 3D variable-coefficient diffusion lies on 6 diagonals: the oracle
    applies the polynomial with the FP16-rounded values, within FP32 rounding
    of the dense FP16 operator and far from the FP32 one."""
    m = problems.diffusion3d(6, sigma=1.0, corr=2.0, seed=3)
    b = spectrum.gershgorin_bound(m.ptr, m.index, m.value)
    c = spectrum.cheb_residual_coeffs(b / 50, b, degree)
    c32 = c.astype(np.float32).astype(np.float64)
    b32 = 2.0 / float(np.float32(2.0 / b))
    B16 = _fp16_poly_operator(m, c32, b32)
    B32 = dense_poly_operator(dense(m), c32, b32)
    r = np.random.default_rng(degree).standard_normal(m.n).astype(np.float32)
    z = oracle.poly_apply((m.ptr, m.index, m.value), r, c, b).astype(np.float64)
    tol = 2e-6 * (degree + 1) * np.abs(c32).sum() * np.abs(r).max()
    e16 = np.abs(z - B16 @ r).max()
    assert e16 <= tol
    assert np.abs(z - B32 @ r).max() > 10 * e16
