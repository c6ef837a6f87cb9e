"""The labelling behind the trained network, against dense linear algebra (CPU).

Each harvested system is labelled from the alpha/beta of one FP64
Jacobi-PCG run: spectral_measure turns them into the Ritz nodes and Gauss
weights of the Jacobi-scaled operator, simulate_iterations runs PCG with a
polynomial preconditioner on that measure, best_lower_end / label_degrees
search the Chebyshev lower end.  Here the same quantities come from a dense
eigendecomposition of D^-1/2 A D^-1/2 and from dense PCG runs with the
preconditioner D^-1/2 p(D^-1/2 A D^-1/2) D^-1/2 written out as a matrix,
stopping on ||D^-1/2 r|| <= tol ||D^-1/2 r_0||.  A wrong tridiagonal index,
weights from the wrong eigenvector row, or a residual norm without the
1/p weighting moves an iteration count or a node, and fails.
"""
import numpy as np
import pytest
import scipy.sparse

from cgxx_amd import learned, problems, spectrum


def _dense(m):
    return scipy.sparse.csr_matrix((m.value, m.index, m.ptr), shape=(m.n, m.n)).toarray()


def _exact_measure(A, k):
    """Eigenvalues of D^-1/2 A D^-1/2 and the squared components of the
    normalised scaled right-hand side D^-1/2 k in its eigenbasis."""
    dh = 1.0 / np.sqrt(np.diag(A))
    lam, V = np.linalg.eigh(dh[:, None] * A * dh[None, :])
    c = V.T @ (dh * k)
    return lam, V, c * c / (c @ c), dh


def _pcg(A, k, minv, dh, steps=None, tol=0.0, max_iter=5000):
    """Textbook PCG from x = 0 with the dense preconditioner minv; alphas,
    betas and ||D^-1/2 r_j|| / ||D^-1/2 r_0|| (j = 0..)."""
    r = k.copy()
    z = minv @ r
    p = z.copy()
    rz = r @ z
    h0 = np.linalg.norm(dh * r)
    al, be, hist = [], [], [1.0]
    for _ in range(steps or max_iter):
        q = A @ p
        a = rz / (p @ q)
        r = r - a * q
        al.append(a)
        hist.append(np.linalg.norm(dh * r) / h0)
        z = minv @ r
        rzn = r @ z
        be.append(rzn / rz)
        p = z + be[-1] * p
        rz = rzn
        if steps is None and hist[-1] <= tol:
            break
    return np.array(al), np.array(be), np.array(hist)


def _poly_minv(V, dh, pvals):
    return dh[:, None] * ((V * pvals) @ V.T) * dh[None, :]


def _agrees(count, hist, tol):
    """count is the first j >= 1 with hist[j] <= tol, or off by one where a
    residual sits within 1e-6 (relative) of the threshold."""
    below = np.nonzero(hist[1:] <= tol)[0]
    assert below.size, "dense PCG did not reach the tolerance"
    e = int(below[0]) + 1
    if count == e:
        return True
    near = any(abs(np.log(hist[i] / tol)) < 1e-6 for i in (e - 1, e) if i >= 1)
    return near and abs(count - e) == 1


@pytest.fixture(scope="module")
def system():
    m = problems.anisotropic_family(100003, sides=(6,))
    A = _dense(m)
    k = m.ones_rhs()
    lam, V, w, dh = _exact_measure(A, k)
    b = spectrum.gershgorin_bound(m.ptr, m.index, m.value)
    minv_j = np.diag(dh * dh)
    a10, b10, _ = _pcg(A, k, minv_j, dh, steps=learned.RITZ_STEPS)
    theta10 = float(spectrum.ritz_values(a10, b10[:-1])[0])
    return {"A": A, "k": k, "lam": lam, "V": V, "w": w, "dh": dh, "b": b, "theta10": theta10}


def test_spectral_measure_of_a_complete_run_is_the_exact_measure():
    """n PCG steps on an n x n system: the Ritz nodes are the eigenvalues of
    D^-1/2 A D^-1/2 and the Gauss weights the squared components of the
    normalised scaled right-hand side."""
    rng = np.random.default_rng(7)
    n = 8
    B = rng.normal(size=(n, n))
    A = B @ B.T + n * np.diag(rng.uniform(0.5, 4.0, n))
    k = rng.normal(size=n)
    lam, _, w, dh = _exact_measure(A, k)
    al, be, _ = _pcg(A, k, np.diag(dh * dh), dh, steps=n)
    theta, wt = learned.spectral_measure(al, be)
    np.testing.assert_allclose(theta, lam, rtol=1e-9)
    np.testing.assert_allclose(wt, w, atol=1e-9)
    assert abs(wt.sum() - 1.0) < 1e-14


@pytest.mark.parametrize("poly", ["jacobi", "cheb4", "cheb8", "cheb6_high"])
def test_simulated_iterations_match_dense_pcg(system, poly):
    s = system
    lam, b = s["lam"], s["b"]
    if poly == "jacobi":
        pv = np.ones_like(lam)
    else:
        deg = int(poly[4:].split("_")[0])
        a = 3.0 * lam[0] if poly.endswith("high") else 0.5 * lam[0]
        pv = spectrum.poly_eval(spectrum.cheb_residual_coeffs(a, b, deg), b, lam)
    assert np.all(pv > 0)
    if poly != "jacobi":
        assert pv.max() / pv.min() > 2.0  # the 1/p weighting matters
    _, _, hist = _pcg(s["A"], s["k"], _poly_minv(s["V"], s["dh"], pv), s["dh"], tol=1e-12)
    for tol in np.logspace(-1, -11, 41):  # every crossing of the residual curve
        count = learned.simulate_iterations(lam, s["w"], pv, tol=tol)
        assert _agrees(count, hist, tol), (poly, tol, count)


@pytest.mark.parametrize("degree", [4, 8])
def test_best_lower_end_brute_force(system, degree):
    """Every ratio of the grid: -1 exactly where the Chebyshev polynomial is
    not positive on the nodes, else the dense PCG count; the label is the
    smallest count, ties to the larger ratio."""
    s = system
    lam, b, theta10 = s["lam"], s["b"], s["theta10"]
    r, it, its = learned.best_lower_end(lam, s["w"], theta10, b, degree)
    assert its.shape == learned.RATIO_GRID.shape
    for i, ratio in enumerate(learned.RATIO_GRID):
        c = spectrum.cheb_residual_coeffs(min(theta10 * ratio, 0.25 * b), b, degree)
        pv = spectrum.poly_eval(c, b, lam)
        if not np.all(pv > 0):
            assert its[i] == -1, ratio
            continue
        _, _, hist = _pcg(s["A"], s["k"], _poly_minv(s["V"], s["dh"], pv), s["dh"], tol=1e-9)
        assert _agrees(int(its[i]), hist, 1e-9), (ratio, its[i])
    ok = its >= 0
    assert it == its[ok].min()
    first = int(np.nonzero(its == it)[0][0])
    assert r == learned.RATIO_GRID[first] and learned.RATIO_GRID[first] == learned.RATIO_GRID[:first + 1].min()


def test_label_degrees_from_a_converged_run(system):
    """Labels from the alpha/beta of one Jacobi-PCG run to 1e-12 (as the
    harvest records them) against labels from the exact measure: the same
    Jacobi count, extreme Ritz values at the extreme eigenvalues (the top one
    carries a small weight of the all-ones right-hand side and converges
    last), and per degree iteration counts within one of the exact ones
    wherever count * (degree + 1) <= 2 m, m the Lanczos steps: the m-node
    Gauss measure reproduces the moments a PCG run of that many operator
    applications uses.  Past that (ratios far below the best) the measure
    undercounts; the labelled best count lies inside."""
    s = system
    al, be, hist = _pcg(s["A"], s["k"], np.diag(s["dh"] ** 2), s["dh"], tol=1e-12)
    m = al.size
    lab = learned.label_degrees(al, be, s["theta10"], s["b"], degrees=(4, 8))
    assert _agrees(lab["jacobi_iters"], hist, 1e-9)
    assert abs(lab["theta_min"] / s["lam"][0] - 1) < 1e-6
    assert abs(lab["theta_max"] / s["lam"][-1] - 1) < 1e-4
    for j, deg in enumerate((4, 8)):
        _, it_exact, its_exact = learned.best_lower_end(s["lam"], s["w"], s["theta10"], s["b"], deg)
        got = np.asarray(lab["grid_iters"][j])
        assert np.array_equal(got >= 0, its_exact >= 0)
        inside = (its_exact >= 0) & (its_exact * (deg + 1) <= 2 * m)
        assert inside.sum() >= 10
        assert np.abs(got - its_exact)[inside].max() <= 1, (deg, got, its_exact)
        assert it_exact * (deg + 1) <= 2 * m and abs(lab["iters"][j] - it_exact) <= 1
