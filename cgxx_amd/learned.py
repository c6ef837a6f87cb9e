"""Learned polynomial preconditioner: features, coefficient encoding, labels.

A small network (include/cghip_mlp.h) maps cheap features of a system to a
degree-8 polynomial p for the preconditioner z = p(D^-1 A) D^-1 r.

Features (42):
  0..35   cghip_matrix_stats: mean/std/min/max of log a_ii, the Gershgorin
          radius sum_{j!=i}|a_ij|/a_ii, log strongest and weakest coupling
          |a_ij|/a_ii and their ratio; the row-length histogram
  36      log N
  37..41  from the first 10 steps of FP64 Jacobi-PCG (the reference's loop,
          CG.cpp:284-385): log of the smallest and largest Ritz values after
          5 and 10 steps, and log10(r2_10 / r2_0) / 10
Outputs (11): log p at the 9 Chebyshev points of [0, b], b = 1 + max
Gershgorin radius (an upper bound of the spectrum of D^-1 A), then the log of
the iteration counts predicted for Jacobi-PCG and for the polynomial PCG.
The coefficients follow by the discrete Chebyshev transform, so p > 0 at
those points by construction; the library checks p > 0 on all of [0, b]
before use.  The iteration predictions, with the per-iteration costs timed
on the device, decide whether the polynomial is used at all.

Labels come from the spectrum the FP64 solver itself reveals: the Lanczos
tridiagonal of a converged Jacobi-PCG run (its alpha/beta) gives Ritz nodes
theta_i and Gauss weights w_i of the initial residual's spectral measure
(Golub & Meurant, Matrices, Moments and Quadrature, 2010, ch. 6).  The
polynomial-preconditioned PCG acts on the same eigenvectors with eigenvalues
theta_i p(theta_i) and weights p(theta_i) w_i, so its convergence is
simulated by CG on that diagonal system.  The label is the Chebyshev
polynomial (on [a, b]) whose simulated iteration count is smallest over a
grid of lower ends a.
"""
from __future__ import annotations

import os
import time

import numpy as np

from . import spectrum

DEGREE = 8
N_STATS = 36
N_FEATURES = N_STATS + 1 + 5
# outputs: log p at the DEGREE + 1 Chebyshev points, then the log of the
# predicted iteration counts of Jacobi-PCG and of the polynomial PCG
DIMS = (N_FEATURES, 128, 128, DEGREE + 3)
RITZ_STEPS = 10
MODEL_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "poly_mlp.npz")
# lower ends a = theta_1 * 2^(s/2), s = 0..A_STEPS-1, theta_1 = smallest Ritz value
A_STEPS = 17


def init_params(dims, seed: int) -> np.ndarray:
    """He-uniform weights, zero biases, float32, in the library's layout."""
    rng = np.random.default_rng(seed)
    parts = []
    i, h1, h2, o = dims
    for fan_in, fan_out in ((i, h1), (h1, h2), (h2, o)):
        lim = np.sqrt(6.0 / fan_in)
        parts.append(rng.uniform(-lim, lim, (fan_in, fan_out)).ravel())
        parts.append(np.zeros(fan_out))
    return np.concatenate(parts).astype(np.float32)


# ---- coefficients <-> log p at Chebyshev points --------------------------

def cheb_nodes(degree: int, b: float) -> np.ndarray:
    y = np.cos(np.pi * (np.arange(degree + 1) + 0.5) / (degree + 1))
    return (y + 1.0) * 0.5 * b


def decode_coeffs(y, b: float) -> np.ndarray:
    """Chebyshev coefficients on [0, b] of the degree-k polynomial with
    p(node_j) = exp(y_j) (discrete Chebyshev transform)."""
    y = np.asarray(y, dtype=np.float64)
    k = y.size - 1
    pv = np.exp(y)
    ang = np.pi * (np.arange(k + 1) + 0.5) / (k + 1)
    c = np.array([np.sum(pv * np.cos(m * ang)) for m in range(k + 1)]) * (2.0 / (k + 1))
    c[0] *= 0.5
    return c


def encode_coeffs(c, b: float) -> np.ndarray:
    lam = cheb_nodes(len(c) - 1, b)
    pv = spectrum.poly_eval(c, b, lam)
    if np.any(pv <= 0):
        raise ValueError("polynomial not positive at the Chebyshev points")
    return np.log(pv)


# ---- features -------------------------------------------------------------

def cg_statistics(alpha, beta, r2) -> np.ndarray:
    """5 features from the first 10 Jacobi-PCG steps (alpha, beta, r2 history)."""
    a = np.asarray(alpha, np.float64)
    bt = np.asarray(beta, np.float64)
    if a.size < RITZ_STEPS or bt.size < RITZ_STEPS - 1 or len(r2) < RITZ_STEPS + 1:
        raise ValueError("need the history of 10 PCG steps")
    t5 = spectrum.ritz_values(a[:5], bt[:4])
    t10 = spectrum.ritz_values(a[:RITZ_STEPS], bt[:RITZ_STEPS - 1])
    rate = np.log10(max(r2[RITZ_STEPS], 1e-300) / r2[0]) / RITZ_STEPS
    return np.array([np.log(t10[0]), np.log(t10[-1]), np.log(t5[0]), np.log(t5[-1]), rate])


def features(stats, n: int, alpha, beta, r2) -> np.ndarray:
    stats = np.asarray(stats, np.float64)
    if stats.size != N_STATS:
        raise ValueError(f"expected {N_STATS} matrix statistics")
    return np.concatenate([stats, [np.log(n)], cg_statistics(alpha, beta, r2)])


def spectrum_bound(stats) -> float:
    """b = 1 + max Gershgorin radius of D^-1 A (stats[7] = max of q1)."""
    return 1.0 + float(stats[7])


def system_features(ctx, n: int):
    """(features, b, ritz10) of the system uploaded to a cghip Context (CRS,
    Jacobi, K set): matrix statistics on the device and 10 FP64 Jacobi-PCG
    steps from x = 0.  Overwrites X."""
    stats = ctx.matrix_stats()
    ctx.fill("X", 0.0)
    ctx.solve(precond="jacobi", max_iter=RITZ_STEPS, tol=1e-300)
    a = ctx.history("alpha", RITZ_STEPS)
    bt = ctx.history("beta", RITZ_STEPS)
    r2 = ctx.history("r2", RITZ_STEPS + 1)
    f = features(stats, n, a, bt, r2)
    return f, spectrum_bound(stats), np.exp(f[37])


# ---- labels from the harvested spectrum ----------------------------------

def spectral_measure(alpha, beta):
    """Ritz nodes and Gauss weights (sum 1) of the Lanczos tridiagonal."""
    from scipy.linalg import eigh_tridiagonal

    a = np.asarray(alpha, np.float64)
    bt = np.asarray(beta, np.float64)
    m = a.size
    d = 1.0 / a
    d[1:] += bt[:m - 1] / a[:m - 1]
    e = np.sqrt(bt[:m - 1]) / a[:m - 1]
    theta, V = eigh_tridiagonal(d, e)
    w = V[0] ** 2
    return theta, w / w.sum()


def simulate_iterations(theta, w, pvals, tol=1e-9, max_iter=20000) -> int:
    """PCG iterations to ||r_hat|| <= tol ||r_hat_0|| on the spectral model:
    CG on diag(theta * p) with residual components sqrt(w * p); r_hat =
    r / sqrt(p) is the Jacobi-scaled residual.  pvals = p(theta) (> 0)."""
    mu = theta * pvals
    r = np.sqrt(w * pvals)
    ip = 1.0 / pvals
    p = r.copy()
    rr = r @ r
    n0 = np.sqrt((r * r) @ ip)
    for it in range(1, max_iter + 1):
        q = mu * p
        al = rr / (p @ q)
        r = r - al * q
        if np.sqrt((r * r) @ ip) <= tol * n0:
            return it
        rn = r @ r
        p = r + (rn / rr) * p
        rr = rn
    return max_iter


def candidate_lower_ends(theta_min: float, b: float) -> np.ndarray:
    a = theta_min * 2.0 ** (0.5 * np.arange(A_STEPS))
    return a[a < 0.5 * b]


def label(alpha, beta, b: float, degree: int = DEGREE):
    """Best Chebyshev lower end on the harvested spectrum.

    Returns dict: a_best, y (log p at the Chebyshev points), the candidate
    ends and their simulated iteration counts, and the simulated Jacobi
    count (p = 1)."""
    theta, w = spectral_measure(alpha, beta)
    theta = np.clip(theta, 1e-300, None)
    cands = candidate_lower_ends(theta[0], b)
    its = []
    for a in cands:
        c = spectrum.cheb_residual_coeffs(a, b, degree)
        its.append(simulate_iterations(theta, w, spectrum.poly_eval(c, b, theta)))
    its = np.array(its)
    k = int(np.argmin(its))
    c = spectrum.cheb_residual_coeffs(cands[k], b, degree)
    return {
        "a_best": float(cands[k]),
        "y": encode_coeffs(c, b),
        "cands": cands,
        "iters": its,
        "jacobi_iters": simulate_iterations(theta, w, np.ones_like(theta)),
        "theta_min": float(theta[0]),
        "theta_max": float(theta[-1]),
    }


# ---- model file and inference ---------------------------------------------

def save_model(path, params, feat_mean, feat_std, out_mean, out_std, **meta):
    np.savez(path, params=np.asarray(params, np.float32), dims=np.array(DIMS),
             feat_mean=feat_mean, feat_std=feat_std, out_mean=out_mean, out_std=out_std,
             degree=DEGREE, **{k: np.asarray(v) for k, v in meta.items()})


def load_model(path=MODEL_PATH) -> dict:
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def normalize(f, model) -> np.ndarray:
    """Standardised features, first clamped to the range seen in training
    (a system outside the family, e.g. a larger grid, is mapped to its
    nearest training edge instead of being extrapolated)."""
    f = np.atleast_2d(np.asarray(f, np.float64))
    if "feat_lo" in model:
        f = np.clip(f, model["feat_lo"], model["feat_hi"])
    return ((f - model["feat_mean"]) / model["feat_std"]).astype(np.float32)


def predict(mlp, model, f, b: float) -> dict:
    """Network output for features f (on the GPU): the coefficients on
    [0, b] and, for a model with 11 outputs, the predicted Jacobi and
    polynomial iteration counts."""
    yn = mlp.forward(normalize(f, model))[0].astype(np.float64)
    y = yn * model["out_std"] + model["out_mean"]
    k = int(model["degree"])
    out = {"coeffs": decode_coeffs(y[: k + 1], b)}
    if y.size >= k + 3:
        out["it_jacobi"] = float(np.exp(y[k + 1]))
        out["it_poly"] = float(np.exp(y[k + 2]))
    return out


def predict_coeffs(mlp, model, f, b: float) -> np.ndarray:
    """Coefficients on [0, b] from features f, with the network on the GPU."""
    return predict(mlp, model, f, b)["coeffs"]


def iteration_costs(ctx, coeffs, b: float) -> tuple[float, float] | None:
    """Algorithmic bytes of one mixed-precision iteration with Jacobi and
    with the polynomial on the uploaded system (the library's byte model,
    cghip_mixed_iteration_bytes: its loop kernels stream at similar rates,
    so bytes rank the two by time).  None if the library refuses the
    polynomial."""
    from .cghip import CGHIPError

    try:
        ctx.set_poly(coeffs, b)
    except CGHIPError:
        return None
    return ctx.mixed_iteration_bytes("jacobi"), ctx.mixed_iteration_bytes("poly")


def fixed_rule_coeffs(theta10: float, b: float, rho: float) -> np.ndarray:
    """The method without the network: Chebyshev on [theta_min(10) / rho, b]."""
    return spectrum.cheb_residual_coeffs(min(theta10 / rho, 0.25 * b), b, DEGREE)


def solve_poly_with_fallback(ctx, coeffs, b: float, tol: float = 1e-9, max_iter: int = 100000,
                             budget: int | None = None) -> dict:
    """Mixed-precision solve of the uploaded system from X = 0 with the
    polynomial preconditioner, falling back to Jacobi.

    The polynomial run may use `budget` SpMVs (default: max_iter, the budget
    a Jacobi run would get).  If the library rejects the polynomial (not
    positive on [0, b]) the solve runs with Jacobi from the start; if the
    polynomial run ends above tol (FP64 true residual), Jacobi continues
    from its X until the true residual relative to the initial one is
    <= tol.  Returns the method used, iteration and SpMV counts, the final
    FP64 true relative residual and the wall time."""
    from .cghip import CGHIPError

    t0 = time.perf_counter()
    k = len(coeffs) - 1
    budget = max_iter if budget is None else int(budget)
    ctx.fill("X", 0.0)
    res = {"method": "poly", "poly_iterations": 0, "jacobi_iterations": 0, "updates": 0}
    try:
        ctx.set_poly(coeffs, b)
        ok = True
    except CGHIPError as e:
        ok = False
        res["method"] = "jacobi"
        res["rejected"] = str(e)
    rel = 1.0
    if ok:
        info = ctx.solve_mixed(precond="poly", max_iter=max(1, budget // (k + 1)), tol=tol)
        res["poly_iterations"] = info["iterations"]
        res["updates"] += info["updates"]
        rel = info["true_residual"]
    # `not rel <= tol` also catches a NaN residual; a non-finite one leaves
    # no usable X, so Jacobi then starts again from X = 0.
    if not rel <= tol:
        if ok:
            res["method"] = "poly+jacobi"
        if not np.isfinite(rel):
            ctx.fill("X", 0.0)
            rel = 1.0
        info = ctx.solve_mixed(precond="jacobi", max_iter=max_iter, tol=tol / rel)
        res["jacobi_iterations"] = info["iterations"]
        res["updates"] += info["updates"]
        rel = rel * info["true_residual"]
    res["true_residual"] = rel
    res["spmvs"] = res["poly_iterations"] * (k + 1) + res["jacobi_iterations"]
    res["wall_s"] = time.perf_counter() - t0
    return res


def solve_learned(ctx, mlp, model, n: int, tol: float = 1e-9, max_iter: int = 100000,
                  budget: int | None = None) -> dict:
    """Features on the device, the network on the GPU, then the solve.

    With iteration predictions (11-output model) the predicted costs decide:
    the polynomial runs only if (predicted polynomial iterations) x (its
    bytes per iteration) is below the same product for Jacobi, and then with
    an SpMV budget worth 1.25x Jacobi's predicted cost; past it, or if the
    library refuses the polynomial, Jacobi takes over
    (solve_poly_with_fallback).  Otherwise Jacobi runs from the start.
    The per-iteration costs are the library's algorithmic bytes per
    iteration (iteration_costs).  The wall time includes features and
    inference."""
    t0 = time.perf_counter()
    f, b, _ = system_features(ctx, n)
    pr = predict(mlp, model, f, b)
    c = pr["coeffs"]
    k = len(c) - 1
    decision = None
    if "it_jacobi" in pr and budget is None:
        costs = iteration_costs(ctx, c, b)
        if costs is not None:
            bj, bp = costs
            cost_j, cost_p = bj * pr["it_jacobi"], bp * pr["it_poly"]
            decision = {"it_jacobi_pred": pr["it_jacobi"], "it_poly_pred": pr["it_poly"],
                        "bytes_jacobi_iter": bj, "bytes_poly_iter": bp,
                        "use_poly": cost_p < cost_j}
            budget = int(np.ceil(1.25 * cost_j / bp)) * (k + 1)
    if decision is not None and not decision["use_poly"]:
        ctx.fill("X", 0.0)
        info = ctx.solve_mixed(precond="jacobi", max_iter=max_iter, tol=tol)
        res = {"method": "jacobi-selected", "poly_iterations": 0,
               "jacobi_iterations": info["iterations"], "updates": info["updates"],
               "true_residual": info["true_residual"], "spmvs": info["iterations"]}
    else:
        res = solve_poly_with_fallback(ctx, c, b, tol, max_iter, budget)
    res["wall_s"] = time.perf_counter() - t0
    res["coeffs"] = c
    res["b"] = b
    res["decision"] = decision
    return res


BIN_MAGIC = b"CGHIPMLP"


def export_bin(model: dict, path: str) -> None:
    """Writes the model for the C++ host (cgxx_amd/host/LearnedPoly.cpp), all
    little-endian: magic "CGHIPMLP", int32 version 1, int32 in, h1, h2, out,
    degree, float64 rho_star, float64 feat_mean[in], feat_std[in],
    feat_lo[in], feat_hi[in], out_mean[out], out_std[out], int32 nparams,
    float32 params[nparams]."""
    dims = [int(v) for v in model["dims"]]
    with open(path, "wb") as f:
        f.write(BIN_MAGIC)
        np.array([1, *dims, int(model["degree"])], "<i4").tofile(f)
        np.array([float(model["rho_star"])], "<f8").tofile(f)
        for key in ("feat_mean", "feat_std", "feat_lo", "feat_hi", "out_mean", "out_std"):
            np.asarray(model[key], "<f8").tofile(f)
        params = np.asarray(model["params"], "<f4")
        np.array([params.size], "<i4").tofile(f)
        params.tofile(f)
