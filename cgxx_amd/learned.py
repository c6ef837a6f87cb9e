"""Learned polynomial preconditioner: features, labels, the model and its use.

A small network (include/cghip_mlp.h, trained on the GPU by the library's
own kernels) maps statistics of the matrix alone -- no solver steps -- to
the bottom of the spectrum of D^-1 A and to what the mixed-precision solve
would cost on this device with Jacobi and with the Chebyshev polynomial
preconditioner of each candidate degree.  The solve then runs the cheapest
choice, the polynomial on [c_k theta_min, b] with theta_min predicted.  The
method without the network must first estimate the spectrum: 10 FP64
Jacobi-PCG steps for the smallest Ritz value theta10, then Chebyshev of one
tuned degree on [theta10 / rho*, b].

Inputs (37):
  0..35   cghip_matrix_stats: mean/std/min/max of log a_ii, the Gershgorin
          radius sum_{j!=i}|a_ij|/a_ii, log strongest and weakest coupling
          |a_ij|/a_ii and their ratio; the row-length histogram
  36      log N
The harvest also records 5 numbers of the first 10 FP64 Jacobi-PCG steps
(the reference's loop, CG.cpp:284-385; features 37..41: log of the smallest
and largest Ritz values after 10 and 5 steps, log10(r2_10 / r2_0) / 10):
the fixed rule's theta10, not an input of the network.
Outputs (2 D + 3, D = len(DEGREES)): log theta_min (the smallest eigenvalue
of D^-1 A); for each candidate degree k the log of the mixed-precision PCG
iterations with Chebyshev on [c_k theta_min, b] and the log of its time per
iteration (ms) on this device; the log of Jacobi's iterations and of its
time per iteration.  b = 1 + max Gershgorin radius bounds the spectrum from
above; c_k (one per degree) is tuned on the training systems.

Labels come from the spectrum the FP64 solver itself reveals: the Lanczos
tridiagonal of a converged Jacobi-PCG run (its alpha/beta) gives Ritz nodes
theta_i and Gauss weights w_i of the initial residual's spectral measure
(Golub & Meurant, Matrices, Moments and Quadrature, 2010, ch. 6); its
smallest node is theta_min.  PCG with the polynomial acts on the same
eigenvectors with eigenvalues theta_i p(theta_i) and weights p(theta_i) w_i,
so its iterations are those of CG on that diagonal system
(simulate_iterations), which matched the device's iteration counts within
one on 256^3 and eight family members for degrees 6-16.  The m-node measure
of an m-step run reproduces the moments of a polynomial PCG run while
iterations x (degree + 1) stay below about 2 m; far below the best lower
end the simulated counts fall short of the true ones
(tests/test_spectral_labels.py), so a label is exact where it matters, near
the best interval.  The times per iteration are measured on the device when
the system is harvested (cgxx_amd.harvest).  On this family the simulated
cost is flat near its optimum: per system, the best degree and lower end
save 0.2-0.3% against one tuned degree and ratio, so the network's gain is
the spectral estimate it replaces, not a per-system choice of interval.

Use (solve_learned): predicted cost = predicted iterations x predicted time
per iteration, for Jacobi and every degree; the cheapest runs.  The
polynomial gets an iteration budget worth 1.25x Jacobi's predicted cost and
the library refuses a polynomial that is not positive on [0, b]; in either
case Jacobi continues from the current x to the FP64 true residual.
"""
from __future__ import annotations

import os
import time

import numpy as np

from . import spectrum

# Candidate degrees of the polynomial.  Lower ends of the Chebyshev
# interval [a, b] searched: for the network, a = c theta_min with c on
# LOWER_END_GRID (quarter octaves, 2^-2 .. 2^8); for the fixed rule,
# a = theta10 r with r on RATIO_GRID (quarter octaves, 2^0.5 .. 2^-6).
DEGREES = (4, 6, 8, 12, 16)
LOWER_END_GRID = 2.0 ** (np.arange(-8, 33) / 4.0)
RATIO_GRID = 2.0 ** (-np.arange(-2, 25) / 4.0)
# the degree of the fixed rule before it is tuned on a training set
DEGREE = 8
N_STATS = 36
N_FEATURES = N_STATS + 1
N_HARVESTED = N_FEATURES + 5
N_OUT = 2 * len(DEGREES) + 3
DIMS = (N_FEATURES, 128, 128, N_OUT)
RITZ_STEPS = 10
MODEL_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "poly_mlp.npz")


def out_index(kind: str, j: int = 0) -> int:
    """Position of an output: "theta_min"; "iters" or "ms" of degree
    DEGREES[j]; "jacobi_iters" / "jacobi_ms"."""
    D = len(DEGREES)
    return {"theta_min": 0, "iters": 1 + j, "ms": 1 + D + j, "jacobi_iters": 1 + 2 * D,
            "jacobi_ms": 2 + 2 * D}[kind]


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
    """The 42 harvested features: 36 statistics, log N, cg_statistics."""
    stats = np.asarray(stats, np.float64)
    if stats.size != N_STATS:
        raise ValueError(f"expected {N_STATS} matrix statistics")
    return np.concatenate([stats, [np.log(n)], cg_statistics(alpha, beta, r2)])


def spectrum_bound(stats) -> float:
    """b = 1 + max Gershgorin radius of D^-1 A (stats[7] = max of q1)."""
    return 1.0 + float(stats[7])


def matrix_features(ctx, n: int):
    """(features, b) from the matrix alone: the device matrix statistics and
    log N (the network's inputs), and b = 1 + max Gershgorin radius."""
    stats = np.asarray(ctx.matrix_stats(), np.float64)
    if stats.size != N_STATS:
        raise ValueError(f"expected {N_STATS} matrix statistics")
    return np.concatenate([stats, [np.log(n)]]), spectrum_bound(stats)


def ritz_features(ctx):
    """(5 features, theta10) of 10 FP64 Jacobi-PCG steps from x = 0 on the
    uploaded system (CRS, Jacobi, K set): what a rule without the network
    needs for its interval.  Overwrites X."""
    ctx.fill("X", 0.0)
    ctx.solve(precond="jacobi", max_iter=RITZ_STEPS, tol=1e-300)
    a = ctx.history("alpha", RITZ_STEPS)
    bt = ctx.history("beta", RITZ_STEPS)
    r2 = ctx.history("r2", RITZ_STEPS + 1)
    c = cg_statistics(a, bt, r2)
    return c, float(np.exp(c[0]))


def system_features(ctx, n: int):
    """(all 42 harvested features, b, theta10): matrix_features then
    ritz_features.  Overwrites X."""
    f, b = matrix_features(ctx, n)
    c, theta10 = ritz_features(ctx)
    return np.concatenate([f, c]), b, theta10


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


def grid_counts(theta, w, lower_ends, b: float, degree: int, tol: float = 1e-9) -> np.ndarray:
    """Simulated PCG iterations (simulate_iterations) of the Chebyshev
    polynomial of the given degree on [a, b] for every a of lower_ends
    (capped at b / 4); -1 where the polynomial is not positive on the
    nodes."""
    its = np.full(len(lower_ends), -1)
    for i, a in enumerate(lower_ends):
        c = spectrum.cheb_residual_coeffs(min(a, 0.25 * b), b, degree)
        pv = spectrum.poly_eval(c, b, theta)
        if np.all(pv > 0):
            its[i] = simulate_iterations(theta, w, pv, tol=tol)
    return its


def best_lower_end(theta, w, theta10: float, b: float, degree: int, tol: float = 1e-9):
    """(ratio r, simulated iterations, counts on RATIO_GRID) of the
    Chebyshev polynomial of the given degree on [theta10 r, b] with the
    fewest simulated PCG iterations on the spectral measure (theta, w); -1
    where the polynomial is not positive on the Ritz nodes.  Ties go to the
    larger r."""
    its = grid_counts(theta, w, theta10 * RATIO_GRID, b, degree, tol)
    ok = its >= 0
    if not ok.any():
        raise ValueError("no positive Chebyshev polynomial on the grid")
    i = int(np.argmin(np.where(ok, its, np.iinfo(its.dtype).max)))
    return float(RATIO_GRID[i]), int(its[i]), its


def label_degrees(alpha, beta, theta10: float, b: float, degrees=DEGREES, tol: float = 1e-9) -> dict:
    """Labels of one harvested system: for each degree the best lower-end
    ratio relative to theta10, its simulated iterations and the counts on
    the whole ratio grid (best_lower_end, for the fixed rule); the counts
    with a = c theta_min on LOWER_END_GRID (lower_end_iters, for the
    network); the simulated Jacobi iterations (p = 1); the extreme Ritz
    values of the full run."""
    theta, w = spectral_measure(alpha, beta)
    theta = np.clip(theta, 1e-300, None)
    out = {"ratio": [], "iters": [], "grid_iters": [], "lower_end_iters": []}
    for k in degrees:
        r, it, its = best_lower_end(theta, w, theta10, b, k, tol)
        out["ratio"].append(r)
        out["iters"].append(it)
        out["grid_iters"].append(its)
        out["lower_end_iters"].append(grid_counts(theta, w, theta[0] * LOWER_END_GRID, b, k, tol))
    out["jacobi_iters"] = simulate_iterations(theta, w, np.ones_like(theta), tol=tol)
    out["theta_min"] = float(theta[0])
    out["theta_max"] = float(theta[-1])
    return out


def _cost(iters, ms):
    it = np.asarray(iters, np.float64).copy()
    it[it < 0] = np.inf
    return it * ms


def tune_lower_ends(d: dict) -> np.ndarray:
    """Per degree the index on LOWER_END_GRID of the multiplier c (a =
    c theta_min) with the least summed cost (simulated iterations x measured
    ms per iteration) over labelled systems."""
    cost = _cost(d["lower_end_iters"], d["poly_ms_per_iter"][:, :, None]).sum(0)
    return np.argmin(cost, axis=1)


def targets(d: dict, lower_end_index) -> np.ndarray:
    """Network targets of labelled systems (cgxx_amd.label output): log
    theta_min; per degree log iterations at a = c_k theta_min (c_k =
    LOWER_END_GRID[lower_end_index[k]]) and log ms per iteration; log Jacobi
    iterations and log Jacobi ms per iteration."""
    S = d["lower_end_iters"].shape[0]
    it = np.stack([d["lower_end_iters"][:, j, g] for j, g in enumerate(lower_end_index)], 1)
    if np.any(it <= 0):
        raise ValueError("a labelled count at the tuned lower end is not positive")
    return np.concatenate([np.log(d["theta_min"]).reshape(S, 1), np.log(it.astype(np.float64)),
                           np.log(d["poly_ms_per_iter"]),
                           np.log(d["jacobi_iters"].astype(np.float64))[:, None],
                           np.log(d["jacobi_ms_per_iter"])[:, None]], axis=1)


def tune_fixed_rule(d: dict) -> tuple[int, float]:
    """The method without the network, tuned on labelled systems: the one
    degree and ratio r (Chebyshev on [theta10 r, b]) with the least summed
    cost (simulated iterations x measured ms per iteration).  Returns
    (degree, rho = 1 / r)."""
    cost = _cost(d["grid_iters"], d["poly_ms_per_iter"][:, :, None]).sum(0)
    j, g = np.unravel_index(int(np.argmin(cost)), cost.shape)
    return int(DEGREES[j]), float(1.0 / RATIO_GRID[g])


# ---- model file and inference ---------------------------------------------

def save_model(path, params, feat_mean, feat_std, out_mean, out_std, **meta):
    np.savez(path, params=np.asarray(params, np.float32), dims=np.array(DIMS),
             feat_mean=feat_mean, feat_std=feat_std, out_mean=out_mean, out_std=out_std,
             degrees=np.array(DEGREES), **{k: np.asarray(v) for k, v in meta.items()})


def load_model(path=MODEL_PATH) -> dict:
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def normalize(f, model) -> np.ndarray:
    """Standardised network inputs (the first N_FEATURES of f), first
    clamped to the range seen in training (a system outside the family, e.g.
    a larger grid, is mapped to its nearest training edge instead of being
    extrapolated)."""
    f = np.atleast_2d(np.asarray(f, np.float64))[:, :len(model["feat_mean"])]
    f = np.clip(f, model["feat_lo"], model["feat_hi"])
    return ((f - model["feat_mean"]) / model["feat_std"]).astype(np.float32)


def decide(y, model) -> dict:
    """The choice from de-standardised network outputs y (one system):
    predicted theta_min, predicted cost (ms) of Jacobi and of every degree,
    and the cheapest with its lower end a = c_k theta_min."""
    degrees = [int(v) for v in model["degrees"]]
    D = len(degrees)
    cost = {"jacobi": float(np.exp(y[1 + 2 * D] + y[2 + 2 * D]))}
    for j, k in enumerate(degrees):
        cost[k] = float(np.exp(y[1 + j] + y[1 + D + j]))
    best = min(cost, key=cost.get)
    theta_min = float(np.exp(y[0]))
    out = {"cost_ms": cost, "choice": best, "theta_min": theta_min,
           "jacobi_ms_per_iter": float(np.exp(y[2 + 2 * D]))}
    if best != "jacobi":
        j = degrees.index(best)
        out.update(degree=best, lower_end=float(model["lower_end"][j]) * theta_min,
                   iters=float(np.exp(y[1 + j])), ms_per_iter=float(np.exp(y[1 + D + j])))
    return out


def predict(mlp, model, f) -> dict:
    """Network outputs for features f (on the GPU), de-standardised, and the
    decision (decide)."""
    yn = mlp.forward(normalize(f, model))[0].astype(np.float64)
    y = yn * model["out_std"] + model["out_mean"]
    return dict(decide(y, model), y=y)


def fixed_rule_coeffs(theta10: float, b: float, rho: float, degree: int = None) -> np.ndarray:
    """The method without the network: Chebyshev on [theta_min(10) / rho, b]."""
    return spectrum.cheb_residual_coeffs(min(theta10 / rho, 0.25 * b), b,
                                         DEGREE if degree is None else degree)


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


def solve_learned(ctx, mlp, model, n: int, tol: float = 1e-9, max_iter: int = 100000) -> dict:
    """Matrix statistics on the device, the network on the GPU, the cheapest
    predicted method, then the solve.

    Jacobi runs from the start when its predicted cost is the lowest;
    otherwise the polynomial of the chosen degree on [c_k theta_min, b]
    (theta_min predicted) runs with an iteration budget worth 1.25x Jacobi's
    predicted cost (solve_poly_with_fallback: past it, or if the library
    refuses the polynomial, Jacobi takes over).  The wall time includes the
    statistics and inference."""
    t0 = time.perf_counter()
    f, b = matrix_features(ctx, n)
    pr = predict(mlp, model, f)
    if pr["choice"] == "jacobi":
        ctx.fill("X", 0.0)
        info = ctx.solve_mixed(precond="jacobi", max_iter=max_iter, tol=tol)
        res = {"method": "jacobi-selected", "poly_iterations": 0,
               "jacobi_iterations": info["iterations"], "updates": info["updates"],
               "true_residual": info["true_residual"], "spmvs": info["iterations"]}
        c = None
    else:
        k = pr["degree"]
        c = spectrum.cheb_residual_coeffs(min(pr["lower_end"], 0.25 * b), b, k)
        budget_iters = int(np.ceil(1.25 * pr["cost_ms"]["jacobi"] / pr["ms_per_iter"]))
        res = solve_poly_with_fallback(ctx, c, b, tol, max_iter, budget=budget_iters * (k + 1))
    res["wall_s"] = time.perf_counter() - t0
    res["coeffs"] = c
    res["b"] = b
    res["decision"] = {key: pr[key] for key in pr if key != "y"}
    return res


def solve_fixed_rule(ctx, n: int, rho: float, degree: int, tol: float = 1e-9,
                     max_iter: int = 100000) -> dict:
    """The method without the network: b from the matrix statistics,
    theta10 from 10 FP64 Jacobi-PCG steps, then Chebyshev of the given
    degree on [theta10 / rho, b] (solve_poly_with_fallback).  The wall time
    includes both spectral estimates."""
    t0 = time.perf_counter()
    _, b = matrix_features(ctx, n)
    _, theta10 = ritz_features(ctx)
    c = fixed_rule_coeffs(theta10, b, rho, degree)
    res = solve_poly_with_fallback(ctx, c, b, tol, max_iter)
    res["spmvs"] += RITZ_STEPS  # the FP64 steps of the estimate
    res["wall_s"] = time.perf_counter() - t0
    res["coeffs"] = c
    res["b"] = b
    res["theta10"] = theta10
    return res


BIN_MAGIC = b"CGHIPMLP"


def export_bin(model: dict, path: str) -> None:
    """Writes the model for the C++ host (cgxx_amd/host/LearnedPoly.cpp), all
    little-endian: magic "CGHIPMLP", int32 version 3, int32 in, h1, h2, out,
    int32 D, int32 degrees[D], int32 fixed degree, float64 fixed rho,
    float64 lower_end[D] (c_k), float64 feat_mean[in], feat_std[in],
    feat_lo[in], feat_hi[in], out_mean[out], out_std[out], int32 nparams,
    float32 params[nparams]."""
    dims = [int(v) for v in model["dims"]]
    degrees = [int(v) for v in model["degrees"]]
    with open(path, "wb") as f:
        f.write(BIN_MAGIC)
        np.array([3, *dims, len(degrees), *degrees, int(model["fixed_degree"])], "<i4").tofile(f)
        np.array([float(model["rho_star"])], "<f8").tofile(f)
        np.asarray(model["lower_end"], "<f8").tofile(f)
        for key in ("feat_mean", "feat_std", "feat_lo", "feat_hi", "out_mean", "out_std"):
            np.asarray(model[key], "<f8").tofile(f)
        params = np.asarray(model["params"], "<f4")
        np.array([params.size], "<i4").tofile(f)
        params.tofile(f)
