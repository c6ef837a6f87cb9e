"""What the trained network gains over the method without it (CPU).

On the committed held-out label files (64^3-160^3 and 176^3-256^3 members
of the anisotropic diffusion family, never seen in training), every method
is priced as the harvest measured it on the device: simulated PCG
iterations of its polynomial (exact on the device within one iteration near
the best interval) times the measured time per iteration of its degree (or
Jacobi's), plus the measured wall time of the spectral estimates it needs:

  learned  matrix statistics (stats_ms), then the network's choice: the
           degree k with the lowest predicted cost on [c_k theta_min_pred, b]
           -- priced at the grid point nearest c_k theta_min_pred on the
           system's own lower-end grid -- or Jacobi
  fixed    matrix statistics for b and 10 FP64 PCG steps for theta10
           (stats_ms + ritz_ms), then Chebyshev of the tuned degree on
           [theta10 / rho*, b]
  ritz     the same estimates, Chebyshev of degree 8 on the raw 10-step
           interval [theta10, b]
  jacobi   mixed-precision Jacobi

The network runs in the float64 restatement of its forward pass
(oracle/mlp_oracle.py).  Also the held-out RMSE of each output group
against predicting the training mean.
"""
import glob
import os

import numpy as np
import pytest

from cgxx_amd import learned
from oracle import mlp_oracle
from tests.conftest import REPO

DATA = os.path.join(REPO, "cgxx_amd", "data")
HELDOUT = sorted(glob.glob(os.path.join(DATA, "labels_c5*_heldout_*.npz")))


def _load(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _outputs(d, model):
    X = learned.normalize(d["features"], model).astype(np.float64)
    dims = tuple(int(v) for v in model["dims"])
    return mlp_oracle.forward(model["params"].astype(np.float64), dims, X)[2] * model["out_std"] \
        + model["out_mean"]


def _prices(d, model):
    """Per-system cost (ms) of each method, the network's choices and how
    many predicted lower ends fell outside the label grid."""
    degrees = list(learned.DEGREES)
    lg = np.log2(learned.LOWER_END_GRID)
    rg = np.log2(learned.RATIO_GRID)
    ms = d["poly_ms_per_iter"]
    Y = _outputs(d, model)
    jac = d["jacobi_iters"] * d["jacobi_ms_per_iter"]
    stats, ritz = d["stats_ms"], d["ritz_ms"]
    fixed_j = degrees.index(int(model["fixed_degree"]))
    g_fixed = int(np.argmin(np.abs(rg - np.log2(1.0 / float(model["rho_star"])))))
    g_ritz = int(np.argmin(np.abs(rg)))
    out = {"learned": [], "fixed": [], "ritz": [], "jacobi": jac, "choice": [], "outside": 0}
    for s in range(len(jac)):
        dec = learned.decide(Y[s], model)
        if dec["choice"] == "jacobi":
            out["choice"].append("jacobi")
            out["learned"].append(stats[s] + jac[s])
        else:
            j = degrees.index(dec["degree"])
            pos = np.log2(dec["lower_end"] / d["theta_min"][s])
            out["outside"] += int(pos < lg[0] - 0.125 or pos > lg[-1] + 0.125)
            g = int(np.argmin(np.abs(lg - pos)))
            out["choice"].append(str(dec["degree"]))
            out["learned"].append(stats[s] + d["lower_end_iters"][s, j, g] * ms[s, j])
        out["fixed"].append(stats[s] + ritz[s] + d["grid_iters"][s, fixed_j, g_fixed] * ms[s, fixed_j])
        out["ritz"].append(stats[s] + ritz[s] + d["grid_iters"][s, degrees.index(8), g_ritz]
                           * ms[s, degrees.index(8)])
    return {k: np.asarray(v) if isinstance(v, list) else v for k, v in out.items()}


@pytest.mark.skipif(not HELDOUT, reason="no held-out label files")
@pytest.mark.parametrize("path", HELDOUT, ids=[os.path.basename(p) for p in HELDOUT])
def test_network_gains_over_the_method_without_it(path):
    model = learned.load_model()
    d = _load(path)
    assert np.all(np.isfinite(d["stats_ms"])) and np.all(np.isfinite(d["ritz_ms"]))
    p = _prices(d, model)
    tot = {k: float(np.sum(p[k])) for k in ("learned", "fixed", "ritz", "jacobi")}
    choices = {c: int(np.sum(p["choice"] == c)) for c in sorted(set(p["choice"].tolist()))}
    print(f"{os.path.basename(path)}: summed ms " + ", ".join(f"{k} {v:.1f}" for k, v in tot.items())
          + f"; learned/fixed {tot['learned'] / tot['fixed']:.4f}, learned/ritz "
            f"{tot['learned'] / tot['ritz']:.4f}; choices {choices}; lower ends off the grid "
            f"{p['outside']}")
    assert np.all(np.isfinite(p["learned"])) and np.all(p["learned"] > 0)
    assert tot["learned"] <= tot["fixed"]
    assert tot["learned"] <= 0.92 * tot["ritz"]
    assert tot["learned"] <= 0.6 * tot["jacobi"]
    assert p["outside"] <= 0.02 * len(p["learned"])


@pytest.mark.skipif(not HELDOUT, reason="no held-out label files")
def test_predictions_beat_the_mean_predictor():
    """Held-out RMSE of each output group (log theta_min, iterations and
    times per iteration per degree, Jacobi's) at least 20% below predicting
    the training mean."""
    model = learned.load_model()
    d = {}
    for path in HELDOUT:
        for k, v in _load(path).items():
            d.setdefault(k, []).append(v)
    d = {k: np.concatenate(v) if v[0].ndim else v[0] for k, v in d.items()}
    Y = _outputs(d, model)
    T = learned.targets(d, model["lower_end_index"])
    D = len(learned.DEGREES)
    groups = {"theta_min": slice(0, 1), "iters": slice(1, 1 + D), "ms_per_iter": slice(1 + D, 1 + 2 * D),
              "jacobi": slice(1 + 2 * D, 3 + 2 * D)}
    for name, sl in groups.items():
        rmse = np.sqrt(np.mean((Y[:, sl] - T[:, sl]) ** 2))
        mean = np.sqrt(np.mean((model["out_mean"][sl] - T[:, sl]) ** 2))
        print(f"{name}: held-out RMSE {rmse:.4f}, mean predictor {mean:.4f}")
        assert rmse <= 0.8 * mean, name
