"""Train the learned preconditioner's network on labelled systems, on the GPU.

    python -m cgxx_amd.train_poly --data labels_train.npz [...] [--heldout labels_heldout.npz ...]
        [--out cgxx_amd/data/poly_mlp.npz] [--steps 6000]

Inputs are the label files of cgxx_amd.label.  First the per-degree lower
ends a = c_k theta_min with the least summed cost on the training systems
(learned.tune_lower_ends).  Inputs (the 37 matrix features) and targets
(learned.targets: log theta_min; per candidate degree the log of the
iterations at c_k theta_min and of the device's ms per iteration; the log
of Jacobi's iterations and ms per iteration) are standardised; the network
(learned.DIMS) is trained by full-batch Adam with the library's HIP
kernels.  Also tunes the method without the network on the same systems:
the one degree and ratio to theta10 with the least summed cost
(learned.tune_fixed_rule).  Prints a JSON summary with, for the held-out
files, the RMSE of each output group against predicting the training mean.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import learned

FIELDS = ("features", "grid_iters", "lower_end_iters", "poly_ms_per_iter", "jacobi_iters",
          "jacobi_ms_per_iter", "theta_min", "n", "seed")


def load(paths) -> dict:
    parts = []
    for p in paths:
        with np.load(p, allow_pickle=False) as z:
            parts.append({k: z[k] for k in FIELDS})
    return {k: np.concatenate([p[k] for p in parts]) for k in FIELDS}


D = len(learned.DEGREES)
GROUPS = {"theta_min": slice(0, 1), "iters": slice(1, 1 + D), "ms_per_iter": slice(1 + D, 1 + 2 * D),
          "jacobi": slice(1 + 2 * D, 3 + 2 * D)}


def group_rmse(pred, Y) -> dict:
    return {g: float(np.sqrt(np.mean((pred[:, s] - Y[:, s]) ** 2))) for g, s in GROUPS.items()}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--data", required=True, nargs="+")
    ap.add_argument("--heldout", nargs="*", default=[])
    ap.add_argument("--out", default=learned.MODEL_PATH)
    ap.add_argument("--steps", type=int, default=6000)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    args = ap.parse_args()

    d = load(args.data)
    X = d["features"][:, :learned.N_FEATURES].astype(np.float64)
    lower = learned.tune_lower_ends(d)
    Y = learned.targets(d, lower)
    fm, fs = X.mean(0), X.std(0)
    fs = np.where(fs < 1e-6, 1.0, fs)
    om, os_ = Y.mean(0), Y.std(0)
    os_ = np.where(os_ < 1e-6, 1.0, os_)
    Xn = ((X - fm) / fs).astype(np.float32)
    Yn = ((Y - om) / os_).astype(np.float32)
    fixed_degree, rho_star = learned.tune_fixed_rule(d)

    from .cghip import Mlp

    t0 = time.perf_counter()
    mlp = Mlp(learned.DIMS, Xn.shape[0])
    mlp.set_params(learned.init_params(learned.DIMS, args.seed))
    n1 = int(args.steps * 0.75)
    h1 = mlp.train(Xn, Yn, n1, lr=args.lr)
    h2 = mlp.train(Xn, Yn, args.steps - n1, lr=args.lr * 0.1)
    hist = np.concatenate([h1, h2])
    params = mlp.get_params()
    train_s = time.perf_counter() - t0
    pred = mlp.forward(Xn).astype(np.float64) * os_ + om
    summary = {"n_train": int(X.shape[0]), "steps": args.steps, "lr": args.lr,
               "train_seconds": train_s, "loss_first": float(hist[0]), "loss_last": float(hist[-1]),
               "train_rmse": group_rmse(pred, Y), "fixed_degree": fixed_degree, "rho_star": rho_star,
               "lower_end": {int(k): float(learned.LOWER_END_GRID[g]) for k, g in zip(learned.DEGREES, lower)}}
    if args.heldout:
        h = load(args.heldout)
        Yh = learned.targets(h, lower)
        Xc = np.clip(h["features"][:, :learned.N_FEATURES].astype(np.float64), X.min(0), X.max(0))  # as inference clamps
        ph = mlp.forward(((Xc - fm) / fs).astype(np.float32)).astype(np.float64) * os_ + om
        summary["n_heldout"] = int(Yh.shape[0])
        summary["heldout_rmse"] = group_rmse(ph, Yh)
        summary["heldout_rmse_mean_predictor"] = group_rmse(np.broadcast_to(om, Yh.shape), Yh)
    mlp.close()
    learned.save_model(args.out, params, fm, fs, om, os_, rho_star=rho_star, fixed_degree=fixed_degree,
                       lower_end=learned.LOWER_END_GRID[lower], lower_end_index=lower,
                       feat_lo=X.min(0), feat_hi=X.max(0),
                       loss_history=hist[:: max(1, len(hist) // 200)])
    learned.export_bin(learned.load_model(args.out), os.path.splitext(args.out)[0] + ".bin")
    print(json.dumps(summary), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
