"""Train the coefficient network on harvested data, on the GPU.

    python -m cgxx_amd.train_poly --data train.npz [--heldout heldout.npz]
        [--out cgxx_amd/data/poly_mlp.npz] [--steps 6000]

Inputs are the harvest files of cgxx_amd.harvest.  Features and targets
(log p at the Chebyshev points of the best Chebyshev polynomial) are
standardised; the network (learned.DIMS) is trained by full-batch Adam with
the library's HIP kernels.  Also picks rho*, the single ratio of the fixed
rule a = theta_min(10 steps) / rho with the fewest simulated iterations
over the training systems: the method without the network.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import learned


def load(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def targets(d) -> np.ndarray:
    """log p at the Chebyshev points of the best Chebyshev polynomial, the
    log of the FP64 Jacobi-PCG iteration count, the log of the best
    simulated polynomial iteration count."""
    it = d["sim_iters"].astype(np.float64)
    it[it < 0] = np.inf
    return np.concatenate([d["y"].astype(np.float64),
                           np.log(d["fp64_iterations"].astype(np.float64))[:, None],
                           np.log(it.min(1))[:, None]], axis=1)


def best_rho(d) -> float:
    it = d["rho_iters"].astype(np.float64)
    it[it < 0] = np.inf
    tot = it.sum(0)
    return float(d["rho"][int(np.argmin(tot))])


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--data", required=True, nargs="+")
    ap.add_argument("--heldout", nargs="*", default=[])
    ap.add_argument("--out", default=learned.MODEL_PATH)
    ap.add_argument("--steps", type=int, default=6000)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    args = ap.parse_args()

    parts = [load(p) for p in args.data]
    d = {k: np.concatenate([p[k] for p in parts])
         for k in ("features", "y", "rho_iters", "fp64_iterations", "sim_iters")}
    d["rho"] = parts[0]["rho"]
    X = d["features"].astype(np.float64)
    Y = targets(d)
    fm, fs = X.mean(0), X.std(0)
    fs = np.where(fs < 1e-6, 1.0, fs)
    om, os_ = Y.mean(0), Y.std(0)
    os_ = np.where(os_ < 1e-6, 1.0, os_)
    Xn = ((X - fm) / fs).astype(np.float32)
    Yn = ((Y - om) / os_).astype(np.float32)

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
    train_rmse = float(np.sqrt(np.mean((pred[:, :learned.DEGREE + 1] - Y[:, :learned.DEGREE + 1]) ** 2)))
    summary = {"n_train": int(X.shape[0]), "steps": args.steps, "lr": args.lr,
               "train_seconds": train_s, "loss_first": float(hist[0]), "loss_last": float(hist[-1]),
               "train_rmse_log_p": train_rmse, "rho_star": best_rho(d)}
    if args.heldout:
        hp = [load(p) for p in args.heldout]
        Xh = np.concatenate([p["features"] for p in hp]).astype(np.float64)
        Yh = targets({k: np.concatenate([p[k] for p in hp])
                      for k in ("y", "fp64_iterations", "sim_iters")})
        Xc = np.clip(Xh, X.min(0), X.max(0))  # as inference clamps to the training range
        ph = mlp.forward(((Xc - fm) / fs).astype(np.float32)).astype(np.float64) * os_ + om
        summary["n_heldout"] = int(Xh.shape[0])
        k = learned.DEGREE + 1
        summary["heldout_rmse_log_p"] = float(np.sqrt(np.mean((ph[:, :k] - Yh[:, :k]) ** 2)))
        summary["heldout_rmse_mean_predictor"] = float(np.sqrt(np.mean((om[:k] - Yh[:, :k]) ** 2)))
        summary["heldout_iterations_rel_err"] = [
            float(np.mean(np.abs(np.exp(ph[:, k + i] - Yh[:, k + i]) - 1))) for i in (0, 1)]
    mlp.close()
    learned.save_model(args.out, params, fm, fs, om, os_, rho_star=summary["rho_star"],
                       feat_lo=X.min(0), feat_hi=X.max(0),
                       loss_history=hist[:: max(1, len(hist) // 200)])
    learned.export_bin(learned.load_model(args.out), os.path.splitext(args.out)[0] + ".bin")
    print(json.dumps(summary), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
