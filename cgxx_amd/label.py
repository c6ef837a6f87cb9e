"""Labels of the learned polynomial preconditioner from raw harvests (CPU).

For every system of the raw files written by cgxx_amd.harvest: the spectral
measure of its FP64 solve (Ritz nodes and Gauss weights of the Lanczos
tridiagonal of its alpha/beta), then for every candidate degree
(learned.DEGREES) the simulated PCG iterations to 1e-9 of the Chebyshev
polynomial on [a, b] for every a = c theta_min (c on
learned.LOWER_END_GRID) and every a = theta10 r (r on learned.RATIO_GRID),
and the simulated Jacobi iterations (learned.label_degrees).  The device's
times per iteration and feature passes measured by the harvest come along
(NaN where a raw file has no such field).

    python -m cgxx_amd.label --raw raw1.npz [raw2.npz ...] --out labels.npz [--workers 8]

The output keeps the features and family parameters of the raw files and
adds: degrees, ratio[s, d], iters[s, d], grid_iters[s, d, g], ratio_grid,
lower_end_iters[s, d, g], lower_end_grid, jacobi_iters[s], theta_min[s],
theta_max[s].
"""
from __future__ import annotations

import argparse
import multiprocessing as mp
import sys
import time

import numpy as np

from . import learned

KEEP = ("seed", "n", "b", "features", "fp64_iterations", "fp64_seconds", "ratio_y", "ratio_z",
        "sigma", "corr", "side", "jacobi_ms_per_iter", "poly_ms_per_iter", "poly_iterations")
OPTIONAL = ("stats_ms", "ritz_ms")


def _one(args):
    alpha, beta, theta10, b = args
    return learned.label_degrees(alpha, beta, theta10, b)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--raw", nargs="+", required=True)
    ap.add_argument("--out", required=True)
    ap.add_argument("--workers", type=int, default=8)
    args = ap.parse_args()
    parts = []
    for path in args.raw:
        with np.load(path, allow_pickle=False) as z:
            parts.append({k: z[k] for k in z.files})
    jobs, keep = [], {k: [] for k in KEEP + OPTIONAL}
    for z in parts:
        for i in range(z["seed"].size):
            a = z["alpha"][z["alpha_offset"][i]:z["alpha_offset"][i + 1]]
            bt = z["beta"][z["beta_offset"][i]:z["beta_offset"][i + 1]]
            theta10 = float(np.exp(z["features"][i][37]))
            jobs.append((a, bt, theta10, float(z["b"][i])))
            for k in KEEP:
                keep[k].append(z[k][i])
            for k in OPTIONAL:
                keep[k].append(z[k][i] if k in z else np.nan)
    t0 = time.perf_counter()
    with mp.get_context("spawn").Pool(args.workers) as pool:
        res = []
        for j, r in enumerate(pool.imap(_one, jobs, chunksize=4)):
            res.append(r)
            if (j + 1) % 100 == 0:
                print(f"[label] {j + 1}/{len(jobs)} systems, {time.perf_counter() - t0:.0f} s", flush=True)
    out = {k: np.array(v) for k, v in keep.items()}
    out["degrees"] = np.array(learned.DEGREES)
    out["ratio_grid"] = learned.RATIO_GRID
    out["ratio"] = np.array([r["ratio"] for r in res])
    out["iters"] = np.array([r["iters"] for r in res], np.int32)
    out["grid_iters"] = np.array([r["grid_iters"] for r in res], np.int32)
    out["lower_end_grid"] = learned.LOWER_END_GRID
    out["lower_end_iters"] = np.array([r["lower_end_iters"] for r in res], np.int32)
    for k in ("jacobi_iters", "theta_min", "theta_max"):
        out[k] = np.array([r[k] for r in res])
    np.savez_compressed(args.out, **out)
    print(f"[label] wrote {args.out}: {len(res)} systems in {time.perf_counter() - t0:.0f} s", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
