"""Harvest training data for the learned polynomial preconditioner.

For each seed of the anisotropic log-normal diffusion family (grids 64^3 to
160^3, axis weight ratios log-uniform in [1, 1e3], sigma in [0.5, 1.5]):
generate the system on the host (worker processes), solve it on the GPU
with the FP64 Jacobi-PCG loop (the reference's method), keep its alpha/beta,
and compute on the host
  - the 42 features (device matrix statistics + the first 10 PCG steps),
  - the label: the Chebyshev lower end whose simulated iteration count on
    the harvested spectrum is smallest (learned.label),
  - the simulated iterations of the fixed rule a = theta_min(10 steps) / rho
    for rho in RHO, the method without the network.

    python -m cgxx_amd.harvest --seeds 0:1000 --out harvest.npz [--workers 16] [--sides 176,256]

Prints one progress line per 20 systems.
"""
from __future__ import annotations

import argparse
import multiprocessing as mp
import sys
import time

import numpy as np

RHO = 2.0 ** np.arange(0, 11)  # 1 .. 1024
MAX_ITER = 8000


def _generate(seed: int, sides=None):
    from cgxx_amd import problems

    m = problems.anisotropic_family(seed) if sides is None else problems.anisotropic_family(seed, sides)
    return seed, m.ptr, m.index, m.value, m.ones_rhs(), 1.0 / m.diagonal(), m.info


def _label(seed, alpha, beta, b, theta10):
    from cgxx_amd import learned, spectrum

    L = learned.label(alpha, beta, b)
    theta, w = learned.spectral_measure(alpha, beta)
    theta = np.clip(theta, 1e-300, None)
    rho_iters = []
    for rho in RHO:
        c = learned.fixed_rule_coeffs(theta10, b, rho)
        rho_iters.append(learned.simulate_iterations(theta, w, spectrum.poly_eval(c, b, theta)))
    L["rho_iters"] = np.array(rho_iters)
    return seed, L


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--seeds", required=True, help="first:last (last excluded)")
    ap.add_argument("--out", required=True)
    ap.add_argument("--workers", type=int, default=16)
    ap.add_argument("--inflight", type=int, default=20, help="systems generated ahead at most")
    ap.add_argument("--sides", default=None,
                    help="comma-separated grid sides to draw from (default: 64..160 by 16)")
    args = ap.parse_args()
    lo, hi = (int(v) for v in args.seeds.split(":"))
    sides = tuple(int(v) for v in args.sides.split(",")) if args.sides else None
    seeds = list(range(lo, hi))

    # Worker processes start before this process touches the GPU.
    pool = mp.get_context("spawn").Pool(args.workers)

    from cgxx_amd import learned
    from cgxx_amd.cghip import Context

    ctx = Context(0)
    pending = {}
    labels = {}
    rec = {k: [] for k in ("seed", "n", "b", "features", "fp64_iterations", "fp64_seconds",
                           "ratio_y", "ratio_z", "sigma", "corr", "side")}
    gen_q = []
    nxt = 0
    t0 = time.perf_counter()

    def top_up():
        nonlocal nxt
        while nxt < len(seeds) and len(gen_q) < args.inflight:
            gen_q.append(pool.apply_async(_generate, (seeds[nxt], sides)))
            nxt += 1

    top_up()
    done = 0
    while gen_q:
        seed, ptr, idx, val, k, C, info = gen_q.pop(0).get()
        top_up()
        n = ptr.size - 1
        ctx.set_size(n)
        ctx.upload_crs(ptr, idx, val)
        ctx.set_format("CRS")
        ctx.upload_jacobi(C)
        ctx.upload("K", k)
        del ptr, idx, val, k, C
        f, b, theta10 = learned.system_features(ctx, n)
        ctx.fill("X", 0.0)
        s = ctx.solve(precond="jacobi", max_iter=MAX_ITER, tol=1e-9)
        it = s["iterations"]
        alpha = ctx.history("alpha", it)
        beta = ctx.history("beta", it)
        pending[seed] = pool.apply_async(_label, (seed, alpha[: it], beta[: max(it - 1, 0)], b, theta10))
        rec["seed"].append(seed)
        rec["n"].append(n)
        rec["b"].append(b)
        rec["features"].append(f)
        rec["fp64_iterations"].append(it)
        rec["fp64_seconds"].append(s["wall_s"])
        for key in ("ratio_y", "ratio_z", "sigma", "corr", "side"):
            rec[key].append(info[key])
        done += 1
        if done % 20 == 0 or done == len(seeds):
            print(f"[harvest] {done}/{len(seeds)} systems, {time.perf_counter() - t0:.0f} s "
                  f"(last: seed {seed}, N={n}, {it} FP64 iterations in {s['wall_s']:.3f} s)",
                  flush=True)
    ctx.close()
    for seed, res in pending.items():
        labels[seed] = res.get()[1]
    pool.close()
    pool.join()

    S = len(rec["seed"])
    A = learned.A_STEPS
    out = {k: np.array(v) for k, v in rec.items()}
    out["y"] = np.stack([labels[s]["y"] for s in rec["seed"]])
    out["a_best"] = np.array([labels[s]["a_best"] for s in rec["seed"]])
    out["theta_min"] = np.array([labels[s]["theta_min"] for s in rec["seed"]])
    out["theta_max"] = np.array([labels[s]["theta_max"] for s in rec["seed"]])
    out["sim_jacobi"] = np.array([labels[s]["jacobi_iters"] for s in rec["seed"]])
    out["sim_iters"] = np.full((S, A), -1)
    out["cands"] = np.full((S, A), np.nan)
    for i, s in enumerate(rec["seed"]):
        L = labels[s]
        out["sim_iters"][i, : L["iters"].size] = L["iters"]
        out["cands"][i, : L["cands"].size] = L["cands"]
    out["rho"] = RHO
    out["rho_iters"] = np.stack([labels[s]["rho_iters"] for s in rec["seed"]])
    np.savez_compressed(args.out, **out)
    print(f"[harvest] wrote {args.out}: {S} systems in {time.perf_counter() - t0:.0f} s", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
