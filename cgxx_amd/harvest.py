"""Harvest raw training data for the learned polynomial preconditioner (GPU).

For each seed of the anisotropic log-normal diffusion family (grids 64^3 to
160^3 by default, axis weight ratios log-uniform in [1, 1e3], sigma in
[0.5, 1.5]): generate the system on the host (worker processes), then on the
GPU compute its 42 features (device matrix statistics + the first 10 FP64
PCG steps) and solve it with the FP64 Jacobi-PCG loop (the reference's
method, CG.cpp:284-385) to relative residual 1e-10, keeping every alpha and
beta: the Lanczos coefficients from which cgxx_amd.label computes, on the
CPU, the spectral measure and the labels.  Then the time per iteration of
the mixed-precision solver on this device: with Jacobi, and with the
Chebyshev polynomial on [theta10 / 16, b] of each candidate degree
(learned.DEGREES), each from x = 0 to 1e-9: after the run that builds its
graph, the fastest of three timed runs (wall time over inner iterations,
reliable updates included).

    python -m cgxx_amd.harvest --seeds 0:1000 --out raw.npz [--workers 12] [--sides 176,256]

Output (npz): seed, n, b, features, fp64_iterations, fp64_seconds, the
family parameters, the concatenated alpha/beta of every solve with their
offsets, jacobi_ms_per_iter, poly_ms_per_iter[s, d], poly_iterations, and
the wall times of the two feature passes (stats_ms: matrix statistics;
ritz_ms: the 10 FP64 PCG steps and their history).  Prints one progress
line per 20 systems.
"""
from __future__ import annotations

import argparse
import multiprocessing as mp
import sys
import time

import numpy as np

MAX_ITER = 20000
HARVEST_TOL = 1e-10
TIMED_RUNS = 3


def _generate(seed: int, sides=None):
    from cgxx_amd import problems

    m = problems.anisotropic_family(seed) if sides is None else problems.anisotropic_family(seed, sides)
    return seed, m.ptr, m.index, m.value, m.ones_rhs(), 1.0 / m.diagonal(), m.info


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--seeds", required=True, help="first:last (last excluded)")
    ap.add_argument("--out", required=True)
    ap.add_argument("--workers", type=int, default=12,
                    help="generator processes (leave host cores to the timed solves)")
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
    rec = {k: [] for k in ("seed", "n", "b", "features", "fp64_iterations", "fp64_seconds",
                           "ratio_y", "ratio_z", "sigma", "corr", "side", "jacobi_ms_per_iter",
                           "poly_ms_per_iter", "poly_iterations", "stats_ms", "ritz_ms")}
    alphas, betas = [], []
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
        f, b, _ = learned.system_features(ctx, n)
        # the two feature passes, timed on a second call
        t1 = time.perf_counter()
        learned.matrix_features(ctx, n)
        t2 = time.perf_counter()
        learned.ritz_features(ctx)
        rec["stats_ms"].append((t2 - t1) * 1e3)
        rec["ritz_ms"].append((time.perf_counter() - t2) * 1e3)
        ctx.fill("X", 0.0)
        s = ctx.solve(precond="jacobi", max_iter=MAX_ITER, tol=HARVEST_TOL)
        it = s["iterations"]
        alphas.append(ctx.history("alpha", it))
        betas.append(ctx.history("beta", it)[: max(it - 1, 0)])
        # time per inner iteration on this device, per method
        ctx.prepare_fp32()
        theta10 = float(np.exp(f[37]))

        def per_iter(precond):
            # the first run builds the graph; the fastest of the next three
            # counts (a run can be delayed on the host while the generator
            # processes work)
            runs = []
            for _ in range(1 + TIMED_RUNS):
                ctx.fill("X", 0.0)
                runs.append(ctx.solve_mixed(precond=precond, max_iter=100000, tol=1e-9))
            best = min(runs[1:], key=lambda r: r["wall_s"])
            return best["wall_s"] * 1e3 / max(best["iterations"], 1), best["iterations"]

        tj, _ = per_iter("jacobi")
        tp, ip = [], []
        for deg in learned.DEGREES:
            ctx.set_poly(learned.fixed_rule_coeffs(theta10, b, 16.0, deg), b)
            t_, i_ = per_iter("poly")
            tp.append(t_)
            ip.append(i_)
        rec["jacobi_ms_per_iter"].append(tj)
        rec["poly_ms_per_iter"].append(tp)
        rec["poly_iterations"].append(ip)
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
    pool.close()
    pool.join()
    out = {k: np.array(v) for k, v in rec.items()}
    out["alpha"] = np.concatenate(alphas)
    out["beta"] = np.concatenate(betas)
    out["alpha_offset"] = np.concatenate([[0], np.cumsum([a.size for a in alphas])])
    out["beta_offset"] = np.concatenate([[0], np.cumsum([b_.size for b_ in betas])])
    out["tol"] = HARVEST_TOL
    out["degrees"] = np.array(learned.DEGREES)
    np.savez_compressed(args.out, **out)
    print(f"[harvest] wrote {args.out}: {len(rec['seed'])} systems in {time.perf_counter() - t0:.0f} s",
          flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
