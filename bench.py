#!/usr/bin/env python3
"""Time-to-solution benchmark of the CG hot path on one MI355X.

Workload: 3D 7-point variable-coefficient diffusion on a 256^3 grid
(N = 16,777,216, nnz = 117,047,296), log-normal cell coefficients
(sigma = 1.0, correlation length 8 cells, seed 20261202), harmonic-mean face
weights, k = A*1, x0 = 0, stop at relative residual 1e-9.  One step = one
complete solve from x = 0 with the matrix and vectors already resident in HBM.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--side 256]

Prints one JSON line.  `value` is the mean time-to-solution (s) of the
measured method; the plain FP64 Jacobi-CG time on the same GPU is reported
beside it.  `roofline` is the fused SpMV+<p,q> kernel's algorithmic HBM
bytes over its average launch duration, timed with HIP events on the
library's stream in a profiled solve inside this run; `cpu_baseline` is the
oracle's serial port of the reference's SerialCG timed on this host over a
bounded number of iterations and scaled to the full iteration count.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "time-to-1e-9 residual (s), SpMV GB/s: mixed+learned vs FP64 Jacobi CG"
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def spmv_bytes(n: int, nnz: int, value_bytes: int = 8, vec_bytes: int = 8) -> int:
    """Algorithmic bytes of one CRS SpMV: values, indices, row pointer, x, y."""
    return value_bytes * nnz + 4 * nnz + 4 * (n + 1) + 2 * vec_bytes * n


def traffic_from_profiles(side: int):
    """HBM bytes per launch of the fused SpMV kernel from committed PMC runs."""
    path = os.path.join(REPO, "profiles", f"traffic_spmv_dot_crs_fp64_{side}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("bytes_per_launch")


def cpu_baseline(m, k, c, iterations: int, sample_iters: int) -> dict:
    import numpy as np

    from oracle import oracle

    t0 = time.perf_counter()
    oracle.pcg(n=m.n, fmt="CRS", precond="jacobi", order=0, k=k, tol=1e-300,
               max_iter=sample_iters, crs=(m.ptr, m.index, m.value), C_jac=c, history=False)
    dt = time.perf_counter() - t0
    per_iter = dt / (sample_iters + 1)  # + the set-up pass (one SpMV, 2 dots)
    return {
        "value": per_iter * (iterations + 1),
        "unit": "s",
        "cores": 1,
        "kind": "port",
        "sample": (f"serial port of SerialCG (oracle/cg_oracle.c, order 0) on this workload: "
                   f"{sample_iters} PCG iterations in {dt:.2f} s, scaled to {iterations} iterations"),
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--side", type=int, default=256)
    ap.add_argument("--cpu-iters", type=int, default=12)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--pmc", action="store_true",
                    help="short run for rocprofv3 --pmc: one D2D copy of a vector (the "
                         "calibration kernel of known size) and a 20-iteration solve")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path "
              f"(--gpus {args.gpus} requested)", file=sys.stderr)
        return 2

    import numpy as np

    from cgxx_amd import problems
    from cgxx_amd.cghip import Context

    t = time.perf_counter()
    m = problems.diffusion3d(args.side, sigma=1.0, corr=8.0, seed=20261202)
    k = m.ones_rhs()
    c = 1.0 / m.diagonal()
    log(f"generated {args.side}^3 diffusion: N={m.n} nnz={m.nnz} in {time.perf_counter() - t:.1f} s")

    ctx = Context(0)
    ctx.upload_crs(m.ptr, m.index, m.value)
    ctx.set_format("CRS")
    ctx.upload_jacobi(c)
    ctx.upload("K", k)
    log("uploaded")

    max_iter = 100000
    if args.pmc:
        ctx.fill("P", 1.0)
        ctx.cpy("R", "P")  # reads and writes exactly 8*N bytes each
        ctx.fill("X", 0.0)
        info = ctx.solve(precond="jacobi", max_iter=20, tol=1e-9, batch=1)
        log(f"pmc run: copy of {8 * m.n} bytes, {info['iterations']} iterations")
        ctx.close()
        return 0

    def solve(profile=0):
        ctx.fill("X", 0.0)
        return ctx.solve(precond="jacobi", max_iter=max_iter, tol=1e-9, profile=profile)

    for i in range(args.warmup):
        info = solve()
        log(f"warmup {i}: {info['iterations']} iterations, {info['wall_s']:.3f} s")

    t0 = time.perf_counter()
    infos = [solve() for _ in range(args.steps)]
    t1 = time.perf_counter()
    sec = (t1 - t0) / args.steps
    iters = infos[-1]["iterations"]
    log(f"timed {args.steps} solves: {sec:.4f} s each, {iters} iterations")

    nr, nb = ctx.true_residual("X", "K")
    x = ctx.download("X")
    err = float(np.abs(x - 1.0).max())
    log(f"true relative residual {nr / nb:.3e}, max |x - 1| = {err:.3e}")

    # Kernel durations: the same solve once more, kernel by kernel, each
    # kernel between two HIP events on the library's stream.
    prof = solve(profile=1)
    spmv_ms = prof["ms_matvec"] / max(prof["n_matvec"], 1)
    upd_ms = prof["ms_update"] / max(prof["n_update"], 1)
    xpay_ms = prof["ms_xpay"] / max(prof["n_xpay"], 1)
    plain_ms = ctx.time_matvec("P", "Q", reps=20)
    nbytes = spmv_bytes(m.n, m.nnz)
    achieved = nbytes / (spmv_ms * 1e-3) / 1e9
    log(f"profiled: spmv+dot {spmv_ms:.4f} ms ({achieved:.0f} GB/s), update {upd_ms:.4f} ms, "
        f"xpay {xpay_ms:.4f} ms, plain spmv {plain_ms:.4f} ms")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(m, k, c, iters, args.cpu_iters)
        log(f"cpu baseline: {cpu['value']:.1f} s ({cpu['sample']})")

    n = m.n
    upd_bytes = 56 * n  # x, r, p, q, C read; x, r written
    xpay_bytes = 32 * n  # C, r, p read; p written
    out = {
        "metric": METRIC,
        "value": sec,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"3D 7-point variable-coefficient diffusion {args.side}^3 "
                        "(log-normal kappa, sigma=1, corr=8, seed 20261202), k=A*1, x0=0, tol 1e-9",
            "method": "fp64 jacobi-pcg, fused device loop",
            "n": n,
            "nnz": m.nnz,
            "format": "CRS",
        },
        "iterations": iters,
        "true_relative_residual": nr / nb,
        "max_abs_error_vs_ones": err,
        "spmv_gbps": achieved,
        "kernel_ms": {"spmv_dot": spmv_ms, "update": upd_ms, "xpay": xpay_ms, "spmv_plain": plain_ms},
        "kernel_gbps": {
            "spmv_dot": achieved,
            "update": upd_bytes / (upd_ms * 1e-3) / 1e9,
            "xpay": xpay_bytes / (xpay_ms * 1e-3) / 1e9,
            "spmv_plain": nbytes / (plain_ms * 1e-3) / 1e9,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic_from_profiles(args.side),
            "kernel": "k_pcg_matvec_dot_crs",
            "algorithmic_bytes_per_launch": nbytes,
        },
        "cpu_baseline": cpu,
    }
    ctx.close()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
