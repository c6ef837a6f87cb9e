#!/usr/bin/env python3
"""Time-to-solution benchmark of the CG hot path on one MI355X.

Workload: 3D 7-point variable-coefficient diffusion on a 256^3 grid
(N = 16,777,216, nnz = 117,047,296), log-normal cell coefficients
(sigma = 1.0, correlation length 8 cells, seed 20261202), harmonic-mean face
weights, k = A*1, x0 = 0, stop at relative residual 1e-9.  One step = one
complete solve from x = 0 with the matrix and vectors already resident in HBM.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--side 256]

Prints one JSON line.  `value` is the mean time-to-solution (s) of the
measured method: FP32-inner PCG with FP64 reliable updates and the Jacobi
preconditioner, stopping on the FP64 true residual.  The plain FP64
Jacobi-PCG (the reference's method, as the fused device loop) is timed on
the same GPU beside it.  Kernel durations come from one more solve of each
method run kernel by kernel with HIP events on the library's stream;
`roofline` is the measured method's dominant kernel (the FP32 SpMV + <p,q>).
`cpu_baseline` is the oracle's serial port of the reference's SerialCG on this
host over a bounded number of iterations, scaled to the FP64 iteration count.
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


def spmv_bytes(n: int, nnz: int, value_bytes: int, vec_bytes: int) -> int:
    """Algorithmic bytes of one CRS SpMV: values, indices, row pointer, x, y."""
    return value_bytes * nnz + 4 * nnz + 4 * (n + 1) + 2 * vec_bytes * n


def traffic_from_profiles(name: str):
    """HBM bytes per launch of a kernel from the committed PMC summary."""
    path = os.path.join(REPO, "profiles", f"traffic_{name}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("bytes_per_launch")


def cpu_baseline(m, k, c, iterations: int, sample_iters: int) -> dict:
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
        "sample": (f"serial port of SerialCG (oracle/cg_oracle.c, order 0), FP64 Jacobi-PCG on this "
                   f"workload: {sample_iters} iterations in {dt:.2f} s, scaled to the {iterations} "
                   f"iterations of the FP64 solve"),
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--side", type=int, default=256)
    ap.add_argument("--cpu-iters", type=int, default=12)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--degree", type=int, default=8, help="polynomial preconditioner degree")
    ap.add_argument("--ritz-div", type=float, default=8.0,
                    help="interval lower end = smallest 10-step Ritz value / this")
    ap.add_argument("--pmc", action="store_true",
                    help="short run for rocprofv3 --pmc: one D2D copy of a vector (the "
                         "calibration kernel of known size), 20 FP64 and 20 mixed iterations")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path "
              f"(--gpus {args.gpus} requested)", file=sys.stderr)
        return 2

    import numpy as np

    from cgxx_amd import problems, spectrum
    from cgxx_amd.cghip import Context

    t = time.perf_counter()
    m = problems.diffusion3d(args.side, sigma=1.0, corr=8.0, seed=20261202)
    k = m.ones_rhs()
    c = 1.0 / m.diagonal()
    n, nnz = m.n, m.nnz
    log(f"generated {args.side}^3 diffusion: N={n} nnz={nnz} in {time.perf_counter() - t:.1f} s")

    ctx = Context(0)
    ctx.upload_crs(m.ptr, m.index, m.value)
    ctx.set_format("CRS")
    ctx.upload_jacobi(c)
    ctx.prepare_fp32()
    ctx.upload("K", k)
    log("uploaded (FP64 CRS, FP32 copy of the values, Jacobi)")

    max_iter = 100000
    if args.pmc:
        ctx.fill("P", 1.0)
        ctx.cpy("R", "P")  # reads and writes exactly 8*N bytes each
        ctx.fill("X", 0.0)
        info = ctx.solve(precond="jacobi", max_iter=20, tol=1e-9, batch=1)
        ctx.fill("X", 0.0)
        info2 = ctx.solve_mixed(precond="jacobi", max_iter=20, tol=1e-9, batch=1)
        log(f"pmc run: copy of {8 * n} bytes, {info['iterations']} FP64 and "
            f"{info2['iterations']} mixed iterations")
        ctx.close()
        return 0

    def fp64(profile=0):
        ctx.fill("X", 0.0)
        return ctx.solve(precond="jacobi", max_iter=max_iter, tol=1e-9, profile=profile)

    def mixed(profile=0):
        ctx.fill("X", 0.0)
        return ctx.solve_mixed(precond="jacobi", max_iter=max_iter, tol=1e-9, profile=profile)

    def timed(fn, label):
        for i in range(args.warmup):
            info = fn()
            log(f"{label} warmup {i}: {info['iterations']} iterations, {info['wall_s']:.4f} s")
        t0 = time.perf_counter()
        infos = [fn() for _ in range(args.steps)]
        sec = (time.perf_counter() - t0) / args.steps
        log(f"{label}: {args.steps} timed solves, {sec:.4f} s each, {infos[-1]['iterations']} iterations")
        return sec, infos[-1]

    # Plain FP64 Jacobi-PCG (the reference's method) on the same GPU.
    sec64, info64 = timed(fp64, "fp64 jacobi")
    nr, nb = ctx.true_residual("X", "K")
    x64 = ctx.download("X")
    prof64 = fp64(profile=1)
    spmv64_ms = prof64["ms_matvec"] / max(prof64["n_matvec"], 1)
    spmv64_bytes = spmv_bytes(n, nnz, 8, 8)
    spmv64_gbps = spmv64_bytes / (spmv64_ms * 1e-3) / 1e9
    log(f"fp64: true residual {nr / nb:.3e}; spmv+dot {spmv64_ms:.4f} ms = {spmv64_gbps:.0f} GB/s, "
        f"update {prof64['ms_update'] / prof64['n_update']:.4f} ms, "
        f"p update {prof64['ms_xpay'] / prof64['n_xpay']:.4f} ms")

    # Measured method: FP32 inner PCG with FP64 reliable updates.
    secmp, infomp = timed(mixed, "mixed")
    nrm, nbm = ctx.true_residual("X", "K")
    xmp = ctx.download("X")
    profmp = mixed(profile=1)
    spmv32_ms = profmp["ms_matvec"] / max(profmp["n_matvec"], 1)
    spmv32_bytes = spmv_bytes(n, nnz, 4, 4)
    spmv32_gbps = spmv32_bytes / (spmv32_ms * 1e-3) / 1e9
    upd32_ms = profmp["ms_update"] / max(profmp["n_update"], 1)
    xpay32_ms = profmp["ms_xpay"] / max(profmp["n_xpay"], 1)
    refresh_ms = profmp["ms_refresh"] / max(profmp["n_refresh"], 1)
    log(f"mixed: true residual {nrm / nbm:.3e} ({infomp['updates']} reliable updates); "
        f"fp32 spmv+dot {spmv32_ms:.4f} ms = {spmv32_gbps:.0f} GB/s, update {upd32_ms:.4f} ms, "
        f"p update {xpay32_ms:.4f} ms, reliable update {refresh_ms:.4f} ms")
    err = float(np.abs(xmp - 1.0).max())
    diff = float(np.abs(xmp - x64).max())

    # Mixed precision with a degree-8 Chebyshev polynomial preconditioner on
    # [a, b]: b = Gershgorin bound of D^-1 A, a from the smallest Ritz value
    # of 10 FP64 Jacobi-PCG steps (placed below it: the 10-step estimate
    # lies above lambda_min).
    ctx.fill("X", 0.0)
    ctx.solve(precond="jacobi", max_iter=10, tol=1e-300)
    theta = spectrum.ritz_values(ctx.history("alpha", 10), ctx.history("beta", 9))
    bnd = spectrum.gershgorin_bound(m.ptr, m.index, m.value)
    poly_a = theta[0] / args.ritz_div
    coeffs = spectrum.cheb_residual_coeffs(poly_a, bnd, args.degree)
    ctx.set_poly(coeffs, bnd)

    def mixed_poly(profile=0):
        ctx.fill("X", 0.0)
        return ctx.solve_mixed(precond="poly", max_iter=max_iter, tol=1e-9, profile=profile)

    secpl, infopl = timed(mixed_poly, "mixed+poly")
    nrp, nbp = ctx.true_residual("X", "K")
    xpl = ctx.download("X")
    profpl = mixed_poly(profile=1)
    log(f"mixed+poly: degree {args.degree} on [{poly_a:.3e}, {bnd:.3f}] (Ritz {theta[0]:.3e}.."
        f"{theta[-1]:.3f}): {infopl['iterations']} iterations ({infopl['iterations'] * (args.degree + 1)} "
        f"FP32 SpMVs), true residual {nrp / nbp:.3e}, {secpl:.4f} s")
    plain64_ms = ctx.time_matvec("P", "Q", reps=20)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(m, k, c, info64["iterations"], args.cpu_iters)
        log(f"cpu baseline: {cpu['value']:.1f} s ({cpu['sample']})")

    out = {
        "metric": METRIC,
        "value": secmp,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": secmp * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": f"3D 7-point variable-coefficient diffusion {args.side}^3 "
                        "(log-normal kappa, sigma=1, corr=8, seed 20261202), k=A*1, x0=0, "
                        "FP64 true relative residual 1e-9",
            "method": "FP32-inner Jacobi-PCG with FP64 reliable updates (delta 0.1), fused "
                      "device loop; learned polynomial preconditioner not in this measurement",
            "n": n,
            "nnz": nnz,
            "format": "CRS",
        },
        "iterations": infomp["iterations"],
        "reliable_updates": infomp["updates"],
        "true_relative_residual": nrm / nbm,
        "max_abs_error_vs_ones": err,
        "max_abs_diff_vs_fp64_solution": diff,
        "fp64_jacobi": {
            "seconds": sec64,
            "iterations": info64["iterations"],
            "true_relative_residual": nr / nb,
            "spmv_dot_ms": spmv64_ms,
            "spmv_gbps": spmv64_gbps,
            "spmv_hbm_frac": spmv64_gbps / HBM_PEAK_GBS,
            "spmv_plain_ms": plain64_ms,
            "spmv_plain_gbps": spmv64_bytes / (plain64_ms * 1e-3) / 1e9,
            "update_ms": prof64["ms_update"] / max(prof64["n_update"], 1),
            "xpay_ms": prof64["ms_xpay"] / max(prof64["n_xpay"], 1),
            "traffic_spmv_dot": traffic_from_profiles("spmv_dot_crs_fp64_256"),
        },
        "mixed_poly": {
            "seconds": secpl,
            "degree": args.degree,
            "interval": [poly_a, bnd],
            "ritz_10": [float(theta[0]), float(theta[-1])],
            "iterations": infopl["iterations"],
            "fp32_spmvs": infopl["iterations"] * (args.degree + 1),
            "reliable_updates": infopl["updates"],
            "true_relative_residual": nrp / nbp,
            "max_abs_error_vs_ones": float(np.abs(xpl - 1.0).max()),
            "max_abs_diff_vs_fp64_solution": float(np.abs(xpl - x64).max()),
            "spmv_dot_ms": profpl["ms_matvec"] / max(profpl["n_matvec"], 1),
            "update_ms": profpl["ms_update"] / max(profpl["n_update"], 1),
            "clenshaw_and_xpay_ms": profpl["ms_xpay"] / max(profpl["n_xpay"], 1),
            "speedup_vs_fp64_jacobi": sec64 / secpl,
        },
        "speedup_vs_fp64_jacobi": sec64 / secmp,
        "spmv_gbps": spmv32_gbps,
        "kernel_ms": {"spmv_dot_fp32": spmv32_ms, "update_fp32": upd32_ms, "xpay_fp32": xpay32_ms,
                      "reliable_update": refresh_ms},
        "kernel_gbps": {
            "spmv_dot_fp32": spmv32_gbps,
            "update_fp32": 28 * n / (upd32_ms * 1e-3) / 1e9,
            "xpay_fp32": 16 * n / (xpay32_ms * 1e-3) / 1e9,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": spmv32_gbps,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": spmv32_gbps / HBM_PEAK_GBS,
            "traffic": traffic_from_profiles("spmv_dot_crs_fp32_256"),
            "kernel": "k_mp_matvec_dot_crs",
            "algorithmic_bytes_per_launch": spmv32_bytes,
        },
        "cpu_baseline": cpu,
    }
    ctx.close()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
