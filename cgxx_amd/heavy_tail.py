"""Heavy-tailed random SPD system: CRS against ELL under plain Jacobi-PCG.

    python -m cgxx_amd.heavy_tail [--n 1000000] [--cap 2000] [--steps 3]

The system of `problems.zipf_spd` (symmetric pattern, off-diagonal row
counts Zipf(2) capped at `cap`, values U(-1, 0), diagonal 1.05 |row sum|,
right-hand side U(-1, 1), seed 20261202).  Both FP64 formats are uploaded
(ELL: width = the longest row, about 2,000 slots per row, 24 GB at n = 1M)
and solved by the fused FP64 Jacobi-PCG loop (CG.cpp:284-385) from x = 0 to
relative residual 1e-9, then by the mixed-precision solver with Jacobi.
CRS and ELL add every row left to right from 0, so their SpMVs are
bit-identical; their fused dot products fold the same row products along
different trees (CRS row blocks end early where rows are long, ELL chunks
always hold 256 rows), so the iterates agree to rounding.  The run checks
that both reach the tolerance within one iteration of each other, reports
the largest relative difference between the two solutions, and the time
each format takes for the same work, i.e. the cost of the load imbalance.
Prints one JSON object (stdout); progress on stderr; exits 1 if the two
formats disagree.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from cgxx_amd import problems
from cgxx_amd.cghip import Context

HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[heavy_tail] {msg}", file=sys.stderr, flush=True)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--cap", type=int, default=2000)
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--steps", type=int, default=3)
    args = ap.parse_args(argv)

    t0 = time.perf_counter()
    m = problems.zipf_spd(args.n, cap=args.cap, seed=args.seed)
    n, nnz = m.n, m.nnz
    lens = m.row_lengths()
    log(f"generated: n={n} nnz={nnz} mean row {lens.mean():.2f} max row {lens.max()} "
        f"in {time.perf_counter() - t0:.1f} s")
    width, length, eidx, edat = m.to_ell()
    log(f"ELL width {width}: {width * n} slots ({12 * width * n / 1e9:.1f} GB)")

    ctx = Context(0)
    t0 = time.perf_counter()
    ctx.upload_crs(m.ptr, m.index, m.value)
    ctx.upload_ell(width, length, eidx, edat)
    del eidx, edat
    ctx.upload_jacobi(1.0 / m.diagonal())
    ctx.upload("K", m.rhs)
    log(f"uploaded CRS + ELL + Jacobi + K in {time.perf_counter() - t0:.1f} s")

    def per(info, key):
        return info[f"ms_{key}"] / max(info[f"n_{key}"], 1)

    spmv_bytes = {"CRS": 12 * nnz + 4 * (n + 1) + 16 * n, "ELL": 12 * nnz + 4 * n + 16 * n}
    out = {"system": dict(m.info), "nnz": int(nnz), "ell_width": int(width),
           "rows_longer_than_256": int((lens > 256).sum()), "formats": {}}
    xs = {}
    for fmt in ("CRS", "ELL"):
        ctx.set_format(fmt)

        def solve(profile=0):
            ctx.fill("X", 0.0)
            return ctx.solve(precond="jacobi", max_iter=100000, tol=1e-9, profile=profile)

        solve()  # warm-up
        t0 = time.perf_counter()
        infos = [solve() for _ in range(args.steps)]
        sec = (time.perf_counter() - t0) / args.steps
        nr, nb = ctx.true_residual("X", "K")
        xs[fmt] = ctx.download("X")
        prof = solve(profile=1)
        spmv_ms = per(prof, "matvec")
        plain_ms = ctx.time_matvec("P", "Q", reps=10)
        by = spmv_bytes[fmt]
        out["formats"][fmt] = {
            "seconds": sec, "iterations": infos[-1]["iterations"], "residual": infos[-1]["residual"],
            "true_relative_residual": nr / nb,
            "spmv_dot_ms": spmv_ms, "spmv_plain_ms": plain_ms,
            "spmv_plain_gbps": by / (plain_ms * 1e-3) / 1e9,
            "spmv_plain_hbm_frac": by / (plain_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "update_ms": per(prof, "update"), "xpay_ms": per(prof, "xpay"),
            "algorithmic_spmv_bytes": by,
        }
        log(f"FP64 {fmt}: {infos[-1]['iterations']} iterations, {sec:.4f} s, true residual "
            f"{nr / nb:.3e}; SpMV {plain_ms:.3f} ms = {by / (plain_ms * 1e-3) / 1e9:.0f} GB/s")
    fc, fe = out["formats"]["CRS"], out["formats"]["ELL"]
    diff = float(np.max(np.abs(xs["CRS"] - xs["ELL"])) / np.max(np.abs(xs["CRS"])))
    same = (abs(fc["iterations"] - fe["iterations"]) <= 1 and fc["residual"] <= 1e-9
            and fe["residual"] <= 1e-9 and diff <= 1e-7)
    out["crs_ell_max_rel_diff"] = diff
    out["crs_ell_agree"] = bool(same)
    out["ell_over_crs_time"] = out["formats"]["ELL"]["seconds"] / out["formats"]["CRS"]["seconds"]

    ctx.set_format("CRS")
    ctx.prepare_fp32()

    def mixed(profile=0):
        ctx.fill("X", 0.0)
        return ctx.solve_mixed(precond="jacobi", max_iter=100000, tol=1e-9, profile=profile)

    mixed()
    t0 = time.perf_counter()
    infos = [mixed() for _ in range(args.steps)]
    sec = (time.perf_counter() - t0) / args.steps
    out["mixed_jacobi"] = {"layout": ctx.fp32_layout(), "seconds": sec,
                           "iterations": infos[-1]["iterations"],
                           "reliable_updates": infos[-1]["updates"],
                           "true_relative_residual": infos[-1]["true_residual"],
                           "speedup_vs_fp64_crs": out["formats"]["CRS"]["seconds"] / sec}
    log(f"mixed Jacobi ({ctx.fp32_layout()}): {infos[-1]['iterations']} iterations, {sec:.4f} s, "
        f"true residual {infos[-1]['true_residual']:.3e}")
    ctx.close()
    print(json.dumps(out), flush=True)
    return 0 if same else 1


if __name__ == "__main__":
    sys.exit(main())
