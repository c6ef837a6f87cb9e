#!/usr/bin/env python3
"""Time-to-solution benchmark of the CG hot path on one MI355X.

Workload: 3D 7-point variable-coefficient diffusion on a 256^3 grid
(N = 16,777,216, nnz = 117,047,296), log-normal cell coefficients
(sigma = 1.0, correlation length 8 cells, seed 20261202), harmonic-mean face
weights, k = A*1, x0 = 0, stop at relative residual 1e-9.  One step = one
complete solve from x = 0 with the matrix and vectors already resident in HBM.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--side 256]

Prints one JSON line.  `value` is the mean time-to-solution (s) of the
measured method, mixed precision with the learned polynomial preconditioner:
device matrix statistics and 10 FP64 Jacobi-PCG steps for the features, the
network on the GPU for the coefficients, then the FP32-inner CG on the
scaled operator with the degree-8 polynomial and FP64 reliable updates,
stopping on the FP64 true relative residual (Jacobi fallback if the
polynomial is refused or misses its SpMV budget); all of it inside the
timed step.  Beside it, on the same GPU: the plain FP64 Jacobi-PCG (the
reference's method, CG.cpp:284-385, as the fused device loop) in CRS and in
ELL, mixed precision with Jacobi, and mixed precision with the polynomial
from the fixed rule that the network replaces.  Kernel durations come from
one more solve of each method run kernel by kernel with HIP events on the
library's stream; `roofline` is the measured method's dominant kernel, the
Clenshaw SpMV step on the FP32 scaled operator.  `cpu_baseline` is the
oracle's serial port of the reference's SerialCG on this host over a bounded
number of iterations, scaled to the FP64 iteration count.
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


def crs_bytes(n: int, nnz: int, value_bytes: int, vec_bytes: int) -> int:
    """Algorithmic bytes of one CRS SpMV: values, indices, row pointer, x, y."""
    return value_bytes * nnz + 4 * nnz + 4 * (n + 1) + 2 * vec_bytes * n


def ell_bytes(n: int, nnz: int, value_bytes: int, vec_bytes: int) -> int:
    """Algorithmic bytes of one ELL SpMV: live slots, lengths, x, y."""
    return value_bytes * nnz + 4 * nnz + 4 * n + 2 * vec_bytes * n


def traffic_from_profiles(prefix: str):
    """HBM bytes per launch of the first kernel whose name starts with
    prefix, from the committed PMC summary."""
    path = os.path.join(REPO, "profiles", "traffic_c4.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        kernels = json.load(f).get("kernels", {})
    for name, k in kernels.items():
        if name.startswith(prefix):
            return k.get("bytes_per_launch")
    return None


def kernel_short_name(name: str) -> str:
    """'void cghip::(anonymous namespace)::k_x<a, b>(args...)' -> 'k_x<a, b>'."""
    name = name.strip()
    if name.startswith("void "):
        name = name[5:]
    name = name.replace("cghip::(anonymous namespace)::", "").replace("cghip::", "")
    depth = 0
    for i, ch in enumerate(name):
        depth += ch == "<"
        depth -= ch == ">"
        if ch == "(" and depth == 0:
            return name[:i]
    return name


def algorithmic_bytes(short: str, meta: dict):
    """Bytes a launch must move at least (each array once, x once per SpMV)."""
    n, nnz, D = meta["n"], meta["nnz"], meta["ndiag"]
    # the Clenshaw steps' FP16 values: the row records, or with the symmetric
    # storage (template arguments FINAL, K, MODE, HALF, NT, ROWS, SYM; SYM
    # true) the upper diagonals only, 2 bytes per stored pair
    args = short[short.find("<") + 1:-1].split(", ") if short.startswith("k_poly_step<") else []
    rec = D * n if len(args) >= 7 and args[6] == "true" else 4 * ((D + 1) // 2) * n
    table = [("k_pcg_matvec_dot_crs", crs_bytes(n, nnz, 8, 8)), ("k_matvec_crs<double", crs_bytes(n, nnz, 8, 8)),
             ("k_pcg_matvec_dot_ell", ell_bytes(n, nnz, 8, 8)), ("k_matvec_ell<double", ell_bytes(n, nnz, 8, 8)),
             ("k_pcg_update<true>", 32 * n), ("k_pcg_update<false>", 24 * n), ("k_pcg_xpay<true>", 48 * n),
             ("k_pcg_xpay<false>", 40 * n), ("k_mp_matvec_dot<2>", 4 * D * n + 8 * n),
             ("k_mp_update", 12 * n), ("k_mp_xpay_poly", 20 * n), ("k_mp_xpay", 20 * n),
             ("k_poly_step<false, 2, 1,", rec + 8 * n),
             ("k_poly_step<false, 2, 2,", rec + 12 * n),
             ("k_poly_step<false, 2, 0,", rec + 16 * n),
             ("k_poly_step<true, 2, 0,", rec + 16 * n)]
    for prefix, by in table:
        if short.startswith(prefix):
            return by
    return None


def summarize_trace(trace_dir: str, bench_json: str, out_stem: str) -> int:
    """profiles/<out_stem>.md from a rocprofv3 --kernel-trace --stats run of
    bench.py (csv output): per kernel the launches and average, min and max
    duration, and for the kernels with a byte model the algorithmic GB/s and
    fraction of the 8 TB/s HBM peak (sizes from bench_json, the bench line
    of a run on the same workload)."""
    import csv
    import glob

    with open(bench_json) as f:
        line = [ln for ln in f if ln.strip().startswith("{")][-1]
    cfg = json.loads(line)["config"]
    fmt = cfg.get("format", "")
    ndiag = int(fmt.split("(")[1].split()[0]) if "diagonals" in fmt else 0
    meta = {"n": cfg["n"], "nnz": cfg["nnz"], "ndiag": ndiag}
    rows = []
    for path in glob.glob(os.path.join(trace_dir, "**", "*kernel_stats.csv"), recursive=True):
        with open(path) as f:
            for row in csv.DictReader(f):
                rows.append(row)
    if not rows:
        raise SystemExit("no kernel_stats.csv under " + trace_dir)
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    lines = []
    for r in rows:
        name = kernel_short_name(r["Name"])
        avg = float(r["AverageNs"]) / 1e3
        alg = algorithmic_bytes(name, meta)
        rate = f"{alg / (avg * 1e-6) / 1e9:.0f} | {alg / (avg * 1e-6) / 1e9 / HBM_PEAK_GBS:.3f}" \
            if alg else "— | —"
        lines.append(f"| {name} | {r['Calls']} | {avg:.1f} | {float(r['MinNs']) / 1e3:.1f} | "
                     f"{float(r['MaxNs']) / 1e3:.1f} | {float(r['Percentage']):.1f} | {rate} |")
    out = os.path.join(REPO, "profiles", out_stem + ".md")
    with open(out, "w") as f:
        f.write(f"# Kernel trace: {cfg['workload']}\n\n"
                "Commands (from the repository root):\n\n"
                "    rocprofv3 --kernel-trace --stats --output-format csv -d gpurun_out/ktrace -o kt -- "
                "python3 bench.py --trace > gpurun_out/trace.json\n"
                f"    python3 bench.py --summarize-trace gpurun_out/ktrace gpurun_out/trace.json {out_stem}\n\n"
                f"N = {meta['n']:,}, nnz = {meta['nnz']:,}, {ndiag} diagonals in the scaled FP32 operator. "
                "`bench.py --trace` launches the bench's kernels directly, not from HIP graphs: 100 FP64 "
                "Jacobi-PCG iterations in CRS and in ELL, 20 plain SpMVs each, 100 mixed Jacobi and 20 "
                "mixed polynomial iterations (the network's degree), kernel by kernel, and 20 "
                "back-to-back polynomial applications. Durations in microseconds over every launch; "
                "GB/s = algorithmic bytes per launch / average duration.\n\n"
                "| kernel | launches | avg us | min us | max us | % of GPU time | algorithmic GB/s | "
                "of 8 TB/s |\n|---|---|---|---|---|---|---|---|\n" + "\n".join(lines) + "\n")
    print(f"wrote {out} ({len(rows)} kernels)")
    return 0


def summarize_pmc(fetch_dir: str, write_dir: str, meta_path: str, out_stem: str) -> int:
    """profiles/<out_stem>.json and .md from two rocprofv3 --pmc runs of
    `bench.py --pmc` (FETCH_SIZE, then WRITE_SIZE; csv output): per kernel
    the median counter per launch in KiB, scaled by what the counters read
    for the D2D copy of 8N bytes (__amd_rocclr_copyBuffer), against the
    kernel's algorithmic bytes."""
    import csv
    import glob
    import statistics

    with open(meta_path) as f:
        meta = json.load(f)

    def load(d, counter):
        vals = {}
        for path in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            with open(path) as f:
                for row in csv.DictReader(f):
                    if row.get("Counter_Name") != counter:
                        continue
                    vals.setdefault(kernel_short_name(row["Kernel_Name"]), []).append(
                        float(row["Counter_Value"]))
        return vals

    fetch, write = load(fetch_dir, "FETCH_SIZE"), load(write_dir, "WRITE_SIZE")
    copy = [k for k in fetch if "copyBuffer" in k]
    if not copy:
        raise SystemExit("no copy kernel in the FETCH_SIZE run")
    # the runtime's copy kernel also moves small buffers: the 8N copy is its
    # largest launch
    ref = 8 * meta["n"]
    copy_f, copy_w = max(fetch[copy[0]]), max(write[copy[0]])
    cal_f = copy_f * 1024 / ref
    cal_w = copy_w * 1024 / ref
    kernels, rows = {}, []
    for name in fetch:
        if name not in write or "copyBuffer" in name:
            continue
        alg = algorithmic_bytes(name, meta)
        if alg is None:
            continue
        fk, wk = statistics.median(fetch[name]), statistics.median(write[name])
        rd, wr = fk * 1024 / cal_f, wk * 1024 / cal_w
        kernels[name] = {"bytes_read": rd, "bytes_written": wr, "bytes_per_launch": rd + wr,
                         "algorithmic": alg, "launches": len(fetch[name])}
        rows.append(f"| {name} | {len(fetch[name])} | {fk:.1f} | {wk:.1f} | {rd:.4g} | {wr:.4g} | "
                    f"{(rd + wr) / alg:.3f} |")
    out = os.path.join(REPO, "profiles", out_stem)
    with open(out + ".json", "w") as f:
        json.dump({"source": f"profiles/{out_stem}.md", "calibration_fetch": cal_f,
                   "calibration_write": cal_w, "meta": meta, "kernels": kernels}, f, indent=1)
    with open(out + ".md", "w") as f:
        f.write(f"# HBM traffic per kernel (PMC): {meta['side']}^3 diffusion, bench kernels\n\n"
                "Commands (from the repository root; one counter per pass):\n\n"
                "    rocprofv3 --pmc FETCH_SIZE --output-format csv -d gpurun_out/pmc_fetch -o f -- "
                "python3 bench.py --pmc\n"
                "    rocprofv3 --pmc WRITE_SIZE --output-format csv -d gpurun_out/pmc_write -o w -- "
                "python3 bench.py --pmc\n"
                f"    python3 bench.py --summarize-pmc gpurun_out/pmc_fetch gpurun_out/pmc_write "
                f"gpurun_out/pmc_meta.json {out_stem}\n\n"
                f"Calibration: the D2D copy of 8N = {ref:,} bytes (the largest launch of the "
                f"runtime's copy kernel) reads FETCH_SIZE = {copy_f:.0f} KiB and WRITE_SIZE = "
                f"{copy_w:.0f} KiB, so the counters are scaled by "
                f"1/{cal_f:.4f} and 1/{cal_w:.4f}. Medians over the launches. N = {meta['n']:,}, "
                f"nnz = {meta['nnz']:,}, {meta['ndiag']} diagonals in the scaled FP32 operator "
                f"({meta['layout']}).\n\n"
                "| kernel | launches | FETCH_SIZE KiB | WRITE_SIZE KiB | bytes read | bytes written | "
                "traffic / algorithmic |\n|---|---|---|---|---|---|---|\n" + "\n".join(rows) + "\n")
    print(f"wrote {out}.json and .md ({len(kernels)} kernels)")
    return 0


def application_traffic(kind: int, degree: int):
    """PMC bytes per Clenshaw step averaged over one application: the first
    step (template arguments <false, kind, 1, ...>), the second (2), the
    plain ones (0) and the last (final, 0)."""
    first = traffic_from_profiles(f"k_poly_step<false, {kind}, 1,")
    second = traffic_from_profiles(f"k_poly_step<false, {kind}, 2,")
    plain = traffic_from_profiles(f"k_poly_step<false, {kind}, 0,")
    last = traffic_from_profiles(f"k_poly_step<true, {kind}, 0,")
    if degree < 3 or None in (first, second, plain, last):
        return None
    return (first + second + (degree - 3) * plain + last) / degree


def cpu_threads() -> int:
    """Threads for the OpenMP port: the CPU share this process is given
    (OMP_NUM_THREADS when set, as on a leased GPU machine), else every CPU
    it may run on."""
    try:
        return max(1, int(os.environ["OMP_NUM_THREADS"]))
    except (KeyError, ValueError):
        return len(os.sched_getaffinity(0))


def cpu_baseline(m, k, c, iterations: int, sample_iters: int) -> dict:
    """The oracle's ports of the reference's CPU solvers on this host: one
    full solve of the OpenMP loop (CGOpenMP.cpp:188-265, order 3) on the
    process's CPU share, threads bound to cores (OMP_PROC_BIND=close,
    OMP_PLACES=cores), to the same stop test; and SerialCG (order 0) on one
    core over `sample_iters` iterations, scaled to the FP64 iteration count
    (a full serial solve takes minutes)."""
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    from oracle import oracle

    threads = cpu_threads()
    host = host_info()
    affinity = os.sched_getaffinity(0)
    t0 = time.perf_counter()
    _, it, res, _ = oracle.pcg(n=m.n, fmt="CRS", precond="jacobi", order=3, nthreads=threads, k=k,
                               tol=1e-9, max_iter=100000, crs=(m.ptr, m.index, m.value), C_jac=c,
                               history=False)
    v_omp = time.perf_counter() - t0
    # with OMP_PROC_BIND the OpenMP runtime binds this thread to one core
    os.sched_setaffinity(0, affinity)
    t0 = time.perf_counter()
    oracle.pcg(n=m.n, fmt="CRS", precond="jacobi", order=0, nthreads=1, k=k, tol=1e-300,
               max_iter=sample_iters, crs=(m.ptr, m.index, m.value), C_jac=c, history=False)
    dt_ser = time.perf_counter() - t0
    v_ser = dt_ser / (sample_iters + 1) * (iterations + 1)  # + the set-up pass
    return {
        "value": v_omp,
        "unit": "s",
        "cores": threads,
        "kind": "port",
        "sample": (f"one full FP64 Jacobi-PCG solve of this workload by the OpenMP port of CGOpenMP "
                   f"(oracle/cg_oracle.c, order 3) on {threads} threads bound to cores: {it} "
                   f"iterations, relative residual {res:.3e}"),
        "host": host,
        "serial": {"value": v_ser, "cores": 1, "kind": "port", "scaled_sample": True,
                   "sample": f"SerialCG port (order 0): {sample_iters} iterations in {dt_ser:.2f} s, "
                             f"scaled to the {iterations} iterations of the FP64 solve"},
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
                         "calibration kernel of known size), 20 FP64 CRS, 20 FP64 ELL, 20 mixed "
                         "Jacobi and 5 mixed polynomial iterations")
    ap.add_argument("--trace", action="store_true",
                    help="launch the bench's kernels directly (no HIP graph replays) for a kernel trace")
    ap.add_argument("--summarize-pmc", nargs=4, metavar=("FETCH_DIR", "WRITE_DIR", "META", "STEM"),
                    help="write profiles/STEM.json/.md from two rocprofv3 --pmc runs of --pmc")
    ap.add_argument("--summarize-trace", nargs=3, metavar=("TRACE_DIR", "BENCH_JSON", "STEM"),
                    help="write profiles/STEM.md from a rocprofv3 --kernel-trace --stats run")
    ap.add_argument("--workload", choices=("diffusion256", "poisson128", "executables"),
                    default="diffusion256",
                    help="poisson128: the plain FP64 path on the 128^3 7-point Poisson system "
                         "(CRS and ELL, per-hook GB/s, the CPU ports' full solves beside it); "
                         "executables: the cg_hip executable on the 256^2 Laplacian and the 128^3 "
                         "Poisson file (loops, COO, the reader's IO time) with the oracle's ports "
                         "of the reference's CPU solvers and reader beside it; "
                         "reports of their own, not the benchmark line")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path "
              f"(--gpus {args.gpus} requested)", file=sys.stderr)
        return 2
    if args.summarize_pmc:
        return summarize_pmc(*args.summarize_pmc)
    if args.summarize_trace:
        return summarize_trace(*args.summarize_trace)
    if args.workload == "poisson128":
        return poisson128(args)
    if args.workload == "executables":
        return executables(args)

    import numpy as np

    from cgxx_amd import learned, problems
    from cgxx_amd.cghip import Context, Mlp

    t = time.perf_counter()
    m = problems.diffusion3d(args.side, sigma=1.0, corr=8.0, seed=20261202)
    k = m.ones_rhs()
    c = 1.0 / m.diagonal()
    n, nnz = m.n, m.nnz
    log(f"generated {args.side}^3 diffusion: N={n} nnz={nnz} in {time.perf_counter() - t:.1f} s")

    ctx = Context(0)
    setup = {}
    t = time.perf_counter()
    ctx.upload_crs(m.ptr, m.index, m.value)
    setup["upload_crs_s"] = time.perf_counter() - t  # host validation + H2D
    ctx.upload_ell(*m.to_ell())
    ctx.upload_jacobi(c)
    ctx.set_format("CRS")
    t = time.perf_counter()
    ctx.prepare_fp32()
    setup["prepare_fp32_s"] = time.perf_counter() - t  # FP32 copy, scaled DIA/ELL/CRS operator
    ctx.upload("K", k)
    nnz_nd = nnz - n  # the scaled operator stores no diagonal
    layout = ctx.fp32_layout()
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(m.ptr))
    offs = np.unique(m.index.astype(np.int64) - rows)
    ndiag = int(offs.size - (offs == 0).sum())
    del rows
    # operator stream of one FP32 SpMV of the scaled operator (without x, y)
    op_bytes = {"DIA": 4 * ndiag * n, "ELL": 8 * nnz_nd + 4 * n,
                "CRS": 8 * nnz_nd + 4 * (n + 1)}[layout]
    # the Clenshaw steps read each row's FP16 record, or with symmetric
    # storage the upper diagonals' values only (2 bytes per stored pair)
    poly_values = ctx.poly_values_layout()
    poly_op_bytes = (ndiag * n if poly_values == "FP16-SYM" else 4 * ((ndiag + 1) // 2) * n) \
        if layout == "DIA" else op_bytes
    kind = {"CRS": 0, "ELL": 1, "DIA": 2}[layout]
    log(f"uploaded FP64 CRS + ELL, Jacobi, FP32 copy and scaled FP32 operator ({layout}"
        + (f", {ndiag} diagonals)" if layout == "DIA" else ")"))

    model = learned.load_model()
    mlp = Mlp(tuple(int(v) for v in model["dims"]), 1)
    mlp.set_params(model["params"])
    rho = float(model["rho_star"])
    fixed_degree = int(model["fixed_degree"])

    max_iter = 100000
    if args.trace:
        # the kernels of the bench, each launched directly (no graph
        # replays), for a profiler's kernel trace
        for fmt in ("CRS", "ELL"):
            ctx.set_format(fmt)
            ctx.fill("X", 0.0)
            ctx.solve(precond="jacobi", max_iter=100, tol=1e-9, profile=1)
            ctx.fill("P", 1.0)
            ctx.time_matvec("P", "Q", reps=20)
        ctx.set_format("CRS")
        ctx.fill("X", 0.0)
        ctx.solve_mixed(precond="jacobi", max_iter=100, tol=1e-9, profile=1)
        f, b = learned.matrix_features(ctx, n)
        pr = learned.predict(mlp, model, f)
        k_deg = pr["degree"] if pr["choice"] != "jacobi" else fixed_degree
        from cgxx_amd import spectrum
        lower = pr["lower_end"] if pr["choice"] != "jacobi" else learned.ritz_features(ctx)[1] / rho
        ctx.set_poly(spectrum.cheb_residual_coeffs(min(lower, 0.25 * b), b, k_deg), b)
        ctx.fill("X", 0.0)
        ctx.solve_mixed(precond="poly", max_iter=20, tol=1e-9, profile=1)
        step_ms = ctx.time_poly_step(reps=20)
        log(f"trace run: 100 FP64 CRS and ELL, 20+20 plain SpMVs, 100 mixed Jacobi and 20 mixed "
            f"degree-{k_deg} polynomial iterations kernel by kernel, 20 applications; Clenshaw step "
            f"{step_ms:.4f} ms")
        print(json.dumps({"config": {"workload": f"3D 7-point variable-coefficient diffusion {args.side}^3",
                                     "n": n, "nnz": nnz,
                                     "format": f"scaled FP32 operator in {layout}"
                                               + (f" ({ndiag} diagonals)" if layout == "DIA" else "")},
                          "degree": k_deg, "clenshaw_step_ms": step_ms}), flush=True)
        mlp.close()
        ctx.close()
        return 0
    if args.pmc:
        ctx.fill("P", 1.0)
        ctx.cpy("R", "P")  # reads and writes exactly 8*N bytes each
        for fmt in ("CRS", "ELL"):
            ctx.set_format(fmt)
            ctx.fill("X", 0.0)
            ctx.solve(precond="jacobi", max_iter=20, tol=1e-9, batch=1)
        ctx.set_format("CRS")
        ctx.fill("X", 0.0)
        ctx.solve_mixed(precond="jacobi", max_iter=20, tol=1e-9, batch=1)
        _, b = learned.matrix_features(ctx, n)
        _, theta10 = learned.ritz_features(ctx)
        ctx.set_poly(learned.fixed_rule_coeffs(theta10, b, rho, fixed_degree), b)  # traffic only
        ctx.fill("X", 0.0)
        ctx.solve_mixed(precond="poly", max_iter=5, tol=1e-9, batch=1)
        for fmt in ("CRS", "ELL"):
            ctx.set_format(fmt)
            ctx.time_matvec("P", "Q", reps=5)
        os.makedirs(os.path.join(REPO, "gpurun_out"), exist_ok=True)
        with open(os.path.join(REPO, "gpurun_out", "pmc_meta.json"), "w") as f:
            json.dump({"n": n, "nnz": nnz, "ndiag": ndiag, "layout": layout, "kind": kind,
                       "side": args.side}, f)
        log(f"pmc run: copy of {8 * n} bytes, 20+20 FP64, 20 mixed Jacobi, 5 mixed poly iterations, "
            "5 plain CRS and ELL SpMVs")
        mlp.close()
        ctx.close()
        return 0

    first_call = {}

    def timed(fn, label):
        for i in range(max(args.warmup, 1)):
            t0 = time.perf_counter()
            info = fn()
            if i == 0:
                first_call[label] = time.perf_counter() - t0
            log(f"{label} warmup {i}: {info['wall_s']:.4f} s")
        t0 = time.perf_counter()
        infos = [fn() for _ in range(args.steps)]
        sec = (time.perf_counter() - t0) / args.steps
        log(f"{label}: {args.steps} timed solves, {sec:.4f} s each")
        return sec, infos[-1]

    def per(info, key):
        return info[f"ms_{key}"] / max(info[f"n_{key}"], 1)

    # ---- plain FP64 Jacobi-PCG (the reference's method), CRS and ELL ----
    fp64 = {}
    xs = {}
    for fmt in ("CRS", "ELL"):
        ctx.set_format(fmt)

        def run64(profile=0):
            ctx.fill("X", 0.0)
            return ctx.solve(precond="jacobi", max_iter=max_iter, tol=1e-9, profile=profile)

        sec, info = timed(run64, f"fp64 {fmt}")
        nr, nb = ctx.true_residual("X", "K")
        xs[fmt] = ctx.download("X")
        prof = run64(profile=1)
        spmv_ms = per(prof, "matvec")
        by = crs_bytes(n, nnz, 8, 8) if fmt == "CRS" else ell_bytes(n, nnz, 8, 8)
        plain_ms = ctx.time_matvec("P", "Q", reps=20)
        fp64[fmt] = {
            "seconds": sec, "iterations": info["iterations"], "true_relative_residual": nr / nb,
            "spmv_dot_ms": spmv_ms, "spmv_dot_gbps": by / (spmv_ms * 1e-3) / 1e9,
            "spmv_plain_ms": plain_ms, "spmv_plain_gbps": by / (plain_ms * 1e-3) / 1e9,
            "spmv_plain_hbm_frac": by / (plain_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "update_ms": per(prof, "update"), "xpay_ms": per(prof, "xpay"),
            "algorithmic_spmv_bytes": by,
        }
        log(f"fp64 {fmt}: {info['iterations']} iterations, true residual {nr / nb:.3e}, "
            f"spmv+dot {spmv_ms:.4f} ms = {fp64[fmt]['spmv_dot_gbps']:.0f} GB/s, plain spmv "
            f"{plain_ms:.4f} ms = {fp64[fmt]['spmv_plain_gbps']:.0f} GB/s")
    base_fmt = min(fp64, key=lambda f: fp64[f]["seconds"])
    sec64 = fp64[base_fmt]["seconds"]
    x64 = xs[base_fmt]
    ctx.set_format("CRS")

    # ---- mixed precision, Jacobi (scaled operator) ----
    def mixed(profile=0):
        ctx.fill("X", 0.0)
        return ctx.solve_mixed(precond="jacobi", max_iter=max_iter, tol=1e-9, profile=profile)

    secmj, infomj = timed(mixed, "mixed jacobi")
    profmj = mixed(profile=1)
    spmv_nd_bytes = op_bytes + 8 * n  # FP32 scaled operator, x and y
    mixed_j = {
        "seconds": secmj, "iterations": infomj["iterations"],
        "reliable_updates": infomj["updates"], "true_relative_residual": infomj["true_residual"],
        "spmv_dot_ms": per(profmj, "matvec"),
        "spmv_dot_gbps": spmv_nd_bytes / (per(profmj, "matvec") * 1e-3) / 1e9,
        "update_ms": per(profmj, "update"), "xpay_ms": per(profmj, "xpay"),
        "reliable_update_ms": per(profmj, "refresh"),
        "speedup_vs_fp64": sec64 / secmj,
    }
    log(f"mixed jacobi: {infomj['iterations']} iterations, {secmj:.4f} s; spmv+dot "
        f"{mixed_j['spmv_dot_ms']:.4f} ms, update {mixed_j['update_ms']:.4f} ms, p update "
        f"{mixed_j['xpay_ms']:.4f} ms")

    # ---- mixed precision, learned polynomial (the measured method) ----
    def learned_solve():
        return learned.solve_learned(ctx, mlp, model, n, tol=1e-9, max_iter=max_iter)

    secl, infol = timed(learned_solve, "mixed+learned")
    nrl, nbl = ctx.true_residual("X", "K")
    xl = ctx.download("X")
    prof_coeffs, prof_src = infol["coeffs"], "learned"
    if infol["coeffs"] is None or infol["method"] == "jacobi":
        # Jacobi chosen or the polynomial refused: time the kernels with the fixed rule's
        _, b = learned.matrix_features(ctx, n)
        _, theta10 = learned.ritz_features(ctx)
        prof_coeffs = learned.fixed_rule_coeffs(theta10, b, rho, fixed_degree)
        prof_src = "fixed rule"
    ctx.set_poly(prof_coeffs, infol["b"])
    ctx.fill("X", 0.0)
    profl = ctx.solve_mixed(precond="poly", max_iter=max_iter, tol=1e-9, profile=1)
    deg = len(prof_coeffs) - 1
    step_ms_events = per(profl, "poly")  # kernel by kernel, events between iterations
    step_ms = ctx.time_poly_step(reps=5)  # steady state: applications back to back
    # Clenshaw steps, averaged over one application: each streams the
    # operator and writes B_j; the first reads r only (gathered, as
    # B_k = c_k r), the second B_{k-1} and r (also as B_k), every later one
    # B_{j+1} (gathered), r and B_{j+2}
    step_bytes = poly_op_bytes + (8 * n + 12 * n + (deg - 2) * 16 * n) / deg
    step_gbps = step_bytes / (step_ms * 1e-3) / 1e9
    log(f"mixed+learned: method {infol['method']}, {infol['poly_iterations']} poly + "
        f"{infol['jacobi_iterations']} Jacobi iterations ({infol['spmvs']} FP32 SpMVs), true residual "
        f"{nrl / nbl:.3e}, {secl:.4f} s; Clenshaw step {step_ms:.4f} ms = {step_gbps:.0f} GB/s")

    # ---- the method without the network: one degree on [theta10 / rho*, b] ----
    def fixed_solve():
        return learned.solve_fixed_rule(ctx, n, rho, fixed_degree, tol=1e-9, max_iter=max_iter)

    secf, infof = timed(fixed_solve, "mixed+fixed-rule poly")

    # ---- the measured method with the polynomial on the FP32 operator values ----
    ctx.set_poly_values(32)
    secl32, infol32 = timed(learned_solve, "mixed+learned, FP32 values")
    ctx.set_poly(infol32["coeffs"], infol32["b"])
    step32_ms = ctx.time_poly_step(reps=5)
    ctx.set_poly_values(16)
    step32_bytes = op_bytes + (8 * n + 12 * n + (deg - 2) * 16 * n) / deg
    log(f"mixed+learned, FP32 values: {infol32['spmvs']} FP32 SpMVs, {secl32:.4f} s; Clenshaw step "
        f"{step32_ms:.4f} ms = {step32_bytes / (step32_ms * 1e-3) / 1e9:.0f} GB/s")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(m, k, c, fp64[base_fmt]["iterations"], args.cpu_iters)
        log(f"cpu baseline: {cpu['value']:.1f} s ({cpu['sample']})")

    out = {
        "metric": METRIC,
        "value": secl,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": secl * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f16+f32+f64",
        "data": "synthetic",
        "config": {
            "workload": f"3D 7-point variable-coefficient diffusion {args.side}^3 "
                        "(log-normal kappa, sigma=1, corr=8, seed 20261202), k=A*1, x0=0, "
                        "FP64 true relative residual 1e-9",
            "method": f"mixed+learned: FP32-inner CG on D^-1/2 A D^-1/2 ({layout}, FP32 values in the CG "
                      "operator), polynomial preconditioner applied by Clenshaw steps reading the "
                      + ("FP16 copy of the DIA values" if layout == "DIA" else "FP32 values")
                      + ", dots accumulated in FP64, FP64 reliable updates and true residual; "
                      "the trained network predicts theta_min and each method's cost from the matrix "
                      "statistics alone and chooses the polynomial (degree, interval "
                      "[c_k theta_min, b]) or Jacobi; statistics, inference, choice and fallback "
                      "inside the timed step",
            "n": n,
            "nnz": nnz,
            "format": f"CRS upload, scaled FP32 operator in {layout}"
                      + (f" ({ndiag} diagonals)" if layout == "DIA" else ""),
        },
        "iterations": infol["poly_iterations"] + infol["jacobi_iterations"],
        "fp32_spmvs": infol["spmvs"],
        "fallback_method": infol["method"],
        "learned_choice": {k: v for k, v in infol["decision"].items() if k != "cost_ms"},
        "learned_predicted_cost_ms": {str(k): v for k, v in infol["decision"]["cost_ms"].items()},
        "reliable_updates": infol["updates"],
        "true_relative_residual": nrl / nbl,
        "max_abs_error_vs_ones": float(np.abs(xl - 1.0).max()),
        "max_abs_diff_vs_fp64_solution": float(np.abs(xl - x64).max()),
        "speedup_vs_fp64_jacobi": sec64 / secl,
        "fp64_baseline_format": base_fmt,
        "fp64_jacobi": fp64,
        "mixed_jacobi": mixed_j,
        "mixed_fixed_rule_poly": {"seconds": secf, "rho": rho, "degree": fixed_degree,
                                  "spmvs": infof["spmvs"], "theta10": infof["theta10"],
                                  "true_relative_residual": infof["true_residual"],
                                  "speedup_vs_fp64": sec64 / secf},
        "mixed_learned_fp32_values": {"seconds": secl32, "fp32_spmvs": infol32["spmvs"],
                                      "true_relative_residual": infol32["true_residual"],
                                      "clenshaw_step_ms": step32_ms,
                                      "clenshaw_step_gbps": step32_bytes / (step32_ms * 1e-3) / 1e9,
                                      "speedup_vs_fp64": sec64 / secl32},
        "setup_s": dict(setup, first_call_extra_s={
            label: first_call[label] - sec for label, sec in
            (("mixed+learned", secl), ("mixed jacobi", secmj), ("mixed+fixed-rule poly", secf))},
            note="outside value (the reference's Converting convention): upload with host "
                 "validation, the FP32/scaled-operator build, and each method's first solve "
                 "beyond a steady one (HIP graph capture, buffers)"),
        "kernel_ms": {"clenshaw_step": step_ms, "clenshaw_step_in_profiled_solve": step_ms_events,
                      "spmv_dot_fp32": per(profl, "matvec"),
                      "update_fp32": per(profl, "update"), "xpay_fp32": per(profl, "xpay"),
                      "reliable_update": per(profl, "refresh")},
        "roofline": {
            "bound": "hbm",
            "achieved": step_gbps,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": step_gbps / HBM_PEAK_GBS,
            "traffic": application_traffic(kind, deg),
            "kernel": f"k_poly_step<*, {kind}, *>: the Clenshaw steps of one degree-{deg} application, "
                      f"per step (scaled operator in {layout}"
                      + (", FP16 values)" if layout == "DIA" else ", FP32 values)"),
            "polynomial": prof_src,
            "algorithmic_bytes_per_launch": step_bytes,
        },
        "cpu_baseline": cpu,
    }
    mlp.close()
    ctx.close()
    print(json.dumps(out), flush=True)
    return 0


def poisson128(args) -> int:
    """The plain FP64 path on the 3D 7-point Poisson system on 128^3
    (N = 2,097,152, 14,581,760 nonzeros, values 6/-1), k = A*1, x0 = 0,
    Jacobi, stop at relative residual 1e-9: the fused device loop and the
    reference's loop hook by hook, in CRS and in ELL; each hook's kernel
    time and GB/s (algorithmic bytes: SpMV as in crs_bytes/ell_bytes, dot
    16 N, axpy/xpay/Jacobi 24 N, copy 16 N); beside them the oracle's
    SerialCG port (1 core) and OpenMP port (up to 16 threads), full solves
    on this host.  The matrix (217 MB) fits in the 256 MB last-level cache,
    so the rates are effective ones.  Prints one JSON object."""
    import numpy as np

    from cgxx_amd import problems
    from cgxx_amd.cghip import Context, reference_loop

    m = problems.laplace3d(128)
    n, nnz = m.n, m.nnz
    k = m.ones_rhs()
    c = 1.0 / m.diagonal()
    ctx = Context(0)
    ctx.upload_crs(m.ptr, m.index, m.value)
    ctx.upload_ell(*m.to_ell())
    ctx.upload_jacobi(c)
    ctx.upload("K", k)
    log(f"128^3 Poisson: N={n} nnz={nnz}")
    out = {"workload": "3D 7-point Poisson 128^3 (6/-1), k=A*1, x0=0, Jacobi, relative residual 1e-9",
           "n": n, "nnz": nnz, "formats": {}}
    vec_bytes = {"dot": 16 * n, "axpy": 24 * n, "xpay": 24 * n, "jacobi": 24 * n, "cpy": 16 * n}
    for fmt in ("CRS", "ELL"):
        ctx.set_format(fmt)

        def fused():
            ctx.fill("X", 0.0)
            return ctx.solve(precond="jacobi", max_iter=1000, tol=1e-9)

        fused()
        t0 = time.perf_counter()
        infos = [fused() for _ in range(args.steps)]
        sec = (time.perf_counter() - t0) / args.steps
        nr, nb = ctx.true_residual("X", "K")
        err = float(np.max(np.abs(ctx.download("X") - 1.0)))
        ctx.fill("X", 0.0)
        t0 = time.perf_counter()
        it_h, res_h, _ = reference_loop(ctx, "jacobi", 1000, 1e-9)
        sec_h = time.perf_counter() - t0
        for v in ("P", "R", "Q", "Z"):
            ctx.fill(v, 1.0)
        by = {"matvec": crs_bytes(n, nnz, 8, 8) if fmt == "CRS" else ell_bytes(n, nnz, 8, 8),
              **vec_bytes}
        hooks = {}
        for h in ("matvec", "dot", "axpy", "xpay", "jacobi", "cpy"):
            ms = ctx.time_hook(h, reps=50)
            hooks[h] = {"ms": ms, "bytes": by[h], "gbps": by[h] / (ms * 1e-3) / 1e9}
        out["formats"][fmt] = {
            "fused_seconds": sec, "iterations": infos[-1]["iterations"],
            "residual": infos[-1]["residual"], "true_relative_residual": nr / nb,
            "max_abs_error_vs_ones": err,
            "hook_by_hook_seconds": sec_h, "hook_by_hook_iterations": it_h, "hooks": hooks}
        log(f"{fmt}: fused {infos[-1]['iterations']} iterations {sec:.4f} s, hook by hook {it_h} "
            f"iterations {sec_h:.3f} s; " + ", ".join(f"{h} {v['gbps']:.0f} GB/s" for h, v in hooks.items()))
    ctx.close()
    if not args.no_cpu:
        from oracle import oracle

        threads = max(1, min(16, len(os.sched_getaffinity(0))))
        cpu = {}
        for name, order, nt in (("serial", 0, 1), ("openmp", 3, threads)):
            t0 = time.perf_counter()
            _, it, res, _ = oracle.pcg(n=n, fmt="CRS", precond="jacobi", order=order, nthreads=nt, k=k,
                                       tol=1e-9, max_iter=1000, crs=(m.ptr, m.index, m.value), C_jac=c,
                                       history=False)
            cpu[name] = {"seconds": time.perf_counter() - t0, "iterations": it, "residual": res,
                         "cores": nt, "kind": "port"}
            log(f"cpu {name} ({nt} cores): {it} iterations {cpu[name]['seconds']:.2f} s")
        out["cpu_baseline"] = cpu
    print(json.dumps(out), flush=True)
    return 0


def executables(args) -> int:
    """The cg_hip executable through its command line, with the oracle's
    ports of the reference's CPU paths beside it on this host, on two files
    written by cgxx_amd.problems.write_mm: the 256^2 5-point Laplacian (C1)
    and the 128^3 7-point Poisson system.  C1: cg_hip's three loops (hook by
    hook, fused with per-iteration kernel timing, fused from HIP graphs) and
    full solves of the SerialCG port (1 core) and the CGOpenMP port (every
    core of this host) from the file read by the oracle's reader.  128^3:
    cg_hip in COO, CRS and ELL and the SerialCG port's COO solve; the
    reader's IO time at 1, 4 and all host threads and with its binary cache,
    beside the oracle's port of the reference's reader (one thread, then the
    CRS conversion).  3 repetitions of each cg_hip run; one JSON object."""
    import re
    import subprocess
    import tempfile

    import numpy as np

    from cgxx_amd import problems
    from oracle import oracle

    exe = os.path.join(REPO, "cgxx_amd", "bin", "cg_hip")
    cores = len(os.sched_getaffinity(0))
    out = {"host": host_info(), "cores_used_openmp": cores}

    def cli(path, reps=3, **env):
        e = dict(os.environ, **{k: str(v) for k, v in env.items()})
        r = subprocess.run([exe, path, str(reps)], capture_output=True, text=True, env=e, timeout=600)
        if r.returncode:
            raise RuntimeError(r.stderr)
        get = lambda label: [ln.split(":", 1)[1].strip() for ln in r.stdout.splitlines()
                             if ln.startswith(label + ":")]
        return {"iterations": [int(v) for v in get("Iterations")], "residual": get("Residual"),
                "solve_s": [float(v) for v in get("Solve time")],
                "io_s": [float(v) for v in get("IO time")],
                "converting_s": [float(v) for v in get("Converting time")],
                "matvec_gflops": [float(v) for v in get("MatVec GFLOP/s")],
                "correct": r.stdout.count("Solution is correct!")}

    def port(mm, fmt, order, threads, k, c, max_iter):
        t0 = time.perf_counter()
        kw = {"mm": mm} if fmt == "COO" else {"crs": mm.crs()}
        x, it, res, _ = oracle.pcg(n=mm.N, fmt=fmt, precond="jacobi", order=order, nthreads=threads,
                                   k=k, tol=1e-9, max_iter=max_iter, C_jac=c, history=False, **kw)
        return {"seconds": time.perf_counter() - t0, "iterations": it, "residual": "%e" % res,
                "correct": bool(np.all(np.abs(x - 1.0) <= 1e-5)), "cores": threads, "kind": "port"}

    with tempfile.TemporaryDirectory() as tmp:
        c1 = os.path.join(tmp, "c1.mtx")
        problems.write_mm(c1, problems.laplace2d(256))
        env = dict(CG_MATRIX_FORMAT="CRS", CG_PRECONDITIONER="jacobi", CG_MAX_ITER=1000)
        r = {"hooks": cli(c1, CG_HIP_FUSED=0, **env), "fused_timed": cli(c1, **env),
             "fused_graph": cli(c1, CG_HIP_GRAPH=1, **env)}
        mm = oracle.MMFile(c1)
        k, (c, _) = mm.rhs_ones(), mm.jacobi()
        r["serial_port"] = [port(mm, "CRS", 0, 1, k, c, 1000) for _ in range(3)]
        r["openmp_port"] = [port(mm, "CRS", 3, cores, k, c, 1000) for _ in range(3)]
        out["c1_lap2d_256"] = r
        log("c1: " + ", ".join(f"{key} {min(v['solve_s']) if isinstance(v, dict) else min(x['seconds'] for x in v):.4f} s"
                               for key, v in r.items()))
        del mm

        c2 = os.path.join(tmp, "c2.mtx")
        problems.write_mm(c2, problems.laplace3d(128))
        env = dict(CG_PRECONDITIONER="jacobi", CG_MAX_ITER=1000)
        r = {fmt: cli(c2, CG_MATRIX_FORMAT=fmt, **env) for fmt in ("COO", "CRS", "ELL")}
        t0 = time.perf_counter()
        mm = oracle.MMFile(c2)
        t_read = time.perf_counter() - t0
        t0 = time.perf_counter()
        crs = mm.crs()
        t_conv = time.perf_counter() - t0
        k, (c, _) = mm.rhs_ones(), mm.jacobi()
        r["serial_port_coo"] = port(mm, "COO", 0, 1, k, c, 1000)
        out["c2_poisson128"] = r
        io = {f"threads_{t}": cli(c2, reps=1, CG_IO_THREADS=t, CG_MAX_ITER=1,
                                  CG_PRECONDITIONER="jacobi") for t in (1, 4, cores)}
        cache = os.path.join(tmp, "c2.cache")
        io["cache_written"] = cli(c2, reps=1, CG_MATRIX_CACHE=cache, CG_MAX_ITER=1,
                                  CG_PRECONDITIONER="jacobi")
        io["cache_loaded"] = cli(c2, reps=1, CG_MATRIX_CACHE=cache, CG_MAX_ITER=1,
                                 CG_PRECONDITIONER="jacobi")
        io["oracle_reader_port"] = {"read_s": t_read, "crs_conversion_s": t_conv,
                                    "entries": int(crs[0][-1]), "cores": 1, "kind": "port"}
        out["reader_c2"] = io
        log("c2: " + ", ".join(f"{fmt} {min(r[fmt]['solve_s']):.4f} s" for fmt in ("COO", "CRS", "ELL"))
            + f", serial port COO {r['serial_port_coo']['seconds']:.2f} s; reader port {t_read:.3f} s")
    print(json.dumps(out), flush=True)
    return 0


def host_info() -> dict:
    """CPU model, sockets and logical CPUs of this host (lscpu), and the CPUs
    this process may run on."""
    import subprocess

    info = {"nproc": len(os.sched_getaffinity(0))}
    try:
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for ln in txt.splitlines():
            key, _, val = ln.partition(":")
            if key.strip() in ("Model name", "Socket(s)", "CPU(s)", "Thread(s) per core",
                               "Core(s) per socket"):
                info[key.strip()] = val.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return info


if __name__ == "__main__":
    sys.exit(main())
