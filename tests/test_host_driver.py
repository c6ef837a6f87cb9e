"""The build's CG driver (cgxx_amd/host/CG.cpp, main.cpp) on the CPU.

tests/host_cg.cpp plugs a host backend into the driver (hooks with
SerialCG's arithmetic), so the driver's own code -- environment, reader and
converters, the reference loop hook by hook, the check, the report and the
CG_REPORT_JSON record -- is checked without a GPU: against the reference's
golden runs (iterations, printed residual, check line, the first entries of
x) for CRS, ELL and COO, and against the reference binary's full output
where it is built (oracle/_ref/cg_serial).
"""
import json
import os
import re
import subprocess

import numpy as np
import pytest

from tests.cases import SANITIZERS, Case, load_cli_fixture, untimed_report, write_cli_matrices
from tests.conftest import REPO, golden_runs

HOST = os.path.join(REPO, "cgxx_amd", "host")
REF = os.path.join(REPO, "oracle", "_ref", "cg_serial")
RUNS = golden_runs(max_n=70_000)
TIMED = re.compile(r"^(IO time|Converting time|Solve time|Check time|Total time \(excl\. IO\)|MatVec time|"
                   r"MatVec GFLOP/s|axpy time|xpay time|vectorDot time|Preconditioner time):")


@pytest.fixture(scope="module", params=["plain", "asan-ubsan"])
def host_cg(request, tmp_path_factory):
    """The driver on a host backend, plain and under AddressSanitizer +
    UndefinedBehaviorSanitizer: every test below runs against both."""
    out = str(tmp_path_factory.mktemp("hostcg") / f"host_cg_{request.param}")
    srcs = [os.path.join(REPO, "tests", "host_cg.cpp")] + [
        os.path.join(HOST, f) for f in ("CG.cpp", "Matrix.cpp", "Preconditioner.cpp", "main.cpp")]
    subprocess.run(["g++", "-std=c++17", *SANITIZERS[request.param], "-ffp-contract=off", "-pthread",
                    "-I" + HOST, "-o", out, *srcs], check=True)
    return out


def run(exe, path, env, reps=None):
    e = dict(os.environ)
    e.update({k: str(v) for k, v in env.items()})
    args = [exe, path] + ([str(reps)] if reps else [])
    return subprocess.run(args, capture_output=True, text=True, env=e, timeout=300)


def field(stdout, label, index=0):
    return [ln.split(":", 1)[1].strip() for ln in stdout.splitlines() if ln.startswith(label + ":")][index]


def untimed(stdout):
    return [TIMED.sub(lambda m: m.group(1) + ":", ln).split(":")[0] if TIMED.match(ln) else ln
            for ln in stdout.splitlines()]


@pytest.mark.parametrize("pair", RUNS, ids=[f"{p[0]['case']}-{p[1]['env']['CG_MATRIX_FORMAT']}-"
                                            f"{p[1]['env']['CG_PRECONDITIONER']}" for p in RUNS])
def test_driver_matches_golden_run(pair, host_cg, tmp_path):
    fx, rn = pair
    case = Case(fx, str(tmp_path))
    env = dict(rn["env"], CG_IO_THREADS=3)
    r = run(host_cg, case.path, env)
    assert r.returncode == 0, r.stderr
    assert int(field(r.stdout, "Iterations")) == rn["iterations"]
    assert field(r.stdout, "Residual") == rn["residual_printed"]
    assert rn["check"] in r.stdout
    xs = re.search(r"^x = \[ (.*) \]$", r.stdout, re.M).group(1).split()
    want = [float.fromhex(v) if isinstance(v, str) else v for v in rn["x_head"]]
    assert [float(v) for v in xs] == pytest.approx(want, rel=1e-5)
    if os.path.exists(REF):
        ref = run(REF, case.path, rn["env"])
        assert untimed(r.stdout) == untimed(ref.stdout)


CLI = load_cli_fixture()


@pytest.mark.parametrize("fmt", ["CRS", "ELL", "COO"])
def test_driver_report_matches_cli_fixture(host_cg, fmt, tmp_path):
    """The driver's whole report (host backend, the reference's arithmetic)
    equals the reference binary's stored report line for line, timings
    blanked: the fixture the GPU tests of cg_hip compare against."""
    files = write_cli_matrices(str(tmp_path))
    for name, path in files.items():
        r = run(host_cg, path, dict(CG_MATRIX_FORMAT=fmt, **CLI["env"]))
        assert r.returncode == 0, r.stderr
        assert untimed_report(r.stdout, path) == CLI["cases"][name]["runs"][fmt], name


def test_driver_solve_record(host_cg, tmp_path):
    """CG_REPORT_JSON on the CPU: one line per repetition with the printed
    figures, and a file name that needs JSON escaping."""
    fx, rn = next(p for p in RUNS if p[0]["case"] == "lap2d_16" and p[1]["env"]["CG_MATRIX_FORMAT"] == "ELL")
    case = Case(fx, str(tmp_path))
    odd = str(tmp_path / 'we"ird\\name.mtx')
    os.rename(case.path, odd)
    rec = tmp_path / "rec.jsonl"
    r = run(host_cg, odd, dict(rn["env"], CG_REPORT_JSON=str(rec)), reps=2)
    assert r.returncode == 0, r.stderr
    lines = rec.read_text().splitlines()
    assert len(lines) == 2
    for i, ln in enumerate(lines):
        d = json.loads(ln)
        assert d["matrix"] == odd and d["repetition"] == i
        assert d["iterations"] == rn["iterations"] == int(field(r.stdout, "Iterations", i))
        assert "%e" % d["residual"] == rn["residual_printed"]
        assert d["format"] == rn["env"]["CG_MATRIX_FORMAT"]
        assert d["incorrect_elements"] == 0
        assert d["backend"] == {"name": "base"}
        assert abs(d["times"]["solve"] - float(field(r.stdout, "Solve time", i))) <= 1e-6


def test_driver_strict_check_and_errors(host_cg, tmp_path):
    fx, rn = next(p for p in RUNS if p[0]["case"] == "lap2d_16")
    case = Case(fx, str(tmp_path))
    bad = run(host_cg, case.path, dict(CG_MAX_ITER=2, CG_STRICT_CHECK=1))
    assert bad.returncode == 2 and "elements are incorrect!" in bad.stdout
    assert run(host_cg, case.path, dict(CG_MAX_ITER=2)).returncode == 0
    r = run(host_cg, case.path, dict(CG_REPORT_JSON=str(tmp_path / "no" / "dir.jsonl")))
    assert r.returncode == 1 and "Could not write the solve record" in r.stderr
    r = run(host_cg, case.path, dict(CG_IO_THREADS="0"))
    assert r.returncode == 1 and r.stderr.strip() == "Invalid value for CG_IO_THREADS!"


@pytest.mark.parametrize("fmt", ["CRS", "ELL", "COO"])
def test_random_rhs_judged_by_true_residual(host_cg, fmt, tmp_path):
    """CG_RHS=random:<seed> (no reference counterpart): k_i = 2 u_i - 1 from
    SplitMix64, the same vector as cgxx_amd.problems.random_rhs (the oracle's
    SerialCG port on that k takes the same iterations and prints the same
    residual), and the check judges the FP64 true relative residual."""
    fx, rn = next(p for p in RUNS if p[0]["case"] == "lap2d_64" and p[1]["env"]["CG_MATRIX_FORMAT"] == "CRS")
    case = Case(fx, str(tmp_path))
    rec = tmp_path / "rec.jsonl"
    r = run(host_cg, case.path, dict(CG_MATRIX_FORMAT=fmt, CG_PRECONDITIONER="jacobi", CG_RHS="random:12345",
                                     CG_REPORT_JSON=str(rec)))
    assert r.returncode == 0, r.stderr
    from cgxx_amd import problems
    from oracle import oracle
    k = problems.random_rhs(case.n, 12345)
    assert k.min() >= -1.0 and k.max() < 1.0 and abs(k.mean()) < 0.05
    x, it, res, _ = oracle.pcg(n=case.n, fmt="CRS", precond="jacobi", order=0, k=k, crs=case.crs,
                               C_jac=case.C)
    assert int(field(r.stdout, "Iterations")) == it
    assert field(r.stdout, "Residual") == "%e" % res
    A = case.crs
    ax = np.array([A[2][A[0][i]:A[0][i + 1]] @ x[A[1][A[0][i]:A[0][i + 1]]] for i in range(case.n)])
    true_res = np.linalg.norm(k - ax) / np.linalg.norm(k)
    printed = float(field(r.stdout, "True relative residual"))
    assert printed == pytest.approx(true_res, rel=1e-3)
    assert "Solution is correct!" in r.stdout
    d = json.loads(rec.read_text())
    assert d["rhs"] == "random:12345" and d["incorrect_elements"] == 0
    assert d["true_relative_residual"] == pytest.approx(true_res, rel=1e-3)
    h = d["residual_history"]
    assert len(h) == it + 1 and h[0] == 1.0 and h[-1] == res
    assert all(v > 0 for v in h)
    assert d["matvec_bytes"] > 0 and d["matvec_gbps"] > 0 and "peak_gbps" not in d
    # a tolerance the solution cannot meet fails the check (strict exit code)
    bad = run(host_cg, case.path, dict(CG_MATRIX_FORMAT=fmt, CG_RHS="random:12345", CG_MAX_ITER=3,
                                       CG_STRICT_CHECK=1))
    assert bad.returncode == 2 and "Solution is incorrect!" in bad.stdout


def test_rhs_variable_errors(host_cg, tmp_path):
    fx, rn = next(p for p in RUNS if p[0]["case"] == "lap2d_16")
    case = Case(fx, str(tmp_path))
    for v in ("bogus", "random:", "random:-3", "random:1x"):
        r = run(host_cg, case.path, dict(CG_RHS=v))
        assert r.returncode == 1 and r.stderr.strip() == "Invalid value for CG_RHS! (ones, or random:<seed>)", v
    ones = run(host_cg, case.path, dict(CG_RHS="ones"))
    assert ones.returncode == 0 and "Solution is correct!" in ones.stdout

