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

from tests.cases import Case
from tests.conftest import REPO, golden_runs

HOST = os.path.join(REPO, "cgxx_amd", "host")
REF = os.path.join(REPO, "oracle", "_ref", "cg_serial")
RUNS = golden_runs(max_n=70_000)
TIMED = re.compile(r"^(IO time|Converting time|Solve time|Check time|Total time \(excl\. IO\)|MatVec time|"
                   r"MatVec GFLOP/s|axpy time|xpay time|vectorDot time|Preconditioner time):")


@pytest.fixture(scope="module")
def host_cg(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("hostcg") / "host_cg")
    srcs = [os.path.join(REPO, "tests", "host_cg.cpp")] + [
        os.path.join(HOST, f) for f in ("CG.cpp", "Matrix.cpp", "Preconditioner.cpp", "main.cpp")]
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-pthread", "-I" + HOST, "-o", out,
                    *srcs], check=True)
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
