"""The cg_hip executable end to end, against the reference's own report.

The reference's serial binary, built from its sources in the development
container, printed its report for the three matrices below in CRS, ELL and
COO; tests/golden/cli/reference_output.json holds every line with the
timing values blanked (tests/golden/make_golden.py cli).  No reference
binary is needed here.  With CG_HIP_FUSED=0 CG_HIP_DOT=serial, cg_hip runs
the reference's loop hook by hook in the reference's summation order, so
every printed line except the timings must be identical to those.  The
default fused device loop must reach the same iteration count (+-1) and
pass the same check.
"""
import os
import subprocess

import pytest

from tests.cases import load_cli_fixture, untimed_report, write_cli_matrices
from tests.conftest import REPO

pytestmark = pytest.mark.gpu

EXE = os.path.join(REPO, "cgxx_amd", "bin", "cg_hip")
REFERENCE = load_cli_fixture()


def run(exe, path, **env):
    reps = env.pop("_reps", None)
    e = dict(os.environ)
    e.update({k: str(v) for k, v in env.items()})
    print(f"run {os.path.basename(exe)} {os.path.basename(path)} {env}", flush=True)
    args = [exe, path] + ([str(reps)] if reps is not None else [])
    return subprocess.run(args, capture_output=True, text=True, env=e, timeout=300)


def reference_lines(name, fmt):
    return REFERENCE["cases"][name]["runs"][fmt]


@pytest.fixture(scope="module")
def mtx(tmp_path_factory):
    return write_cli_matrices(str(tmp_path_factory.mktemp("cli")))


@pytest.mark.parametrize("name", ["lap2d_256", "rand_sym_40", "diff3d_24"])
@pytest.mark.parametrize("fmt", ["CRS", "ELL", "COO"])
def test_reference_order_output_identical(mtx, name, fmt):
    env = dict(CG_MATRIX_FORMAT=fmt, **REFERENCE["env"])
    ours = run(EXE, mtx[name], CG_HIP_FUSED=0, CG_HIP_DOT="serial", **env)
    assert ours.returncode == 0, ours.stderr
    assert untimed_report(ours.stdout, mtx[name]) == reference_lines(name, fmt)


@pytest.mark.parametrize("name", ["lap2d_256", "rand_sym_40", "diff3d_24"])
@pytest.mark.parametrize("fmt", ["CRS", "ELL", "COO"])
def test_fused_matches_reference_run(mtx, name, fmt):
    env = dict(CG_MATRIX_FORMAT=fmt, **REFERENCE["env"])
    ours = run(EXE, mtx[name], **env)
    assert ours.returncode == 0, ours.stderr

    def field(lines, label):
        return [ln.split(":", 1)[1].strip() for ln in lines if ln.startswith(label)][0]

    ref_it = int(field(reference_lines(name, fmt), "Iterations"))
    assert abs(int(field(ours.stdout.splitlines(), "Iterations")) - ref_it) <= 1
    assert "Solution is correct!" in reference_lines(name, fmt)
    assert "Solution is correct!" in ours.stdout
    # every timing category is reported as a number
    for label in ("Solve time", "MatVec time", "axpy time", "xpay time", "MatVec GFLOP/s"):
        float(field(ours.stdout.splitlines(), label))
    assert float(field(ours.stdout.splitlines(), "MatVec time")) > 0


def test_repetitions_and_strict_check(mtx):
    ours = run(EXE, mtx["lap2d_256"], _reps=2, CG_MATRIX_FORMAT="CRS")
    assert ours.returncode == 0
    assert ours.stdout.count("Solution is correct!") == 2
    assert ours.stdout.count("Iterations:") == 2
    bad = run(EXE, mtx["lap2d_256"], CG_MAX_ITER=5, CG_STRICT_CHECK=1)
    assert bad.returncode == 2 and "elements are incorrect!" in bad.stdout
    lenient = run(EXE, mtx["lap2d_256"], CG_MAX_ITER=5)
    assert lenient.returncode == 0 and "elements are incorrect!" in lenient.stdout


@pytest.mark.parametrize("poly", ["none", "learned", "fixed"])
@pytest.mark.parametrize("name,fmt", [("lap2d_256", "CRS"), ("diff3d_24", "CRS"), ("diff3d_24", "COO"),
                                      ("rand_sym_40", "COO")])
def test_mixed_precision_cli(mtx, name, fmt, poly):
    """CG_HIP_MIXED=1: the FP32-inner solver reaches the FP64 true residual
    (the printed residual) and passes the reference's check; CG_HIP_POLY adds
    the polynomial preconditioner (learned: the committed network).  COO
    files run on the same device CRS (sorted rows: the file's own arrays;
    the symmetric file: the stable conversion)."""
    env = dict(CG_MATRIX_FORMAT=fmt, CG_PRECONDITIONER="jacobi", CG_MAX_ITER=20000,
               CG_HIP_MIXED=1, CG_HIP_POLY=poly, CG_STRICT_CHECK=1)
    ours = run(EXE, mtx[name], **env)
    assert ours.returncode == 0, ours.stderr + ours.stdout
    assert "Solution is correct!" in ours.stdout
    res = [ln.split(":", 1)[1].strip() for ln in ours.stdout.splitlines() if ln.startswith("Residual")]
    assert float(res[0]) <= 1e-9
    if poly != "none":
        assert "HIP mixed precision, polynomial (" + poly + ")" in ours.stdout


def test_mixed_precision_cli_rejects(mtx):
    ell = run(EXE, mtx["lap2d_256"], CG_MATRIX_FORMAT="ELL", CG_HIP_MIXED=1)
    assert ell.returncode == 1 and "CG_HIP_MIXED needs CG_MATRIX_FORMAT=CRS" in ell.stderr
    nomix = run(EXE, mtx["lap2d_256"], CG_HIP_POLY="learned")
    assert nomix.returncode == 1 and "needs CG_HIP_MIXED=1" in nomix.stderr
    bad = run(EXE, mtx["lap2d_256"], CG_HIP_MIXED=1, CG_HIP_POLY="cubic")
    assert bad.returncode == 1 and "Invalid value for CG_HIP_POLY" in bad.stderr
    nomodel = run(EXE, mtx["lap2d_256"], CG_MATRIX_FORMAT="CRS", CG_HIP_MIXED=1,
                  CG_HIP_POLY="learned", CG_HIP_MODEL="/nonexistent/model.bin")
    assert nomodel.returncode == 1 and "cannot open model file" in nomodel.stderr


def field(stdout, label, index=0):
    vals = [ln.split(":", 1)[1].strip() for ln in stdout.splitlines() if ln.startswith(label + ":")]
    return vals[index]


@pytest.mark.parametrize("mode", ["fused", "hooks", "mixed", "learned"])
def test_solve_record(mtx, mode, tmp_path):
    """CG_REPORT_JSON: one JSON line per repetition carrying the printed
    report's figures unrounded, plus what the backend ran."""
    rec = tmp_path / "solves.jsonl"
    env = dict(CG_MATRIX_FORMAT="CRS", CG_PRECONDITIONER="jacobi", CG_REPORT_JSON=str(rec), _reps=2)
    if mode == "hooks":
        env["CG_HIP_FUSED"] = 0
    if mode in ("mixed", "learned"):
        env.update(CG_HIP_MIXED=1, CG_MAX_ITER=20000)
    if mode == "learned":
        env["CG_HIP_POLY"] = "learned"
    ours = run(EXE, mtx["diff3d_24"], **env)
    assert ours.returncode == 0, ours.stderr
    import json
    lines = rec.read_text().splitlines()
    assert len(lines) == 2
    for i, ln in enumerate(lines):
        r = json.loads(ln)
        assert r["repetition"] == i and r["matrix"] == mtx["diff3d_24"]
        assert (r["rows"], r["nonzeros"]) == tuple(int(v) for v in field(ours.stdout, "# rows / # nonzeros", i).split(" / "))
        assert r["format"] == "CRS" and r["preconditioner"] == "jacobi"
        assert r["iterations"] == int(field(ours.stdout, "Iterations", i))
        assert "%e" % r["residual"] == field(ours.stdout, "Residual", i)
        assert r["incorrect_elements"] == 0
        assert abs(r["times"]["solve"] - float(field(ours.stdout, "Solve time", i))) <= 1e-6
        assert abs(r["times"]["io"] - float(field(ours.stdout, "IO time", i))) <= 1e-6
        # per-iteration residuals of the loop's recursion and the SpMV rate
        h = r["residual_history"]
        if mode != "learned":
            assert len(h) == r["iterations"] + 1 and h[0] == 1.0
            assert all(v > 0 for v in h)
        if mode in ("fused", "hooks"):
            assert h[-1] == r["residual"]
        assert r["rhs"] == "ones" and r["matvec_bytes"] > 0
        assert r["peak_gbps"] == 8000.0
        if mode != "learned":  # the learned driver does not time its kernels: null
            assert 0 < r["matvec_peak_fraction"] < 2
        b = r["backend"]
        assert b["name"] == "hip" and b["dot"] == "tree"
        assert b["loop"] == {"fused": "fused", "hooks": "hooks", "mixed": "mixed", "learned": "mixed"}[mode]
        if mode == "mixed":
            assert b["mixed"]["polynomial"] == "none" and b["mixed"]["true_residual"] == r["residual"]
            assert b["mixed"]["inner_iterations"] == r["iterations"]
        if mode == "learned":
            m = b["mixed"]
            assert m["polynomial"] == "learned" and m["method"] in ("poly", "poly+jacobi", "jacobi (refused)",
                                                                   "jacobi-selected")
            assert m["poly_iterations"] + m["jacobi_iterations"] == r["iterations"]


def test_solve_record_unwritable(mtx, tmp_path):
    ours = run(EXE, mtx["lap2d_256"], CG_REPORT_JSON=str(tmp_path / "no" / "such" / "dir.jsonl"))
    assert ours.returncode == 1 and "Could not write the solve record" in ours.stderr


@pytest.mark.parametrize("mode", ["fused", "mixed"])
def test_random_rhs_cli(mtx, mode, tmp_path):
    """CG_RHS=random:<seed> on the device: the FP64 true relative residual
    (computed on the host from the returned x) passes the check, and the
    record names the right-hand side."""
    rec = tmp_path / "solves.jsonl"
    env = dict(CG_MATRIX_FORMAT="CRS", CG_PRECONDITIONER="jacobi", CG_RHS="random:7",
               CG_REPORT_JSON=str(rec), CG_MAX_ITER=20000, CG_STRICT_CHECK=1)
    if mode == "mixed":
        env["CG_HIP_MIXED"] = 1
    ours = run(EXE, mtx["diff3d_24"], **env)
    assert ours.returncode == 0, ours.stderr + ours.stdout
    tr = float(field(ours.stdout, "True relative residual"))
    assert tr <= 1e-8 and "Solution is correct!" in ours.stdout
    import json
    r = json.loads(rec.read_text())
    assert r["rhs"] == "random:7" and r["true_relative_residual"] == pytest.approx(tr, rel=1e-6)

