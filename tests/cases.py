"""Shared helpers: rebuild a golden fixture's inputs and compare histories."""
from __future__ import annotations

import json
import os
import re

import numpy as np

from cgxx_amd import problems
from oracle import oracle

# Matrices larger than this are rebuilt in memory instead of through a .mtx
# file: the generator's rows are column-sorted, so the reader's CRS of a
# general file is exactly the generator's CRS (pinned on the small cases).
MM_LIMIT = 200_000


class Case:
    """Inputs of one golden fixture, in every storage format."""

    def __init__(self, fx: dict, tmpdir: str):
        self.fx = fx
        m = getattr(problems, fx["generator"])(**fx["kwargs"])
        self.gen = m
        self.n = m.n
        self.mm = None
        if fx["N"] <= MM_LIMIT:
            path = os.path.join(tmpdir, fx["case"] + ".mtx")
            sha = problems.write_mm(path, m, symmetric=fx["symmetric"])
            assert sha == fx["mtx_sha256"], "generator or writer changed"
            self.path = path
            self.mm = oracle.MMFile(path)
            self.crs = self.mm.crs()
            self.ell = self.mm.ell()
            self.C, missing = self.mm.jacobi()
            assert missing == 0
            self.k = self.mm.rhs_ones()
        else:
            assert not fx["symmetric"]
            self.path = None
            self.crs = (m.ptr, m.index, m.value)
            self.ell = to_ell(m.ptr, m.index, m.value)
            self.C = 1.0 / m.diagonal()
            self.k = m.ones_rhs()
        self.nz = int(self.crs[0][-1])


def to_ell(ptr, index, value):
    """Column-major ELL (slot j*N + i) of a CRS matrix, padding = 0."""
    n = ptr.size - 1
    lens = np.diff(ptr).astype(np.int32)
    w = int(lens.max()) if n else 0
    idx = np.zeros(n * w, np.int32)
    dat = np.zeros(n * w, np.float64)
    starts = ptr[:-1].astype(np.int64)
    rows = np.arange(n, dtype=np.int64)
    for j in range(w):
        live = lens > j
        idx[j * n + rows[live]] = index[starts[live] + j]
        dat[j * n + rows[live]] = value[starts[live] + j]
    return w, lens, idx, dat


def run_env(run: dict):
    env = run["env"]
    return (env["CG_MATRIX_FORMAT"], env["CG_PRECONDITIONER"],
            int(env.get("CG_MAX_ITER", 1000)), float(env.get("CG_TOLERANCE", 1e-9)))


def hexs(a) -> list[float]:
    return [float.fromhex(v) for v in a]


def golden_sequences(run: dict) -> dict:
    return {
        "r2": np.array(hexs(run["r2"])),
        "rho0": float.fromhex(run["rho0"]),
        "pq": np.array(hexs(run["dot_pq"])),
        "alpha": np.array(hexs(run["a"])),
        "beta": np.array(hexs(run["b"])),
        "x_head": np.array(hexs(run["x_head"])),
    }


def assert_bitwise(a, b, what):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    assert a.shape == b.shape, f"{what}: shape {a.shape} != {b.shape}"
    same = (a.view(np.uint64) == b.view(np.uint64)) | (np.isnan(a) & np.isnan(b))
    if not same.all():
        i = int(np.flatnonzero(~same)[0])
        raise AssertionError(f"{what}: first difference at {i}: {a[i]!r} vs {b[i]!r}")


def max_rel(a, b) -> float:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    n = min(a.size, b.size)
    if n == 0:
        return 0.0
    d = np.abs(a[:n] - b[:n]) / np.maximum(np.abs(b[:n]), np.finfo(float).tiny)
    return float(d.max())


# ---- the reference's printed report (tests/golden/cli) -------------------

TIMED = re.compile(r"^(IO time|Converting time|Transfer to time|Solve time|Transfer from time|"
                   r"Check time|Total time \(excl\. IO\)|MatVec time|MatVec GFLOP/s|axpy time|"
                   r"xpay time|vectorDot time|Preconditioner time):")
CLI_FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cli",
                           "reference_output.json")


def load_cli_fixture() -> dict:
    with open(CLI_FIXTURE) as f:
        return json.load(f)


def untimed_report(stdout: str, path: str) -> list[str]:
    """A CG report's lines with every timing value blanked ("<t>"), the
    device-only transfer lines dropped and the matrix path as <matrix>: the
    form of the fixture written by tests/golden/make_golden.py cli."""
    out = []
    for ln in stdout.splitlines():
        if ln.startswith("Transfer to time:") or ln.startswith("Transfer from time:"):
            continue  # only device backends transfer
        out.append(TIMED.match(ln).group(1) + ": <t>" if TIMED.match(ln) else ln)
    return [ln.replace(path, "<matrix>") for ln in out]


def write_cli_matrices(tmpdir: str) -> dict:
    """The fixture's matrices as .mtx files (checked against their sha256)."""
    files = {}
    for name, case in load_cli_fixture()["cases"].items():
        p = os.path.join(tmpdir, f"{name}.mtx")
        m = getattr(problems, case["generator"])(**case["kwargs"])
        assert problems.write_mm(p, m, symmetric=case["symmetric"]) == case["mtx_sha256"]
        files[name] = p
    return files


# ---- sanitizer builds of the C++ host tools --------------------------------

# Compiler flags of each build variant of the host test tools: the plain
# build, AddressSanitizer + UndefinedBehaviorSanitizer (any report aborts the
# run, so a test that expects exit 0 or a given message fails), and
# ThreadSanitizer for the parallel reader.
SANITIZERS = {
    "plain": ["-O2"],
    "asan-ubsan": ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
                   "-fno-sanitize-recover=all"],
    "tsan": ["-O1", "-g", "-fsanitize=thread"],
}
