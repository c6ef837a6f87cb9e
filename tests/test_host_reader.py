"""The C++ host's Matrix Market reader and converters against the oracle's.

Builds tests/host_dump.cpp with the product's Matrix.cpp/Preconditioner.cpp
(g++, a few seconds) and compares every array bit for bit on the golden
fixtures' matrices, including the symmetric one (mirroring order), plus the
reader's error messages.
"""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle
from tests.cases import Case
from tests.conftest import REPO, load_golden

HOST = os.path.join(REPO, "cgxx_amd", "host")


@pytest.fixture(scope="module")
def dump_bin(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("hostdump") / "host_dump")
    subprocess.run(["g++", "-std=c++17", "-O2", "-I" + HOST, "-o", out,
                    os.path.join(REPO, "tests", "host_dump.cpp"),
                    os.path.join(HOST, "Matrix.cpp"), os.path.join(HOST, "Preconditioner.cpp")],
                   check=True)
    return out


def read_dump(path):
    raw = open(path, "rb").read()
    n, nz, w = np.frombuffer(raw[:12], np.int32)
    off = 12

    def take(dtype, count):
        nonlocal off
        a = np.frombuffer(raw[off: off + count * np.dtype(dtype).itemsize], dtype)
        off += count * np.dtype(dtype).itemsize
        return a

    ptr = take(np.int32, n + 1)
    idx = take(np.int32, nz)
    val = take(np.float64, nz)
    length = take(np.int32, n)
    eidx = take(np.int32, n * w)
    edat = take(np.float64, n * w)
    C = take(np.float64, n)
    return n, nz, w, (ptr, idx, val), (length, eidx, edat), C


CASES = [fx for fx in load_golden() if fx["N"] <= 200_000]


@pytest.mark.parametrize("fx", CASES, ids=[fx["case"] for fx in CASES])
def test_reader_matches_oracle(fx, tmp_path, dump_bin):
    case = Case(fx, str(tmp_path))
    out = str(tmp_path / "dump.bin")
    subprocess.run([dump_bin, case.path, out], check=True)
    n, nz, w, crs, ell, C = read_dump(out)
    assert (n, nz) == (case.n, case.nz)
    for a, b in zip(crs, case.crs):
        assert np.array_equal(np.asarray(a).view(np.uint8), np.asarray(b).view(np.uint8))
    ow, olen, oidx, odat = case.ell
    assert w == ow
    assert np.array_equal(ell[0], olen)
    assert np.array_equal(ell[1], oidx) and np.array_equal(ell[2].view(np.uint64), odat.view(np.uint64))
    assert np.array_equal(C.view(np.uint64), case.C.view(np.uint64))


@pytest.mark.parametrize("text,msg", [
    ("%%MatrixMarket matrix coordinate pattern general\n2 2 1\n1 1\n",
     "Only supporting real matrices in coordinate format!"),
    ("%%MatrixMarket matrix coordinate real skew-symmetric\n2 2 1\n1 1 1\n",
     "Only supporting general or symmetric matrices!"),
    ("%%MatrixMarket matrix coordinate real general\n2 3 1\n1 1 1\n", "Need a square matrix!"),
    ("%%MatrixMarket matrix coordinate real general\n2 2 2\n1 1 1\n",
     "An exception occurred while reading the matrix!"),
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1\n",
     "An exception occurred while reading the matrix!"),
])
def test_reader_errors(text, msg, tmp_path, dump_bin):
    path = tmp_path / "bad.mtx"
    path.write_text(text)
    r = subprocess.run([dump_bin, str(path), str(tmp_path / "o.bin")], capture_output=True, text=True)
    assert r.returncode == 1
    assert r.stderr.strip() == msg


def test_missing_file(tmp_path, dump_bin):
    r = subprocess.run([dump_bin, str(tmp_path / "none.mtx"), str(tmp_path / "o.bin")],
                       capture_output=True, text=True)
    assert r.returncode == 1 and r.stderr.strip() == "Can't open file with matrix!"


def test_cli_and_environment_errors():
    """Usage and CG_* validation happen before any device call."""
    exe = os.path.join(REPO, "cgxx_amd", "bin", "cg_hip")
    if not os.path.exists(exe):
        pytest.skip("cg_hip not built")
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 1 and r.stderr.startswith("Usage: ")
    cases = {
        "CG_MAX_ITER": ("0", "Invalid value for CG_MAX_ITER!"),
        "CG_TOLERANCE": ("-1", "Invalid value for CG_TOLERANCE!"),
        "CG_MATRIX_FORMAT": ("CSR", "Invalid value for CG_MATRIX_FORMAT! (COO, CRS, or ELL)"),
        "CG_PRECONDITIONER": ("ilu", "Invalid value for CG_PRECONDITIONER! (none, or jacobi)"),
        "CG_WORK_DISTRIBUTION": ("col", "Invalid value for CG_WORK_DISTRIBUTION! (row, or nz)"),
        "CG_OVERLAPPED_GATHER": ("1", "No support for overlapped gather!"),
    }
    for var, (val, msg) in cases.items():
        env = dict(os.environ, **{var: val})
        r = subprocess.run([exe, "x.mtx"], capture_output=True, text=True, env=env)
        assert r.returncode == 1 and r.stderr.strip() == msg, (var, r.stderr)
    env = dict(os.environ, CG_MATRIX_FORMAT="coo")
    r = subprocess.run([exe, "x.mtx"], capture_output=True, text=True, env=env)
    assert r.returncode == 1 and r.stderr.strip() == "No support for this matrix format!"
