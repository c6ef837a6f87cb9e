"""The C++ host's Matrix Market reader and converters against the oracle's.

Builds tests/host_dump.cpp with the product's Matrix.cpp/Preconditioner.cpp
(g++, a few seconds; plain, with AddressSanitizer + UndefinedBehaviorSanitizer
and with ThreadSanitizer, every test running against each) and compares every array bit for bit on the golden
fixtures' matrices, including the symmetric one (mirroring order), at several
thread counts (CG_IO_THREADS), plus the reader's error messages, its decoding
of values against Python's correctly rounded float(), line-ending and
piece-boundary cases, and the binary cache (CG_MATRIX_CACHE).
"""
import os
import random
import struct
import subprocess

import numpy as np
import pytest

from oracle import oracle
from tests.cases import SANITIZERS, Case
from tests.conftest import REPO, load_golden

HOST = os.path.join(REPO, "cgxx_amd", "host")


@pytest.fixture(scope="module", params=["plain", "asan-ubsan", "tsan"])
def dump_bin(request, tmp_path_factory):
    """The reader tool, plain and under each sanitizer: every test below runs
    against all three builds."""
    out = str(tmp_path_factory.mktemp("hostdump") / f"host_dump_{request.param}")
    subprocess.run(["g++", "-std=c++17", *SANITIZERS[request.param], "-pthread", "-I" + HOST, "-o", out,
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


def run_dump(dump_bin, path, out, threads=None, cache=None, check=True):
    env = dict(os.environ)
    env.pop("CG_IO_THREADS", None)
    if threads is not None:
        env["CG_IO_THREADS"] = str(threads)
    args = [dump_bin, str(path), str(out)] + ([str(cache)] if cache else [])
    return subprocess.run(args, capture_output=True, text=True, env=env, check=check)


@pytest.mark.parametrize("threads", [1, 3, 16])
@pytest.mark.parametrize("fx", CASES, ids=[fx["case"] for fx in CASES])
def test_reader_matches_oracle(fx, threads, tmp_path, dump_bin):
    case = Case(fx, str(tmp_path))
    out = str(tmp_path / "dump.bin")
    run_dump(dump_bin, case.path, out, threads)
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
@pytest.mark.parametrize("threads", [1, 4])
def test_reader_errors(text, msg, threads, tmp_path, dump_bin):
    path = tmp_path / "bad.mtx"
    path.write_text(text)
    r = run_dump(dump_bin, path, tmp_path / "o.bin", threads, check=False)
    assert r.returncode == 1
    assert r.stderr.strip() == msg


def diag_file(path, texts, line_end="\n", final_newline=True, extra=""):
    """A general n x n file whose row i holds the single entry (i, i) written
    as texts[i]."""
    n = len(texts)
    lines = ["%%MatrixMarket matrix coordinate real general", "% comment", f"{n} {n} {n}"]
    lines += [f"{i + 1} {i + 1} {t}" for i, t in enumerate(texts)]
    body = line_end.join(lines) + (line_end if final_newline else "") + extra
    path.write_bytes(body.encode())


VALUE_TEXTS = [
    "0", "-0", "+3", "1", "-1", "6", "0.1", ".5", "5.", "1.50", "0.000123", "1E5", "1e+05", "1e-05",
    "123456789012345", "1234567890123456", "12345678901234567890", "9007199254740993",
    "0.30000000000000004", "1e22", "1e23", "-1e-22", "1e-23", "2.2250738585072014e-308",
    "4.9e-324", "1.7976931348623157e308", "1e-400", "3.14159265358979323846264338327950288",
    "100000000000000000000000", "0.000000000000000000000000000001", "7e-10", "1.000000000000001",
]


def random_texts(rng, count):
    out = []
    for _ in range(count):
        v = rng.uniform(-1, 1) * 10.0 ** rng.randint(-30, 30)
        form = rng.randrange(5)
        if form == 0:
            out.append(repr(v))
        elif form == 1:
            out.append("%.*g" % (rng.randint(1, 17), v))
        elif form == 2:
            out.append("%.*e" % (rng.randint(0, 16), v))
        elif form == 3:
            out.append("%.*f" % (rng.randint(0, 20), v / 10.0 ** rng.randint(0, 25)))
        else:
            out.append(str(rng.randint(-10 ** 18, 10 ** 18)))
    return out


def midpoint_texts(rng, count):
    """Decimal texts at and next to the exact midpoint of two neighbouring
    doubles, with 16 to 19 significant digits: the cases where rounding an
    intermediate result twice can go wrong."""
    import decimal
    import math
    ctx = decimal.Context(prec=2000)
    out = []
    for _ in range(count):
        d = rng.uniform(1, 10) * 10.0 ** rng.randint(-24, 24)
        mid = ctx.divide(ctx.add(decimal.Decimal(d), decimal.Decimal(math.nextafter(d, math.inf))), 2)
        digits = rng.randint(16, 19)
        q = decimal.Decimal(1).scaleb(mid.adjusted() - digits + 1)
        for rounding in (decimal.ROUND_FLOOR, decimal.ROUND_CEILING, decimal.ROUND_HALF_EVEN):
            out.append(str(mid.quantize(q, rounding=rounding, context=ctx)))
    # exact midpoints that fit in 19 digits: small integers between doubles
    for _ in range(count // 10):
        base = rng.randint(2 ** 53, 2 ** 60)
        step = 1 << (base.bit_length() - 53)
        lo = base - base % step
        out.append(str(lo + step // 2))
    return out


@pytest.mark.parametrize("threads", [1, 5])
def test_reader_values_correctly_rounded(threads, tmp_path, dump_bin):
    """Every value equals Python's correctly rounded float() of its text,
    whichever of the short cuts and strtod decodes it."""
    rng = random.Random(20261202)
    texts = VALUE_TEXTS + random_texts(rng, 3000) + midpoint_texts(rng, 1000)
    path = tmp_path / "values.mtx"
    diag_file(path, texts)
    out = tmp_path / "o.bin"
    run_dump(dump_bin, path, out, threads)
    n, nz, w, (ptr, idx, val), ell, C = read_dump(str(out))
    want = np.array([float(t) for t in texts])
    assert n == len(texts) and np.array_equal(idx, np.arange(n))
    bad = np.nonzero(val.view(np.uint64) != want.view(np.uint64))[0]
    assert bad.size == 0, [(texts[i], val[i], want[i]) for i in bad[:5]]


@pytest.mark.parametrize("threads", [1, 2, 7, 64])
@pytest.mark.parametrize("variant", ["lf", "crlf", "no_final_newline", "trailing_lines", "tabs_extra"])
def test_reader_line_forms(variant, threads, tmp_path, dump_bin):
    """Line endings, a last line without newline, lines after the nz-th, tabs
    and extra tokens; more threads than lines leaves empty pieces."""
    texts = [str(1.0 + i / 8) for i in range(23)]
    path = tmp_path / "m.mtx"
    if variant == "lf":
        diag_file(path, texts)
    elif variant == "crlf":
        diag_file(path, texts, line_end="\r\n")
    elif variant == "no_final_newline":
        diag_file(path, texts, final_newline=False)
    elif variant == "trailing_lines":
        diag_file(path, texts, extra="\n% trailing\n1 1 not-a-number\n")
    else:
        n = len(texts)
        body = f"%%MatrixMarket matrix coordinate real general\n{n} {n} {n}\n"
        body += "".join(f"\t{i + 1}\t {i + 1}  {t} extra tokens\n" for i, t in enumerate(texts))
        path.write_text(body)
    out = tmp_path / "o.bin"
    run_dump(dump_bin, path, out, threads)
    n, nz, w, (ptr, idx, val), ell, C = read_dump(str(out))
    assert (n, nz) == (23, 23)
    assert np.array_equal(val, np.array([float(t) for t in texts]))


@pytest.mark.parametrize("threads", [1, 3])
@pytest.mark.parametrize("text", [
    "1 1 1\n2 2\n3 3 3\n",          # missing value mid-file
    "1 1 1\n2 2 2\n3 3\n",          # missing value on the last line
    "1 1 1\n\n2 2 2\n3 3 3\n",     # empty line among the entries
    "1 1 1\n2 2 .\n3 3 3\n",        # no digits
    "1 1 1\n2 2 2\n",                # fewer lines than nz
    "1 1 1\n2 0 2\n3 3 3\n",        # index below 1
    "1 1 1\n2 99999999999999999999 2\n3 3 3\n",  # index overflow
])
def test_reader_malformed_entries(text, threads, tmp_path, dump_bin):
    path = tmp_path / "bad.mtx"
    path.write_text("%%MatrixMarket matrix coordinate real general\n3 3 3\n" + text)
    r = run_dump(dump_bin, path, tmp_path / "o.bin", threads, check=False)
    assert r.returncode == 1 and r.stderr.strip() == "An exception occurred while reading the matrix!"


def test_reader_symmetric_missing_diagonal(tmp_path, dump_bin):
    path = tmp_path / "sym.mtx"
    path.write_text("%%MatrixMarket matrix coordinate real symmetric\n3 3 3\n1 1 4\n2 1 -1\n3 3 4\n")
    for threads in (1, 2):
        r = run_dump(dump_bin, path, tmp_path / "o.bin", threads, check=False)
        assert r.returncode == 1 and r.stderr.strip() == "An exception occurred while reading the matrix!"


def test_thread_count_validation(tmp_path, dump_bin):
    path = tmp_path / "m.mtx"
    diag_file(path, ["1", "2"])
    for bad in ("0", "-2", "abc", "257"):
        r = run_dump(dump_bin, path, tmp_path / "o.bin", bad, check=False)
        assert r.returncode == 1 and r.stderr.strip() == "Invalid value for CG_IO_THREADS!"


def test_matrix_cache(tmp_path, dump_bin):
    """The first read writes the cache, the second loads it (shown by a
    source replaced under the same size and time stamp, which only the
    cache can answer), a changed time stamp re-reads the source, and a cache
    with an index out of range is ignored."""
    fx = next((f for f in CASES if f["symmetric"]), CASES[0])
    case = Case(fx, str(tmp_path))
    cache = tmp_path / "m.cache"
    first, second = tmp_path / "a.bin", tmp_path / "b.bin"
    run_dump(dump_bin, case.path, first, 3, cache)
    assert cache.exists()
    run_dump(dump_bin, case.path, second, 5, cache)
    assert first.read_bytes() == second.read_bytes()

    # Same size and time stamp, different content: the cache answers.
    texts_a = ["1.25"] * 40
    texts_b = ["2.50"] * 40
    p = tmp_path / "swap.mtx"
    diag_file(p, texts_a)
    run_dump(dump_bin, p, first, 2, tmp_path / "swap.cache")
    st = os.stat(p)
    diag_file(p, texts_b)
    os.utime(p, ns=(st.st_atime_ns, st.st_mtime_ns))
    run_dump(dump_bin, p, second, 2, tmp_path / "swap.cache")
    assert read_dump(str(second))[3][2][0] == 1.25
    # A new time stamp: the source is read again and the cache rewritten.
    os.utime(p, ns=(st.st_atime_ns, st.st_mtime_ns + 1_000_000_000))
    run_dump(dump_bin, p, second, 2, tmp_path / "swap.cache")
    assert read_dump(str(second))[3][2][0] == 2.5

    # Corrupt an index in the cache: it is not used, the source is parsed.
    raw = bytearray((tmp_path / "swap.cache").read_bytes())
    hdr = 8 + 4 + 4 + 8 + 8 + 8 + 8 + 8
    raw[hdr: hdr + 4] = struct.pack("<i", 1 << 30)
    (tmp_path / "swap.cache").write_bytes(bytes(raw))
    r = run_dump(dump_bin, p, second, 2, tmp_path / "swap.cache")
    assert r.returncode == 0 and read_dump(str(second))[3][2][0] == 2.5

    # Another file (other inode) of the same size and time stamp does not use the cache.
    other = tmp_path / "other.mtx"
    diag_file(other, ["3.75"] * 40)
    os.utime(other, ns=(st.st_atime_ns, st.st_mtime_ns + 1_000_000_000))
    run_dump(dump_bin, other, second, 2, tmp_path / "swap.cache")
    assert read_dump(str(second))[3][2][0] == 3.75


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
    # COO is accepted (it runs on the device CRS); the run then stops later,
    # at the missing file or, without a GPU, at the device.
    env = dict(os.environ, CG_MATRIX_FORMAT="coo")
    r = subprocess.run([exe, "x.mtx"], capture_output=True, text=True, env=env)
    assert r.returncode == 1 and "matrix format" not in r.stderr
