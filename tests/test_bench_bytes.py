"""The algorithmic byte model of the bench's kernel-trace and traffic tables.

Each Clenshaw step reads its operator values once (the FP16 row records,
4 ceil(D/2) bytes per row, or with the symmetric storage the upper
diagonals only, D bytes per row) and its vectors: the first step r and
B_{j-1} (8 bytes per row), the second also B_{j+1} (12), the plain and
last steps B_{j+1}, r, B_{j+2} and B_j (16).  The kernel names are the
demangled ones rocprofv3 reports; template arguments FINAL, K, MODE, HALF,
NT, ROWS, SYM.
"""
import importlib.util
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


META = {"n": 1000, "nnz": 7000, "ndiag": 6}


@pytest.mark.parametrize("mode, vec", [(1, 8), (2, 12), (0, 16)])
@pytest.mark.parametrize("sym, values", [("false", 12), ("true", 6)])
@pytest.mark.parametrize("nt, rows", [("false", 2), ("true", 4)])
def test_clenshaw_step_bytes(bench, mode, vec, sym, values, nt, rows):
    name = f"k_poly_step<false, 2, {mode}, true, {nt}, {rows}, {sym}>"
    assert bench.algorithmic_bytes(name, META) == (values + vec) * META["n"]


@pytest.mark.parametrize("sym, values", [("false", 12), ("true", 6)])
def test_last_step_bytes(bench, sym, values):
    name = f"k_poly_step<true, 2, 0, true, true, 4, {sym}>"
    assert bench.algorithmic_bytes(name, META) == (values + 16) * META["n"]


def test_records_without_the_trailing_arguments(bench):
    # instances whose trailing template arguments are defaults print without them
    assert bench.algorithmic_bytes("k_poly_step<false, 2, 0, true, false>", META) == 28 * META["n"]


def test_other_kernels_unchanged(bench):
    n = META["n"]
    assert bench.algorithmic_bytes("k_mp_matvec_dot<2>", META) == 4 * 6 * n + 8 * n
    assert bench.algorithmic_bytes("k_pcg_xpay<true>", META) == 48 * n
    assert bench.algorithmic_bytes("k_unknown", META) is None
