"""ctypes binding of the CPU oracle (liboracle.so, built by oracle/Makefile).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg -- never by the product.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class COO(C.Structure):
    _fields_ = [
        ("N", C.c_int),
        ("nz", C.c_int),
        ("I", C.POINTER(C.c_int)),
        ("J", C.POINTER(C.c_int)),
        ("V", C.POINTER(C.c_double)),
        ("nzPerRow", C.POINTER(C.c_int)),
    ]


class Problem(C.Structure):
    _fields_ = [
        ("format", C.c_int),
        ("precond", C.c_int),
        ("order", C.c_int),
        ("nthreads", C.c_int),
        ("max_iter", C.c_int),
        ("tol", C.c_double),
        ("N", C.c_int),
        ("nz", C.c_int),
        ("coo", C.POINTER(COO)),
        ("ptr", C.c_void_p),
        ("index", C.c_void_p),
        ("value", C.c_void_p),
        ("ell_width", C.c_int),
        ("length", C.c_void_p),
        ("ell_index", C.c_void_p),
        ("ell_data", C.c_void_p),
        ("C", C.c_void_p),
    ]


class Result(C.Structure):
    _fields_ = [("iterations", C.c_int), ("residual", C.c_double)]


_lib = None


def build() -> None:
    """Compile liboracle.so (and the reference binaries where the sources exist)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_mm_read.argtypes = [C.c_char_p, C.POINTER(COO)]
        L.orc_coo_free.argtypes = [C.POINTER(COO)]
        L.orc_max_nz.argtypes = [C.POINTER(COO)]
        L.orc_coo_to_crs.argtypes = [C.POINTER(COO), _i32p, _i32p, _f64p]
        L.orc_coo_to_ell.argtypes = [C.POINTER(COO), C.c_int, _i32p, _i32p, _f64p]
        L.orc_jacobi_init.argtypes = [C.POINTER(COO), _f64p]
        L.orc_rhs_ones.argtypes = [C.POINTER(COO), _f64p]
        L.orc_check.argtypes = [C.c_int, _f64p, C.c_double]
        L.orc_spmv_crs.argtypes = [C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p]
        L.orc_spmv_ell.argtypes = [C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p]
        L.orc_spmv_coo.argtypes = [C.POINTER(COO), _f64p, _f64p]
        L.orc_axpy.argtypes = [C.c_int, C.c_double, _f64p, _f64p]
        L.orc_xpay.argtypes = [C.c_int, _f64p, C.c_double, _f64p]
        L.orc_jacobi_apply.argtypes = [C.c_int, _f64p, _f64p, _f64p]
        L.orc_dot_serial.argtypes = [C.c_int, _f64p, _f64p]
        L.orc_dot_serial.restype = C.c_double
        L.orc_dot_tree.argtypes = [C.c_int, _f64p, _f64p]
        L.orc_dot_tree.restype = C.c_double
        L.orc_vec_grid.argtypes = [C.c_int]
        L.orc_crs_rowblocks.argtypes = [C.c_int, _i32p, _i32p]
        L.orc_rowblock_dot_tree.argtypes = [C.c_int, _i32p, _f64p, _f64p]
        L.orc_rowblock_dot_tree.restype = C.c_double
        L.orc_rowblock_dot_tree_walk.argtypes = [C.c_int, _i32p, _f64p, _f64p, C.c_int, C.c_int]
        L.orc_rowblock_dot_tree_walk.restype = C.c_double
        L.orc_rowblock_dot_tree_groups.argtypes = [C.c_int, _i32p, _f64p, _f64p]
        L.orc_rowblock_dot_tree_groups.restype = C.c_double
        L.orc_pcg.argtypes = [C.POINTER(Problem), _f64p, _f64p, C.POINTER(Result)] + [C.c_void_p] * 5
        _lib = L
    return _lib


GEOMETRY = {"block": 256, "vec_gmax": 2048, "rows_max": 256, "nnz_cap": 2048, "spmv_gmax": 2048,
            "ell_gmax": 0, "dia_max": 8}


class MMFile:
    """A Matrix Market file read by the oracle's restatement of the reader."""

    def __init__(self, path: str):
        self._coo = COO()
        rc = lib().orc_mm_read(os.fsencode(path), C.byref(self._coo))
        if rc:
            raise ValueError(f"orc_mm_read failed with code {rc}")
        self.N = self._coo.N
        self.nz = self._coo.nz

    def __del__(self):
        if _lib is not None and getattr(self, "_coo", None) is not None:
            _lib.orc_coo_free(C.byref(self._coo))

    @property
    def coo(self):
        n = self.nz
        I = np.ctypeslib.as_array(self._coo.I, shape=(n,)).copy()
        J = np.ctypeslib.as_array(self._coo.J, shape=(n,)).copy()
        V = np.ctypeslib.as_array(self._coo.V, shape=(n,)).copy()
        return I, J, V

    def crs(self):
        ptr = np.zeros(self.N + 1, np.int32)
        idx = np.zeros(max(self.nz, 1), np.int32)
        val = np.zeros(max(self.nz, 1), np.float64)
        lib().orc_coo_to_crs(C.byref(self._coo), ptr, idx, val)
        return ptr, idx[: self.nz], val[: self.nz]

    def ell(self):
        w = lib().orc_max_nz(C.byref(self._coo))
        length = np.zeros(self.N, np.int32)
        idx = np.zeros(max(self.N * w, 1), np.int32)
        dat = np.zeros(max(self.N * w, 1), np.float64)
        lib().orc_coo_to_ell(C.byref(self._coo), w, length, idx, dat)
        return w, length, idx, dat

    def jacobi(self):
        c = np.zeros(self.N)
        missing = lib().orc_jacobi_init(C.byref(self._coo), c)
        return c, missing

    def rhs_ones(self):
        k = np.zeros(self.N)
        lib().orc_rhs_ones(C.byref(self._coo), k)
        return k


def pcg(*, n, fmt, precond, order, k, x0=None, tol=1e-9, max_iter=1000, nthreads=1,
        crs=None, ell=None, mm: MMFile | None = None, C_jac=None, history=True):
    """Run the oracle's PCG; returns (x, iterations, residual, history dict).

    fmt: "COO" (needs mm), "CRS" (crs=(ptr, index, value)), "ELL"
    (ell=(width, length, index, data)); precond: "none" | "jacobi";
    order: 0 serial, 1 fused HIP, 2 HIP hooks, 3 OpenMP.
    """
    L = lib()
    p = Problem()
    p.format = {"COO": 0, "CRS": 1, "ELL": 2}[fmt]
    p.precond = {"none": 0, "jacobi": 1}[precond]
    p.order = order
    p.nthreads = nthreads
    p.max_iter = max_iter
    p.tol = tol
    p.N = n
    keep = []
    if fmt == "COO":
        p.coo = C.pointer(mm._coo)
        p.nz = mm.nz
    elif fmt == "CRS":
        ptr, idx, val = (np.ascontiguousarray(a) for a in crs)
        keep += [ptr, idx, val]
        p.ptr, p.index, p.value = ptr.ctypes.data, idx.ctypes.data, val.ctypes.data
        p.nz = int(ptr[-1])
    else:
        w, length, idx, dat = ell
        length, idx, dat = (np.ascontiguousarray(a) for a in (length, idx, dat))
        keep += [length, idx, dat]
        p.ell_width = w
        p.length, p.ell_index, p.ell_data = length.ctypes.data, idx.ctypes.data, dat.ctypes.data
    if precond == "jacobi":
        cj = np.ascontiguousarray(C_jac, dtype=np.float64)
        keep.append(cj)
        p.C = cj.ctypes.data
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
    k = np.ascontiguousarray(k, dtype=np.float64)
    hist = {key: np.full(max_iter + 1, np.nan) for key in ("r2", "rho", "pq", "alpha", "beta")}
    ptrs = [hist[key].ctypes.data if history else None for key in ("r2", "rho", "pq", "alpha", "beta")]
    res = Result()
    rc = L.orc_pcg(C.byref(p), k, x, C.byref(res), *ptrs)
    if rc:
        raise ValueError(f"orc_pcg failed with code {rc}")
    it = res.iterations
    h = {
        "r2": hist["r2"][: it + 1],
        "rho": hist["rho"][: it + 1],
        "pq": hist["pq"][:it],
        "alpha": hist["alpha"][:it],
        "beta": hist["beta"][:it],
    }
    return x, it, res.residual, h


class MixedResult(C.Structure):
    _fields_ = [("iterations", C.c_int), ("updates", C.c_int), ("true_residual", C.c_double),
                ("inner_residual", C.c_double)]


class Poly(C.Structure):
    _fields_ = [("degree", C.c_int), ("c", C.c_float * 17), ("two_over_b", C.c_float),
                ("fp32_values", C.c_int)]


def make_poly(coeffs, b, fp32_values=False):
    """orc_poly from Chebyshev coefficients on [0, b], rounded as the library
    does; fp32_values: the Clenshaw steps read the FP32 operator values
    (cghip_set_poly_values 32) instead of the FP16 copy of a DIA operator."""
    pl = Poly()
    pl.fp32_values = int(fp32_values)
    pl.degree = len(coeffs) - 1
    if not 1 <= pl.degree <= 16:
        raise ValueError("degree must be in [1, 16]")
    for j, cj in enumerate(coeffs):
        pl.c[j] = float(np.float32(cj))
    pl.two_over_b = float(np.float32(2.0 / b))
    return pl


def mixed_pcg(*, n, crs, k, precond="jacobi", C_jac=None, x0=None, tol=1e-9, max_iter=100000,
              delta=0.0, history=True, poly=None, poly_fp32_values=False):
    """The oracle's FP32-inner / FP64 reliable-update PCG (CRS).

    precond "jacobi" and "poly" run on the scaled operator
    D^-1/2 A D^-1/2 built from the matrix's own diagonal (C_jac is not
    used); "poly" needs poly=(coeffs, b) (Chebyshev coefficients on [0, b]).
    Returns (x, info dict, history dict)."""
    L = lib()
    if not hasattr(L, "_mixed_bound"):
        L.orc_mixed_pcg_poly.argtypes = [C.POINTER(Problem), C.c_void_p, _f64p, _f64p, C.c_double,
                                         C.POINTER(MixedResult)] + [C.c_void_p] * 5
        L.orc_vec_grid4.argtypes = [C.c_int]
        L._mixed_bound = True
    p = Problem()
    p.format = 1
    p.precond = {"none": 0, "jacobi": 1, "poly": 2}[precond]
    pl = None
    if precond == "poly":
        if poly is None:
            raise ValueError("precond 'poly' needs poly=(coeffs, b)")
        pl = make_poly(*poly, fp32_values=poly_fp32_values)
    p.max_iter = max_iter
    p.tol = tol
    p.N = n
    ptr, idx, val = (np.ascontiguousarray(a) for a in crs)
    p.ptr, p.index, p.value = ptr.ctypes.data, idx.ctypes.data, val.ctypes.data
    p.nz = int(ptr[-1])
    keep = [ptr, idx, val]
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
    k = np.ascontiguousarray(k, dtype=np.float64)
    hist = {key: np.full(max_iter + 1, np.nan) for key in ("r2", "rho", "pq", "alpha", "beta")}
    ptrs = [hist[key].ctypes.data if history else None for key in ("r2", "rho", "pq", "alpha", "beta")]
    res = MixedResult()
    rc = L.orc_mixed_pcg_poly(C.byref(p), C.cast(C.byref(pl), C.c_void_p) if pl else None, k, x,
                              delta, C.byref(res), *ptrs)
    if rc:
        raise ValueError(f"orc_mixed_pcg failed with code {rc}")
    it = res.iterations
    h = {"r2": hist["r2"][: it + 1], "rho": hist["rho"][: it + 1], "pq": hist["pq"][:it],
         "alpha": hist["alpha"][:it], "beta": hist["beta"][:it]}
    info = {"iterations": it, "updates": res.updates, "true_residual": res.true_residual,
            "inner_residual": res.inner_residual}
    return x, info, h


def poly_apply(crs, r32, coeffs, b, fp32_values=False):
    """z = p(A_hat) r in FP32 on the scaled operator A_hat = D^-1/2 A D^-1/2,
    the library's expression order."""
    L = lib()
    ptr = np.ascontiguousarray(crs[0], dtype=np.int32)
    idx = np.ascontiguousarray(crs[1], dtype=np.int32)
    val = np.ascontiguousarray(crs[2], dtype=np.float64)
    n = ptr.size - 1
    r = np.ascontiguousarray(r32, dtype=np.float32)
    z = np.zeros(n, dtype=np.float32)
    pl = make_poly(coeffs, b, fp32_values)
    L.orc_poly_apply.argtypes = [C.c_int] + [C.c_void_p] * 6
    rc = L.orc_poly_apply(n, ptr.ctypes.data, idx.ctypes.data, val.ctypes.data, r.ctypes.data,
                          C.cast(C.byref(pl), C.c_void_p), z.ctypes.data)
    if rc:
        raise ValueError(f"orc_poly_apply failed with code {rc}")
    return z


def op32_apply(crs, x32, half=False):
    """(y, is_dia): y = A_hat x in FP32 on the scaled operator
    D^-1/2 A D^-1/2 (unit diagonal implicit), with the FP32 values or, with
    half, the FP16-rounded ones the Clenshaw steps read when the library
    stores A_hat in DIA."""
    L = lib()
    ptr = np.ascontiguousarray(crs[0], dtype=np.int32)
    idx = np.ascontiguousarray(crs[1], dtype=np.int32)
    val = np.ascontiguousarray(crs[2], dtype=np.float64)
    n = ptr.size - 1
    x = np.ascontiguousarray(x32, dtype=np.float32)
    y = np.zeros(n, dtype=np.float32)
    dia = C.c_int(0)
    L.orc_op32_apply.argtypes = [C.c_int] + [C.c_void_p] * 4 + [C.c_int, C.c_void_p, C.c_void_p]
    rc = L.orc_op32_apply(n, ptr.ctypes.data, idx.ctypes.data, val.ctypes.data, x.ctypes.data,
                          int(half), y.ctypes.data, C.byref(dia))
    if rc:
        raise ValueError(f"orc_op32_apply failed with code {rc}")
    return y, bool(dia.value)
