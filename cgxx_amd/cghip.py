"""ctypes binding of the C ABI in include/cghip.h (libcghip.so).

The library is built in-tree by ``cgxx_amd.build`` (``__graft_entry__.build``)
into ``cgxx_amd/lib/libcghip.so``.  Loading fails loudly when it is missing:
there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libcghip.so")

# CG::Vector (CG.h:22-35)
VEC = {"K": 0, "X": 1, "P": 2, "Q": 3, "R": 4, "Z": 5}
FORMAT = {"CRS": 1, "ELL": 2}
PRECOND = {"none": 0, "jacobi": 1, "poly": 2}
DOT = {"tree": 0, "serial": 1}

SYMBOLS = [
    "cghip_create", "cghip_destroy", "cghip_last_error", "cghip_geometry",
    "cghip_set_size", "cghip_upload_crs", "cghip_upload_ell", "cghip_upload_jacobi",
    "cghip_set_format", "cghip_upload_vector", "cghip_download_vector", "cghip_set_dot_mode",
    "cghip_vector_pointer", "cghip_dot_noalloc",
    "cghip_cpy", "cghip_matvec", "cghip_axpy", "cghip_xpay", "cghip_vector_dot",
    "cghip_apply_preconditioner", "cghip_solve", "cghip_history", "cghip_time_matvec", "cghip_time_hook",
    "cghip_fill_vector", "cghip_true_residual", "cghip_prepare_fp32", "cghip_solve_mixed",
    "cghip_set_poly", "cghip_set_poly_values", "cghip_matrix_stats", "cghip_time_poly_step", "cghip_fp32_layout",
    "cghip_poly_values_layout",
    "cghip_mixed_iteration_bytes",
    "cghip_mlp_param_count", "cghip_mlp_create", "cghip_mlp_destroy", "cghip_mlp_last_error",
    "cghip_mlp_set_params", "cghip_mlp_get_params", "cghip_mlp_forward", "cghip_mlp_gradient",
    "cghip_mlp_train",
]


class CGHIPError(RuntimeError):
    pass


class AdamOpts(C.Structure):
    _fields_ = [("lr", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float), ("eps", C.c_float)]


class SolveOpts(C.Structure):
    _fields_ = [("precond", C.c_int), ("max_iter", C.c_int), ("tol", C.c_double),
                ("batch", C.c_int), ("profile", C.c_int)]


class SolveInfo(C.Structure):
    _fields_ = [("iterations", C.c_int), ("residual", C.c_double), ("r2", C.c_double),
                ("wall_s", C.c_double), ("ms_matvec", C.c_double), ("ms_update", C.c_double),
                ("ms_xpay", C.c_double), ("ms_init", C.c_double), ("n_matvec", C.c_int),
                ("n_update", C.c_int), ("n_xpay", C.c_int), ("n_init", C.c_int)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class MixedOpts(C.Structure):
    _fields_ = [("precond", C.c_int), ("max_iter", C.c_int), ("tol", C.c_double),
                ("delta", C.c_double), ("batch", C.c_int), ("profile", C.c_int)]


class MixedInfo(C.Structure):
    _fields_ = [("iterations", C.c_int), ("updates", C.c_int), ("true_residual", C.c_double),
                ("inner_residual", C.c_double), ("wall_s", C.c_double),
                ("ms_matvec", C.c_double), ("ms_update", C.c_double), ("ms_xpay", C.c_double),
                ("ms_refresh", C.c_double), ("ms_poly", C.c_double), ("n_matvec", C.c_int),
                ("n_update", C.c_int), ("n_xpay", C.c_int), ("n_refresh", C.c_int),
                ("n_poly", C.c_int), ("restarts", C.c_int)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_lib = None


def load() -> C.CDLL:
    """Load libcghip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise CGHIPError(f"{LIB_PATH} is missing: run __graft_entry__.build() (no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.cghip_create.argtypes = [C.c_int, C.POINTER(vp)]
    L.cghip_destroy.argtypes = [vp]
    L.cghip_last_error.argtypes = [vp]
    L.cghip_last_error.restype = C.c_char_p
    L.cghip_geometry.argtypes = [C.POINTER(C.c_int)] * 7
    L.cghip_set_size.argtypes = [vp, C.c_int]
    L.cghip_upload_crs.argtypes = [vp, C.c_int, C.c_int, _i32p, _i32p, _f64p]
    L.cghip_upload_ell.argtypes = [vp, C.c_int, C.c_int, _i32p, _i32p, _f64p]
    L.cghip_upload_jacobi.argtypes = [vp, _f64p]
    L.cghip_set_format.argtypes = [vp, C.c_int]
    L.cghip_upload_vector.argtypes = [vp, C.c_int, _f64p]
    L.cghip_download_vector.argtypes = [vp, C.c_int, _f64p]
    L.cghip_set_dot_mode.argtypes = [vp, C.c_int]
    L.cghip_vector_pointer.argtypes = [vp, C.c_int, C.POINTER(vp)]
    L.cghip_dot_noalloc.argtypes = [vp, vp, C.c_int, vp, vp]
    L.cghip_cpy.argtypes = [vp, C.c_int, C.c_int]
    L.cghip_matvec.argtypes = [vp, C.c_int, C.c_int]
    L.cghip_axpy.argtypes = [vp, C.c_double, C.c_int, C.c_int]
    L.cghip_xpay.argtypes = [vp, C.c_int, C.c_double, C.c_int]
    L.cghip_vector_dot.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_double)]
    L.cghip_apply_preconditioner.argtypes = [vp, C.c_int, C.c_int]
    L.cghip_solve.argtypes = [vp, C.POINTER(SolveOpts), C.POINTER(SolveInfo)]
    L.cghip_history.argtypes = [vp, C.c_int, _f64p, C.c_int, C.POINTER(C.c_int)]
    L.cghip_time_matvec.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double)]
    L.cghip_time_hook.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_double)]
    L.cghip_fill_vector.argtypes = [vp, C.c_int, C.c_double]
    L.cghip_true_residual.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.cghip_prepare_fp32.argtypes = [vp]
    L.cghip_set_poly.argtypes = [vp, C.c_int, _f64p, C.c_double]
    L.cghip_set_poly_values.argtypes = [vp, C.c_int]
    L.cghip_matrix_stats.argtypes = [vp, _f64p, C.c_int]
    L.cghip_time_poly_step.argtypes = [vp, C.c_int, C.POINTER(C.c_double)]
    L.cghip_fp32_layout.argtypes = [vp, C.POINTER(C.c_int)]
    L.cghip_poly_values_layout.argtypes = [vp, C.POINTER(C.c_int)]
    L.cghip_mixed_iteration_bytes.argtypes = [vp, C.c_int, C.POINTER(C.c_double)]
    _f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
    L.cghip_mlp_param_count.argtypes = [C.c_int] * 4
    L.cghip_mlp_create.argtypes = [C.c_int] * 6 + [C.POINTER(vp)]
    L.cghip_mlp_destroy.argtypes = [vp]
    L.cghip_mlp_last_error.argtypes = [vp]
    L.cghip_mlp_last_error.restype = C.c_char_p
    L.cghip_mlp_set_params.argtypes = [vp, _f32p]
    L.cghip_mlp_get_params.argtypes = [vp, _f32p]
    L.cghip_mlp_forward.argtypes = [vp, _f32p, C.c_int, _f32p]
    L.cghip_mlp_gradient.argtypes = [vp, _f32p, _f32p, C.c_int, _f32p, C.POINTER(C.c_double)]
    L.cghip_mlp_train.argtypes = [vp, _f32p, _f32p, C.c_int, C.c_int, C.POINTER(AdamOpts), vp]
    L.cghip_solve_mixed.argtypes = [vp, C.POINTER(MixedOpts), C.POINTER(MixedInfo)]
    _lib = L
    return L


def geometry() -> dict:
    vals = [C.c_int() for _ in range(7)]
    load().cghip_geometry(*[C.byref(v) for v in vals])
    return dict(zip(["block", "vec_gmax", "rows_max", "nnz_cap", "spmv_gmax", "ell_gmax",
                     "dia_max"], [v.value for v in vals]))


class Context:
    """One device context: the GPU-side state of a CG backend."""

    def __init__(self, device: int = 0):
        self.L = load()
        self._h = C.c_void_p()
        rc = self.L.cghip_create(device, C.byref(self._h))
        if rc:
            raise CGHIPError(f"cghip_create failed ({rc}): {self.L.cghip_last_error(None).decode()}")
        self.n = 0

    def close(self):
        if self._h:
            self.L.cghip_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _chk(self, rc: int, what: str):
        if rc:
            raise CGHIPError(f"{what} failed ({rc}): {self.L.cghip_last_error(self._h).decode()}")

    # data
    def set_size(self, n: int):
        self._chk(self.L.cghip_set_size(self._h, n), "cghip_set_size")
        self.n = n

    def upload_crs(self, ptr, index, value):
        ptr = np.ascontiguousarray(ptr, np.int32)
        index = np.ascontiguousarray(index, np.int32)
        value = np.ascontiguousarray(value, np.float64)
        n = ptr.size - 1
        if index.size == 0:
            index, value = np.zeros(1, np.int32), np.zeros(1, np.float64)
        self._chk(self.L.cghip_upload_crs(self._h, n, int(ptr[-1]), ptr, index, value), "cghip_upload_crs")
        self.n = n

    def upload_ell(self, width, length, index, data):
        length = np.ascontiguousarray(length, np.int32)
        index = np.ascontiguousarray(index, np.int32)
        data = np.ascontiguousarray(data, np.float64)
        n = length.size
        if index.size == 0:
            index, data = np.zeros(1, np.int32), np.zeros(1, np.float64)
        self._chk(self.L.cghip_upload_ell(self._h, n, int(width), length, index, data), "cghip_upload_ell")
        self.n = n

    def upload_jacobi(self, c):
        self._chk(self.L.cghip_upload_jacobi(self._h, np.ascontiguousarray(c, np.float64)), "cghip_upload_jacobi")

    def set_format(self, fmt: str):
        self._chk(self.L.cghip_set_format(self._h, FORMAT[fmt]), "cghip_set_format")

    def set_dot_mode(self, mode: str):
        self._chk(self.L.cghip_set_dot_mode(self._h, DOT[mode]), "cghip_set_dot_mode")

    def upload(self, vec: str, a):
        a = np.ascontiguousarray(a, np.float64)
        assert a.size == self.n
        self._chk(self.L.cghip_upload_vector(self._h, VEC[vec], a), "cghip_upload_vector")

    def pointer(self, vec: str) -> int:
        """Device address of a vector (cghip_vector_pointer)."""
        out = C.c_void_p()
        self._chk(self.L.cghip_vector_pointer(self._h, VEC[vec], C.byref(out)), "cghip_vector_pointer")
        return out.value

    def download(self, vec: str) -> np.ndarray:
        out = np.empty(self.n, np.float64)
        self._chk(self.L.cghip_download_vector(self._h, VEC[vec], out), "cghip_download_vector")
        return out

    # hooks (CG.h:224-237)
    def cpy(self, dst, src):
        self._chk(self.L.cghip_cpy(self._h, VEC[dst], VEC[src]), "cghip_cpy")

    def matvec(self, x, y):
        self._chk(self.L.cghip_matvec(self._h, VEC[x], VEC[y]), "cghip_matvec")

    def axpy(self, a, x, y):
        self._chk(self.L.cghip_axpy(self._h, a, VEC[x], VEC[y]), "cghip_axpy")

    def xpay(self, x, a, y):
        self._chk(self.L.cghip_xpay(self._h, VEC[x], a, VEC[y]), "cghip_xpay")

    def dot(self, a, b) -> float:
        out = C.c_double()
        self._chk(self.L.cghip_vector_dot(self._h, VEC[a], VEC[b], C.byref(out)), "cghip_vector_dot")
        return out.value

    def apply_preconditioner(self, x, y):
        self._chk(self.L.cghip_apply_preconditioner(self._h, VEC[x], VEC[y]), "cghip_apply_preconditioner")

    # fused solver
    def solve(self, precond="jacobi", max_iter=1000, tol=1e-9, batch=0, profile=False) -> dict:
        o = SolveOpts(PRECOND[precond], max_iter, tol, batch, int(profile))
        info = SolveInfo()
        self._chk(self.L.cghip_solve(self._h, C.byref(o), C.byref(info)), "cghip_solve")
        return info.as_dict()

    def prepare_fp32(self):
        self._chk(self.L.cghip_prepare_fp32(self._h), "cghip_prepare_fp32")

    def mixed_iteration_bytes(self, precond: str) -> float:
        """Algorithmic bytes of one mixed-precision iteration (cost model)."""
        v = C.c_double()
        self._chk(self.L.cghip_mixed_iteration_bytes(self._h, PRECOND[precond], C.byref(v)),
                  "cghip_mixed_iteration_bytes")
        return v.value

    def fp32_layout(self) -> str | None:
        """Layout of the scaled FP32 operator: "CRS", "ELL", "DIA" or None."""
        v = C.c_int()
        self._chk(self.L.cghip_fp32_layout(self._h, C.byref(v)), "cghip_fp32_layout")
        return {0: "CRS", 1: "ELL", 2: "DIA"}.get(v.value)

    def poly_values_layout(self) -> str | None:
        """Values the Clenshaw steps read: "FP32", "FP16" (DIA row records),
        "FP16-SYM" (upper diagonals only) or None."""
        v = C.c_int()
        self._chk(self.L.cghip_poly_values_layout(self._h, C.byref(v)), "cghip_poly_values_layout")
        return {0: "FP32", 1: "FP16", 2: "FP16-SYM"}.get(v.value)

    def set_poly(self, coeffs, b: float):
        """Polynomial preconditioner p(D^-1 A) D^-1, p = sum_j coeffs[j] T_j(2 l / b - 1)."""
        c = np.ascontiguousarray(coeffs, dtype=np.float64)
        self._chk(self.L.cghip_set_poly(self._h, int(c.size - 1), c, float(b)), "cghip_set_poly")

    def set_poly_values(self, bits: int):
        """Precision (16 or 32) of the DIA operator values the Clenshaw
        steps read; drops the polynomial (set it again)."""
        self._chk(self.L.cghip_set_poly_values(self._h, int(bits)), "cghip_set_poly_values")

    def time_poly_step(self, reps: int = 5) -> float:
        """Average ms of one Clenshaw step (back-to-back applications)."""
        ms = C.c_double()
        self._chk(self.L.cghip_time_poly_step(self._h, reps, C.byref(ms)), "cghip_time_poly_step")
        return ms.value

    def matrix_stats(self) -> np.ndarray:
        """36 statistics of the CRS matrix (include/cghip.h, cghip_matrix_stats)."""
        out = np.zeros(36)
        self._chk(self.L.cghip_matrix_stats(self._h, out, out.size), "cghip_matrix_stats")
        return out

    def solve_mixed(self, precond="jacobi", max_iter=100000, tol=1e-9, delta=0.0, batch=0,
                    profile=False) -> dict:
        o = MixedOpts(PRECOND[precond], max_iter, tol, delta, batch, int(profile))
        info = MixedInfo()
        self._chk(self.L.cghip_solve_mixed(self._h, C.byref(o), C.byref(info)), "cghip_solve_mixed")
        return info.as_dict()

    def history(self, which: str, count: int) -> np.ndarray:
        idx = {"r2": 0, "rho": 1, "pq": 2, "alpha": 3, "beta": 4}[which]
        out = np.full(max(count, 1), np.nan)
        w = C.c_int()
        self._chk(self.L.cghip_history(self._h, idx, out, count, C.byref(w)), "cghip_history")
        return out[: w.value]

    def fill(self, vec: str, value: float):
        self._chk(self.L.cghip_fill_vector(self._h, VEC[vec], value), "cghip_fill_vector")

    def true_residual(self, x="X", b="K") -> tuple[float, float]:
        """(||b - A x||, ||b||) in FP64 on the device; overwrites Q."""
        nr, nb = C.c_double(), C.c_double()
        self._chk(self.L.cghip_true_residual(self._h, VEC[x], VEC[b], C.byref(nr), C.byref(nb)),
                  "cghip_true_residual")
        return nr.value, nb.value

    def time_matvec(self, x="P", y="Q", reps=20) -> float:
        ms = C.c_double()
        self._chk(self.L.cghip_time_matvec(self._h, VEC[x], VEC[y], reps, C.byref(ms)), "cghip_time_matvec")
        return ms.value

    HOOKS = {"matvec": 0, "axpy": 1, "xpay": 2, "dot": 3, "jacobi": 4, "cpy": 5}

    def time_hook(self, hook: str, reps=20) -> float:
        """Average kernel time (ms) of one hook's device work (cghip_time_hook)."""
        ms = C.c_double()
        self._chk(self.L.cghip_time_hook(self._h, self.HOOKS[hook], reps, C.byref(ms)), "cghip_time_hook")
        return ms.value


def reference_loop(ctx: Context, precond: str, max_iter=1000, tol=1e-9):
    """CG::solve (CG.cpp:284-385) driven hook by hook through the C ABI.

    This is the reference's loop structure unchanged (8 synchronous hooks and
    3 host round trips per iteration); returns (iterations, residual, history).
    """
    import math

    hist = {"r2": [], "rho": [], "pq": [], "alpha": [], "beta": []}
    ctx.matvec("X", "R")
    ctx.xpay("K", -1.0, "R")
    if precond == "none":
        ctx.cpy("P", "R")
    else:
        ctx.apply_preconditioner("R", "P")
    r2 = ctx.dot("R", "R")
    hist["r2"].append(r2)
    nrm2_0 = math.sqrt(r2)
    rho = r2 if precond == "none" else ctx.dot("P", "R")
    hist["rho"].append(rho)
    residual = float("nan")
    it = 0
    while it < max_iter:
        ctx.matvec("P", "Q")
        dot_pq = ctx.dot("P", "Q")
        a = rho / dot_pq
        hist["pq"].append(dot_pq)
        hist["alpha"].append(a)
        ctx.axpy(a, "P", "X")
        ctx.axpy(-a, "Q", "R")
        r2 = ctx.dot("R", "R")
        hist["r2"].append(r2)
        residual = math.sqrt(r2) / nrm2_0
        if residual <= tol:
            it += 1
            break
        rho_old = rho
        if precond == "none":
            rho = r2
        else:
            ctx.apply_preconditioner("R", "Z")
            rho = ctx.dot("R", "Z")
        hist["rho"].append(rho)
        b = rho / rho_old
        hist["beta"].append(b)
        ctx.xpay("R" if precond == "none" else "Z", b, "P")
        it += 1
    return it, residual, {k: np.array(v) for k, v in hist.items()}


class Mlp:
    """The FP32 network of include/cghip_mlp.h: in -> h1 -> h2 -> out, ReLU."""

    def __init__(self, dims, max_rows: int, device: int = 0):
        self.L = load()
        self.dims = tuple(int(v) for v in dims)
        self.total = self.L.cghip_mlp_param_count(*self.dims)
        self._h = C.c_void_p()
        rc = self.L.cghip_mlp_create(device, *self.dims, int(max_rows), C.byref(self._h))
        if rc:
            raise CGHIPError(f"cghip_mlp_create failed ({rc}): "
                             f"{self.L.cghip_mlp_last_error(None).decode()}")

    def close(self):
        if self._h:
            self.L.cghip_mlp_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc:
            raise CGHIPError(f"{what} failed ({rc}): {self.L.cghip_mlp_last_error(self._h).decode()}")

    def set_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        if p.size != self.total:
            raise ValueError(f"expected {self.total} parameters, got {p.size}")
        self._chk(self.L.cghip_mlp_set_params(self._h, p), "cghip_mlp_set_params")

    def get_params(self) -> np.ndarray:
        p = np.zeros(self.total, np.float32)
        self._chk(self.L.cghip_mlp_get_params(self._h, p), "cghip_mlp_get_params")
        return p

    def forward(self, X) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.zeros((X.shape[0], self.dims[3]), np.float32)
        self._chk(self.L.cghip_mlp_forward(self._h, X, X.shape[0], Y), "cghip_mlp_forward")
        return Y

    def gradient(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        g = np.zeros(self.total, np.float32)
        loss = C.c_double()
        self._chk(self.L.cghip_mlp_gradient(self._h, X, T, X.shape[0], g, C.byref(loss)),
                  "cghip_mlp_gradient")
        return g, loss.value

    def train(self, X, T, steps: int, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        hist = np.zeros(max(steps, 1))
        o = AdamOpts(lr, beta1, beta2, eps)
        self._chk(self.L.cghip_mlp_train(self._h, X, T, X.shape[0], int(steps), C.byref(o),
                                         hist.ctypes.data), "cghip_mlp_train")
        return hist[:steps]
