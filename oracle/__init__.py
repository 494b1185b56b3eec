"""ctypes front end of the oracle (a CPU restatement of monolish's reference
CPU path; see oracle.h for what each function restates and how it is pinned).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

SUCCESS, SIZE_ERROR, MAXITER, BREAKDOWN, RESIDUAL_NAN, NOT_IMPL = 0, -1, -2, -3, -4, -10
FACTOR_ERROR = -11
PRECOND_DEFAULT, PRECOND_NONE, PRECOND_JACOBI, PRECOND_SOR = 0, 1, 2, 3
PRECOND_IC, PRECOND_ILU = 4, 5


def build() -> None:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


class _COO(C.Structure):
    _fields_ = [("rows", C.c_int64), ("cols", C.c_int64), ("nnz", C.c_int64),
                ("alloc", C.c_int64), ("row", C.POINTER(C.c_int32)),
                ("col", C.POINTER(C.c_int32)), ("val", C.POINTER(C.c_double))]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_P = C.c_void_p
_I64 = C.c_int64


def _declare(L):
    L.or_mt_seed.argtypes = [_P, C.c_uint32]
    L.or_mt_next.argtypes = [_P]
    L.or_mt_next.restype = C.c_uint32
    L.or_random_vector_d.argtypes = [_P, _I64, C.c_double, C.c_double, C.c_uint32]
    L.or_random_vector_s.argtypes = [_P, _I64, C.c_float, C.c_float, C.c_uint32]
    L.or_coo_free.argtypes = [C.POINTER(_COO)]
    L.or_laplacian_2d5p.argtypes = [C.POINTER(_COO), C.c_int, C.c_int]
    L.or_input_mm.argtypes = [C.POINTER(_COO), C.c_char_p, C.c_int]
    L.or_input_mm.restype = C.c_int
    L.or_coo_sort_arrays.argtypes = [_I64, _P, _P, _P, C.c_int]
    L.or_coo_sort_arrays.restype = _I64
    L.or_crs_from_coo.argtypes = [C.POINTER(_COO), _P, _P, _P]
    L.or_num_threads.restype = C.c_int
    L.or_set_sor_omega.argtypes = [C.c_double]
    L.or_get_sor_omega.restype = C.c_double
    L.or_factor_bad_row.restype = _I64
    for s, T in (("d", C.c_double), ("s", C.c_float)):
        getattr(L, f"or_crs_diag_{s}").argtypes = [_I64, _I64, _P, _P, _P, _P]
        getattr(L, f"or_spmv_{s}").argtypes = [_I64, _P, _P, _P, _P, _P]
        getattr(L, f"or_spmv_t_{s}").argtypes = [_I64, _I64, _P, _P, _P, _P, _P]
        for f in ("dot",):
            getattr(L, f"or_{f}_{s}").argtypes = [_I64, _P, _P]
            getattr(L, f"or_{f}_{s}").restype = T
        getattr(L, f"or_nrm2_{s}").argtypes = [_I64, _P]
        getattr(L, f"or_nrm2_{s}").restype = T
        for f in ("axpy", "xpay"):
            getattr(L, f"or_{f}_{s}").argtypes = [_I64, T, _P, _P]
        getattr(L, f"or_axpyz_{s}").argtypes = [_I64, T, _P, _P, _P]
        getattr(L, f"or_scal_{s}").argtypes = [_I64, T, _P]
        getattr(L, f"or_copy_{s}").argtypes = [_I64, _P, _P]
        for f in ("vadd", "vsub", "vmul"):
            getattr(L, f"or_{f}_{s}").argtypes = [_I64, _P, _P, _P]
        getattr(L, f"or_vreciprocal_{s}").argtypes = [_I64, _P, _P]
        getattr(L, f"or_residual_l2_{s}").argtypes = [_I64, _P, _P, _P, _P, _P]
        getattr(L, f"or_residual_l2_{s}").restype = T
        getattr(L, f"or_sor_precond_{s}").argtypes = [_I64, _P, _P, _P, _P, _P]
        getattr(L, f"or_ilu_factor_{s}").argtypes = [_I64, _P, _P, _P, C.c_int, _P]
        getattr(L, f"or_ilu_factor_{s}").restype = C.c_int
        getattr(L, f"or_ilu_precond_{s}").argtypes = [_I64, _P, _P, _P, C.c_int, _P, _P]
        getattr(L, f"or_ilu_precond_{s}").restype = C.c_int
        for f in ("cg", "bicgstab", "jacobi", "sor"):
            fn = getattr(L, f"or_{f}_{s}")
            fn.argtypes = [_I64, _P, _P, _P, _P, _P, C.c_int, C.c_double, _I64, _I64,
                           C.POINTER(_I64), C.POINTER(C.c_double), _P]
            fn.restype = C.c_int


def _sfx(dtype) -> str:
    dt = np.dtype(dtype)
    if dt == np.float64:
        return "d"
    if dt == np.float32:
        return "s"
    raise TypeError(f"unsupported dtype {dt}")


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


# ---------------------------------------------------------------- random ----
class MT19937:
    """std::mt19937 restated (oracle.c)."""

    def __init__(self, seed: int = 5489):
        self._buf = C.create_string_buffer(624 * 4 + 8)
        lib().or_mt_seed(self._buf, seed)

    def __call__(self) -> int:
        return lib().or_mt_next(self._buf)


def random_vector(n: int, vmin: float, vmax: float, seed: int, dtype=np.float64) -> np.ndarray:
    """vector<T>(N, min, max, seed) (src/utils/vector_constructor.cpp:98-113)."""
    out = np.empty(n, dtype=dtype)
    getattr(lib(), f"or_random_vector_{_sfx(dtype)}")(_ptr(out), n, vmin, vmax, seed)
    return out


# ------------------------------------------------------------------- COO ----
def _take_coo(c: _COO):
    nnz = c.nnz
    row = np.ctypeslib.as_array(c.row, shape=(nnz,)).copy() if nnz else np.zeros(0, np.int32)
    col = np.ctypeslib.as_array(c.col, shape=(nnz,)).copy() if nnz else np.zeros(0, np.int32)
    val = np.ctypeslib.as_array(c.val, shape=(nnz,)).copy() if nnz else np.zeros(0, np.float64)
    rows, cols = c.rows, c.cols
    lib().or_coo_free(C.byref(c))
    return rows, cols, row, col, val


def laplacian_2d5p(m: int, n: int):
    """util::laplacian_matrix_2D_5p(m, n) as COO (rows, cols, row, col, val)."""
    c = _COO()
    lib().or_laplacian_2d5p(C.byref(c), m, n)
    return _take_coo(c)


def input_mm(path: str, is_float: bool = False):
    """COO::input_mm(path) as (rows, cols, row, col, val); raises on rejection."""
    c = _COO()
    rc = lib().or_input_mm(C.byref(c), path.encode(), int(is_float))
    if rc != 0:
        raise ValueError(f"or_input_mm({path}) rejected with {rc}")
    return _take_coo(c)


def coo_sort(row, col, val, merge: bool):
    """COO::sort(merge) on triplets; returns sorted (row, col, val)."""
    row = _c(row, np.int32).copy()
    col = _c(col, np.int32).copy()
    val = _c(val, np.float64).copy()
    nnz = lib().or_coo_sort_arrays(len(row), _ptr(row), _ptr(col), _ptr(val), int(merge))
    return row[:nnz], col[:nnz], val[:nnz]


def crs_from_coo(rows, cols, row, col, val, dtype=np.float64):
    """CRS(COO&) (src/utils/convert/convert_crs.cpp:9-43)."""
    row = _c(row, np.int32)
    col = _c(col, np.int32)
    v = _c(val, np.float64)
    c = _COO()
    c.rows, c.cols, c.nnz, c.alloc = rows, cols, len(row), len(row)
    c.row = row.ctypes.data_as(C.POINTER(C.c_int32))
    c.col = col.ctypes.data_as(C.POINTER(C.c_int32))
    c.val = v.ctypes.data_as(C.POINTER(C.c_double))
    rp = np.zeros(rows + 1, np.int32)
    ci = np.zeros(max(len(row), 1), np.int32)
    vv = np.zeros(max(len(row), 1), np.float64)
    lib().or_crs_from_coo(C.byref(c), _ptr(rp), _ptr(ci), _ptr(vv))
    nnz = len(row)
    return rp, ci[:nnz].copy(), vv[:nnz].astype(dtype)


# ------------------------------------------------------------ BLAS / VML ----
def crs_diag(rows, cols, rp, ci, val):
    s = _sfx(val.dtype)
    d = np.zeros(min(rows, cols), dtype=val.dtype)
    getattr(lib(), f"or_crs_diag_{s}")(rows, cols, _ptr(_c(rp, np.int32)), _ptr(_c(ci, np.int32)),
                                       _ptr(val), _ptr(d))
    return d


def spmv(rp, ci, val, x):
    rows = len(rp) - 1
    y = np.zeros(rows, dtype=val.dtype)
    x = _c(x, val.dtype)
    getattr(lib(), f"or_spmv_{_sfx(val.dtype)}")(rows, _ptr(_c(rp, np.int32)), _ptr(_c(ci, np.int32)),
                                                 _ptr(val), _ptr(x), _ptr(y))
    return y


def spmv_t(rp, ci, val, x, cols):
    rows = len(rp) - 1
    y = np.zeros(cols, dtype=val.dtype)
    x = _c(x, val.dtype)
    getattr(lib(), f"or_spmv_t_{_sfx(val.dtype)}")(rows, cols, _ptr(_c(rp, np.int32)),
                                                   _ptr(_c(ci, np.int32)), _ptr(val), _ptr(x), _ptr(y))
    return y


def dot(x, y):
    x = np.ascontiguousarray(x)
    y = _c(y, x.dtype)
    return getattr(lib(), f"or_dot_{_sfx(x.dtype)}")(len(x), _ptr(x), _ptr(y))


def nrm2(x):
    x = np.ascontiguousarray(x)
    return getattr(lib(), f"or_nrm2_{_sfx(x.dtype)}")(len(x), _ptr(x))


def axpy(a, x, y):
    y = y.copy()
    getattr(lib(), f"or_axpy_{_sfx(y.dtype)}")(len(y), a, _ptr(_c(x, y.dtype)), _ptr(y))
    return y


def xpay(a, x, y):
    y = y.copy()
    getattr(lib(), f"or_xpay_{_sfx(y.dtype)}")(len(y), a, _ptr(_c(x, y.dtype)), _ptr(y))
    return y


def axpyz(a, x, y):
    z = np.empty_like(y)
    getattr(lib(), f"or_axpyz_{_sfx(y.dtype)}")(len(y), a, _ptr(_c(x, y.dtype)),
                                                _ptr(_c(y, y.dtype)), _ptr(z))
    return z


def scal(a, x):
    x = x.copy()
    getattr(lib(), f"or_scal_{_sfx(x.dtype)}")(len(x), a, _ptr(x))
    return x


def _vv(name, a, b):
    y = np.empty_like(a)
    getattr(lib(), f"or_{name}_{_sfx(a.dtype)}")(len(a), _ptr(_c(a, a.dtype)), _ptr(_c(b, a.dtype)), _ptr(y))
    return y


def vadd(a, b):
    return _vv("vadd", a, b)


def vsub(a, b):
    return _vv("vsub", a, b)


def vmul(a, b):
    return _vv("vmul", a, b)


def vreciprocal(a):
    y = np.empty_like(a)
    getattr(lib(), f"or_vreciprocal_{_sfx(a.dtype)}")(len(a), _ptr(_c(a, a.dtype)), _ptr(y))
    return y


def residual_l2(rp, ci, val, x, b):
    n = len(rp) - 1
    return getattr(lib(), f"or_residual_l2_{_sfx(val.dtype)}")(
        n, _ptr(_c(rp, np.int32)), _ptr(_c(ci, np.int32)), _ptr(val),
        _ptr(_c(x, val.dtype)), _ptr(_c(b, val.dtype)))


# --------------------------------------------------------------- solvers ----
def solve(method: str, rp, ci, val, x0, b, precond=PRECOND_JACOBI, tol=1e-8,
          miniter=0, maxiter=2**62, history=False):
    """Run equation::{CG,BiCGSTAB,Jacobi,SOR}::solve restated.  Returns
    (status, x, final_iter, final_resid, rhistory or None)."""
    n = len(rp) - 1
    dt = val.dtype
    x = _c(x0, dt).copy()
    fi = C.c_int64(0)
    fr = C.c_double(0.0)
    hist = None
    if history:
        hist = np.full(int(min(maxiter, 1_000_000)), np.nan)
    fn = getattr(lib(), f"or_{method}_{_sfx(dt)}")
    rc = fn(n, _ptr(_c(rp, np.int32)), _ptr(_c(ci, np.int32)), _ptr(val), _ptr(x),
            _ptr(_c(b, dt)), precond, float(tol), int(miniter), int(maxiter),
            C.byref(fi), C.byref(fr), _ptr(hist) if hist is not None else None)
    if hist is not None:
        hist = hist[: min(fi.value, len(hist))]
    return rc, x, fi.value, fr.value, hist


def set_sor_omega(w: float):
    """Relaxation coefficient of the SOR preconditioner / solver (default 1.9)."""
    lib().or_set_sor_omega(float(w))


def sor_precond(rp, ci, val, r):
    """SOR::create_precond + apply_precond (sor_kernel_precond) on r."""
    n = len(rp) - 1
    z = np.zeros(n, val.dtype)
    getattr(lib(), f"or_sor_precond_{_sfx(val.dtype)}")(n, _ptr(_c(rp, np.int32)),
                                                        _ptr(_c(ci, np.int32)), _ptr(val),
                                                        _ptr(_c(r, val.dtype)), _ptr(z))
    return z


def num_threads() -> int:
    return lib().or_num_threads()


def ilu_factor(rp, ci, val, kind):
    """IC::create_precond (kind PRECOND_IC) / ILU::create_precond
    (PRECOND_ILU): the factored copy of val.  Returns (status, f); status
    FACTOR_ERROR for unsorted columns, a missing diagonal or a zero pivot
    (factor_bad_row() names the row)."""
    n = len(rp) - 1
    f = np.zeros(max(int(rp[-1]) if n > 0 else 0, 1), val.dtype)
    rc = getattr(lib(), f"or_ilu_factor_{_sfx(val.dtype)}")(
        n, _ptr(_c(rp, np.int32)), _ptr(_c(ci, np.int32)), _ptr(val), int(kind), _ptr(f))
    return rc, f[: int(rp[-1]) if n > 0 else 0]


def ilu_precond(rp, ci, val, kind, r):
    """create_precond + apply_precond of IC / ILU on r.  Returns (status, z)."""
    n = len(rp) - 1
    z = np.zeros(n, val.dtype)
    rc = getattr(lib(), f"or_ilu_precond_{_sfx(val.dtype)}")(
        n, _ptr(_c(rp, np.int32)), _ptr(_c(ci, np.int32)), _ptr(val), int(kind),
        _ptr(_c(r, val.dtype)), _ptr(z))
    return rc, z


def factor_bad_row() -> int:
    """Row at which the last IC / ILU set-up failed (-1: none)."""
    return int(lib().or_factor_bad_row())


# ---------------------------------------------------------------- LOBPCG ----
def _chol_fail_order(B):
    """Order (1-based) of the first leading minor of B (lower triangle) that
    is not positive definite, 0 when B is positive definite."""
    k = B.shape[0]
    L = np.zeros_like(B)
    for j in range(k):
        d = B[j, j] - np.dot(L[j, :j], L[j, :j])
        if not d > 0:
            return j + 1
        L[j, j] = np.sqrt(d)
        for i in range(j + 1, k):
            L[i, j] = (B[i, j] - np.dot(L[i, :j], L[j, :j])) / L[j, j]
    return 0


def lobpcg(rp, ci, val, X0, m, tol=1e-8, maxiter=1000, miniter=0, precond=PRECOND_DEFAULT):
    """standard_eigen::LOBPCG<CRS<T>,T>::solve (src/eigen/lobpcg_standard.cpp:
    10-300) with initvec_scheme::USER, restated with numpy: vectors in T,
    Gram matrices in double, the 3m x 3m problem Sa v = lambda Sb v by
    scipy.linalg.eigh (LAPACK's sygvd, uplo 'L', as the reference), each
    eigenvector signed so that its largest component is positive.
    Preconditioners: default (z = r), none (copy) and Jacobi.  Returns
    (status, eigenvalues, X (m x n), iterations, max residual, history)."""
    import scipy.linalg as sla
    dt = val.dtype
    n = len(rp) - 1
    T = dt.type
    Minv = None
    if precond == PRECOND_JACOBI:
        d = np.zeros(n, dt)
        lib_diag = getattr(lib(), f"or_crs_diag_{_sfx(dt)}")
        lib_diag(n, n, _ptr(_c(rp, np.int32)), _ptr(_c(ci, np.int32)), _ptr(val), _ptr(d))
        Minv = (T(1) / d).astype(dt)
    mv = lambda v: spmv(rp, ci, val, v.astype(dt))
    nrm = lambda v: T(np.sqrt(T(np.dot(v, v))))
    wxp = np.zeros((3 * m, n), dt)
    WXP = np.zeros((3 * m, n), dt)
    wxp[m:2 * m] = np.asarray(X0, dt).reshape(m, n)
    for i in range(m):
        x = wxp[m + i]
        WXP[m + i] = mv(x)
        mu = T(np.dot(x, WXP[m + i]))
        wxp[i] = (WXP[m + i] - mu * x).astype(dt)
        wxp[i] = (wxp[i] * T(1.0 / float(nrm(wxp[i])))).astype(dt)
        WXP[i] = mv(wxp[i])
    singular = False
    hist = []
    resid = T(0)
    lam = np.zeros(m)
    it = 0
    while it < maxiter:
        k = 2 * m if (it == 0 or singular) else 3 * m
        a = wxp[:k].astype(np.float64)
        A_ = WXP[:k].astype(np.float64)
        Sa = A_ @ a.T
        Sb = a @ a.T
        Sa = np.tril(Sa) + np.tril(Sa, -1).T
        Sb = np.tril(Sb) + np.tril(Sb, -1).T
        q = _chol_fail_order(Sb)
        if q:
            if q > k - m and k == 3 * m:
                singular = True
                continue
            raise RuntimeError("LOBPCG: Gram matrix not positive definite")
        w, V = sla.eigh(Sa, Sb, lower=True)
        V = V.T.copy()
        for e in range(k):
            big = int(np.argmax(np.abs(V[e])))
            if V[e, big] < 0:
                V[e] = -V[e]
        wxp_p = wxp[:k].copy()
        WXP_p = WXP[:k].copy()
        residual = np.zeros(m, dt)
        for i in range(m):
            lam[i] = float(T(w[i]))
            b = np.zeros(3 * m, dt)
            b[:k] = V[i].astype(dt)
            p = np.zeros(n, dt)
            x = np.zeros(n, dt)
            P = np.zeros(n, dt)
            X = np.zeros(n, dt)
            for j in range(m):
                p = p + b[j] * wxp_p[j]
                x = x + b[j] * wxp_p[j]
                P = P + b[j] * WXP_p[j]
                X = X + b[j] * WXP_p[j]
                if k == 3 * m:
                    p = p + b[2 * m + j] * wxp_p[2 * m + j]
                    x = x + b[2 * m + j] * wxp_p[2 * m + j]
                    P = P + b[2 * m + j] * WXP_p[2 * m + j]
                    X = X + b[2 * m + j] * WXP_p[2 * m + j]
                x = x + b[m + j] * wxp_p[m + j]
                X = X + b[m + j] * WXP_p[m + j]
            sp = T(1.0 / float(nrm(p)))
            p, P = (p * sp).astype(dt), (P * sp).astype(dt)
            sx = T(1.0 / float(nrm(x)))
            x, X = (x * sx).astype(dt), (X * sx).astype(dt)
            wv = (X - T(lam[i]) * x).astype(dt)
            if Minv is not None:
                wv = (Minv * wv).astype(dt)
            residual[i] = nrm(wv)
            wxp[m + i], wxp[2 * m + i], WXP[m + i], WXP[2 * m + i] = x, p, X, P
            if np.isnan(residual[i]):
                return RESIDUAL_NAN, lam, wxp[m:2 * m].copy(), it + 1, float(residual[i]), hist
            wxp[i] = (wv * T(1.0 / float(residual[i]))).astype(dt)
            WXP[i] = mv(wxp[i])
        resid = residual.max()
        hist.append(float(resid))
        if float(resid) < tol and miniter < it + 1:
            return SUCCESS, lam, wxp[m:2 * m].copy(), it + 1, float(resid), hist
        singular = False
        it += 1
    return MAXITER, lam, wxp[m:2 * m].copy(), maxiter, float(resid), hist


def tridiagonal_toeplitz(n, a, b):
    """util::tridiagonal_toeplitz_matrix (src/utils/create_matrix/
    tridiagonal_toeplitz.cpp:6-23) as CRS arrays, and its closed-form
    eigenvalues (:30-33): a - 2|b| cos(pi (k+1) / (n+1))."""
    rows, cols, vals = [], [], []
    for i in range(n):
        for j in (i - 1, i, i + 1):
            if 0 <= j < n:
                rows.append(i)
                cols.append(j)
                vals.append(a if i == j else b)
    rp = np.zeros(n + 1, np.int32)
    np.add.at(rp, np.array(rows) + 1, 1)
    ev = a - 2.0 * abs(b) * np.cos(np.pi * (np.arange(n) + 1) / (n + 1))
    return np.cumsum(rp).astype(np.int32), np.array(cols, np.int32), np.array(vals), ev
