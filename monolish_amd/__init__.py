"""monolish_amd -- Python front end of the gfx950 CRS sparse BLAS and Krylov
library (libmonolish_amd.so, C-ABI declared in include/monolish_amd.h).

The objects mirror monolish's device-resident types: ``Vector`` is a
``vector<T>`` after ``send()``, ``CRS`` a ``matrix::CRS<T>`` after ``send()``;
``matvec``, ``dot``, ``nrm2``, ``axpy`` ... are ``monolish::blas`` / ``vml``
functions, and ``cg`` / ``bicgstab`` / ``jacobi`` are ``equation::*::solve``.
All arithmetic runs in HIP kernels; importing the package raises if the
library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmonolish_amd.so")

OK = 0
SOLVER_SIZE_ERROR, SOLVER_MAXITER, SOLVER_BREAKDOWN, SOLVER_RESIDUAL_NAN, SOLVER_NOT_IMPL = -1, -2, -3, -4, -10
ERR_INVALID, ERR_HIP, ERR_NOMEM, ERR_BUDGET, ERR_NOT_FOUND, ERR_TIMEOUT, ERR_CANCELLED = (
    -101, -102, -103, -104, -105, -106, -107)
REQ_QUEUED, REQ_ACTIVE, REQ_DONE = 0, 1, 2
F64, F32 = 0, 1
PRECOND_DEFAULT, PRECOND_NONE, PRECOND_JACOBI, PRECOND_SOR = 0, 1, 2, 3
PRECOND_IC, PRECOND_ILU = 4, 5
MODE_REFERENCE, MODE_FUSED, MODE_MIXED_IR = 0, 1, 2


class MgxError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is missing: build it with `python -m monolish_amd.build`")

_lib = C.CDLL(LIB_PATH)

_P, _I64, _I32, _D = C.c_void_p, C.c_int64, C.c_int32, C.c_double


class SolveParams(C.Structure):
    _fields_ = [("tol", _D), ("maxiter", _I64), ("miniter", _I64), ("precond", _I32),
                ("mode", _I32), ("tol_is_relative", _I32), ("check_every", _I32),
                ("rhistory", C.POINTER(_D)), ("rhistory_len", _I64), ("omega", _D)]


class SolveResultC(C.Structure):
    _fields_ = [("status", _I32), ("pad", _I32), ("final_iter", _I64), ("final_resid", _D),
                ("seconds", _D), ("inner_iter", _I64), ("outer_iter", _I64)]


class SpmvTuning(C.Structure):
    _fields_ = [("rows_per_block", _I32), ("nnz_per_block", _I32), ("long_row", _I32),
                ("xcd_remap", _I32), ("us", _D), ("candidates", _I32), ("column_panels", _I32),
                ("row_walk", _I32), ("pad", _I32)]


class CsrStats(C.Structure):
    _fields_ = [("rows", _I64), ("cols", _I64), ("nnz", _I64), ("max_row", _I32),
                ("min_row", _I32), ("mean_row", _D), ("stream_blocks", _I64), ("long_rows", _I64),
                ("gather_lines", _D)]


def _sig(name, args, res=C.c_int):
    f = getattr(_lib, name)
    f.argtypes = args
    f.restype = res
    return f


_sig("mgx_last_error", [], C.c_char_p)
_sig("mgx_version", [C.POINTER(C.c_int)] * 3)
_sig("mgx_device_count", [C.POINTER(C.c_int)])
_sig("mgx_init", [C.c_int, C.POINTER(_P)])
_sig("mgx_finalize", [_P])
_sig("mgx_sync", [_P])
_sig("mgx_stream", [_P], _P)
_sig("mgx_device_info", [_P, C.c_char_p, C.c_size_t, C.POINTER(C.c_int), C.POINTER(C.c_size_t)])
_sig("mgx_set_spmv_tuning", [_P, C.c_int, C.c_int, C.c_int, C.c_int])
_sig("mgx_vec_create", [_P, _I64, C.c_int, _P, C.POINTER(_P)])
_sig("mgx_vec_view", [_P, _I64, _I64, C.POINTER(_P)])
_sig("mgx_vec_upload", [_P, _P])
_sig("mgx_vec_download", [_P, _P])
_sig("mgx_vec_destroy", [_P])
_sig("mgx_vec_fill", [_P, _D])
_sig("mgx_vec_size", [_P], _I64)
_sig("mgx_vec_device_ptr", [_P], _P)
_sig("mgx_csr_create", [_P, _I64, _I64, _I64, _P, _P, _P, C.c_int, C.POINTER(_P)])
_sig("mgx_csr_from_coo", [_P, _I64, _I64, _I64, _P, _P, _P, C.c_int, C.c_int, C.POINTER(_P)])
_sig("mgx_csr_read_mm", [_P, C.c_char_p, C.c_int, C.c_int, C.POINTER(_P)])
_sig("mgx_csr_write_mm", [_P, C.c_char_p])
_sig("mgx_csr_update_values", [_P, _P])
_sig("mgx_csr_download", [_P, _P, _P, _P])
_sig("mgx_csr_destroy", [_P])
_sig("mgx_csr_diag", [_P, _P])
_sig("mgx_csr_stats_get", [_P, C.POINTER(CsrStats)])
_sig("mgx_csr_autotune", [_P, C.c_int, C.POINTER(SpmvTuning)])
_sig("mgx_csr_set_plan", [_P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int])
_sig("mgx_matvec", [_P, _P, _P])
_sig("mgx_matvec_ab", [_D, _P, _P, _D, _P])
_sig("mgx_matvec_t", [_P, _P, _P])
_sig("mgx_dot", [_P, _P, C.POINTER(_D)])
_sig("mgx_nrm2", [_P, C.POINTER(_D)])
_sig("mgx_dot_device", [_P, _P, _P])
_sig("mgx_nrm2_device", [_P, _P])
_sig("mgx_dot_device_noalloc", [_P, _P, _P])
_sig("mgx_set_stream", [_P, _P])
_sig("mgx_axpy", [_D, _P, _P])
_sig("mgx_xpay", [_D, _P, _P])
_sig("mgx_scal", [_D, _P])
_sig("mgx_copy", [_P, _P])
_sig("mgx_axpyz", [_D, _P, _P, _P])
for _n in ("mgx_vadd", "mgx_vsub", "mgx_vmul"):
    _sig(_n, [_P, _P, _P])
_sig("mgx_vreciprocal", [_P, _P])
_sig("mgx_solve_params_default", [C.POINTER(SolveParams)], None)
for _n in ("mgx_cg_solve", "mgx_bicgstab_solve", "mgx_jacobi_solve", "mgx_sor_solve",
           "mgx_ic_solve", "mgx_ilu_solve"):
    _sig(_n, [_P, _P, _P, C.POINTER(SolveParams), C.POINTER(SolveResultC)])
_sig("mgx_residual_l2", [_P, _P, _P, C.POINTER(_D)])
_sig("mgx_lobpcg_solve", [_P, _P, _I64, _P, C.POINTER(SolveParams), C.POINTER(SolveResultC)])
_sig("mgx_ir_cg_solve", [_P, _P, _P, C.POINTER(SolveParams), _D, _I64, C.POINTER(SolveResultC)])
_sig("mgx_mixed_cg_solve", [_P, _P, _P, C.POINTER(SolveParams), _D, _I64,
                            C.POINTER(SolveResultC)])
_sig("mgx_gen_random_vector", [_I64, _D, _D, C.c_uint32, C.c_int, _P])
_sig("mgx_gen_stencil3d", [_I64, _I64, _I64, C.c_int, _D, C.POINTER(_I64), _P, _P, _P])
_sig("mgx_gen_random_lognormal", [_I64, _I64, _D, _D, _I32, C.c_uint64, _P, _P, _P])
_sig("mgx_gen_laplacian2d", [_I64, _I64, C.POINTER(_I64), _P, _P, _P])
_sig("mgx_gen_convdiff2d", [_I64, _D, C.POINTER(_I64), _P, _P, _P])
_sig("mgx_gen_mmatrix", [_I64, _I32, _I32, _D, C.c_uint64, _P, _P, _P])
_sig("mgx_set_graphs", [_P, C.c_int])
_sig("mgx_set_persistent", [_P, C.c_int])
_sig("mgx_batch_solve", [_P, C.c_int, _I64, C.POINTER(_P), C.POINTER(_P), C.POINTER(_P),
                         C.POINTER(SolveParams), C.POINTER(SolveResultC)])
_sig("mgx_batch_ir_cg_solve", [_P, _I64, C.POINTER(_P), C.POINTER(_P), C.POINTER(_P),
                               C.POINTER(SolveParams), C.c_double, _I64, C.POINTER(SolveResultC)])


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_size_t), ("max_active", _I32), ("iters_per_poll", _I32),
                ("ir_inner_tol", _D), ("ir_max_outer", _I32), ("pad_", _I32)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", _I64), ("completed", _I64), ("active", _I64), ("queued", _I64),
                ("p50_latency_s", _D), ("p99_latency_s", _D), ("mean_latency_s", _D),
                ("max_latency_s", _D), ("mean_queue_s", _D), ("solves_per_s", _D),
                ("peak_bytes", C.c_size_t), ("cancelled", _I64), ("expired", _I64)]


_sig("mgx_server_create", [_P, C.POINTER(ServerConfig), C.POINTER(_P)])
_sig("mgx_server_submit", [_P, C.c_int, _I64, _I64, _P, _P, _P, _P, _P, C.POINTER(SolveParams),
                           C.POINTER(_I64)])
_sig("mgx_server_submit_resident", [_P, C.c_int, _P, _P, _P, C.POINTER(SolveParams),
                                    C.POINTER(_I64)])
_sig("mgx_server_cancel", [_P, _I64])
_sig("mgx_server_set_deadline", [_P, _I64, _D])
_sig("mgx_server_poll", [_P, _I64, C.POINTER(_I32)])
_sig("mgx_server_wait", [_P, _I64, _D, C.POINTER(SolveResultC), _P])
_sig("mgx_server_stats_get", [_P, C.POINTER(ServerStats)])
_sig("mgx_server_destroy", [_P])
METHOD_CG, METHOD_BICGSTAB = 0, 1


def lib() -> C.CDLL:
    return _lib


def last_error() -> str:
    return (_lib.mgx_last_error() or b"").decode()


def check(rc: int) -> int:
    if rc <= -100:
        raise MgxError(rc, last_error())
    return rc


def _dt(dtype) -> int:
    dt = np.dtype(dtype)
    if dt == np.float64:
        return F64
    if dt == np.float32:
        return F32
    raise TypeError(f"unsupported dtype {dt}")


def _np_dtype(code: int):
    return np.float64 if code == F64 else np.float32


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def device_count() -> int:
    n = C.c_int(0)
    check(_lib.mgx_device_count(C.byref(n)))
    return n.value


class Context:
    """One device and one HIP stream (util::set_default_device analogue)."""

    def __init__(self, device: int = 0):
        self._h = _P()
        check(_lib.mgx_init(device, C.byref(self._h)))
        self.device = device

    @property
    def handle(self):
        return self._h

    def sync(self):
        check(_lib.mgx_sync(self._h))

    @property
    def stream(self) -> int:
        return _lib.mgx_stream(self._h) or 0

    def info(self):
        name = C.create_string_buffer(256)
        cus = C.c_int(0)
        hbm = C.c_size_t(0)
        check(_lib.mgx_device_info(self._h, name, 256, C.byref(cus), C.byref(hbm)))
        return {"name": name.value.decode(), "compute_units": cus.value, "hbm_bytes": hbm.value}

    def set_spmv_tuning(self, rows_per_block=256, nnz_per_block=2048, long_row=256, xcd_remap=True):
        check(_lib.mgx_set_spmv_tuning(self._h, rows_per_block, nnz_per_block, long_row, int(xcd_remap)))

    def close(self):
        if self._h:
            check(_lib.mgx_finalize(self._h))
            self._h = _P()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Vector:
    """A device-resident vector<T> (after send())."""

    def __init__(self, ctx: Context, data=None, n: int | None = None, dtype=np.float64, _handle=None,
                 _base=None):
        self.ctx = ctx
        self._base = _base
        if _handle is not None:
            self._h = _handle
        else:
            self._h = _P()
            if data is not None:
                arr = np.ascontiguousarray(data, dtype=dtype)
                check(_lib.mgx_vec_create(ctx.handle, arr.size, _dt(dtype), _ptr(arr), C.byref(self._h)))
            else:
                check(_lib.mgx_vec_create(ctx.handle, int(n or 0), _dt(dtype), None, C.byref(self._h)))
        self.dtype = np.dtype(dtype)

    @property
    def handle(self):
        return self._h

    @property
    def size(self) -> int:
        return int(_lib.mgx_vec_size(self._h))

    def __len__(self):
        return self.size

    @property
    def device_ptr(self) -> int:
        return _lib.mgx_vec_device_ptr(self._h) or 0

    def download(self) -> np.ndarray:
        out = np.empty(self.size, dtype=self.dtype)
        check(_lib.mgx_vec_download(self._h, _ptr(out)))
        return out

    def upload(self, data):
        arr = np.ascontiguousarray(data, dtype=self.dtype)
        if arr.size != self.size:
            raise ValueError("size mismatch")
        check(_lib.mgx_vec_upload(self._h, _ptr(arr)))

    def fill(self, value: float):
        check(_lib.mgx_vec_fill(self._h, value))

    def view(self, offset: int, n: int) -> "Vector":
        h = _P()
        check(_lib.mgx_vec_view(self._h, offset, n, C.byref(h)))
        return Vector(self.ctx, dtype=self.dtype, _handle=h, _base=self)

    def free(self):
        if getattr(self, "_h", None):
            # Context.close() already released it (mgx_finalize)
            if self.ctx.handle:
                check(_lib.mgx_vec_destroy(self._h))
            self._h = _P()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class CRS:
    """A device-resident matrix::CRS<T> (after send())."""

    def __init__(self, ctx: Context, rows: int, cols: int, row_ptr, col_ind, val, dtype=None):
        self.ctx = ctx
        val = np.asarray(val)
        dtype = np.dtype(dtype or val.dtype)
        rp = np.ascontiguousarray(row_ptr, dtype=np.int32)
        ci = np.ascontiguousarray(col_ind, dtype=np.int32)
        v = np.ascontiguousarray(val, dtype=dtype)
        if rp.size != rows + 1 or ci.size != v.size:
            raise ValueError("inconsistent CRS arrays")
        self._h = _P()
        check(_lib.mgx_csr_create(ctx.handle, rows, cols, v.size, _ptr(rp), _ptr(ci), _ptr(v),
                                  _dt(dtype), C.byref(self._h)))
        self.rows, self.cols, self.nnz, self.dtype = rows, cols, int(v.size), dtype

    @classmethod
    def _adopt(cls, ctx: Context, h, dtype):
        self = cls.__new__(cls)
        self.ctx, self._h, self.dtype = ctx, h, np.dtype(dtype)
        s = CsrStats()
        check(_lib.mgx_csr_stats_get(h, C.byref(s)))
        self.rows, self.cols, self.nnz = int(s.rows), int(s.cols), int(s.nnz)
        return self

    @classmethod
    def from_coo(cls, ctx: Context, rows: int, cols: int, row, col, val, sort: int = 0,
                 dtype=None):
        """CRS(COO&) built on the device from COO triplets.  sort: 0 the
        given (row-ordered) order, 1 stable (row, col) sort, 2 sort and merge
        duplicates (the last in input order wins)."""
        val = np.asarray(val)
        dtype = np.dtype(dtype or val.dtype)
        r = np.ascontiguousarray(row, dtype=np.int32)
        c = np.ascontiguousarray(col, dtype=np.int32)
        v = np.ascontiguousarray(val, dtype=dtype)
        if not (r.size == c.size == v.size):
            raise ValueError("inconsistent COO arrays")
        h = _P()
        check(_lib.mgx_csr_from_coo(ctx.handle, rows, cols, v.size, _ptr(r), _ptr(c), _ptr(v),
                                    _dt(dtype), int(sort), C.byref(h)))
        return cls._adopt(ctx, h, dtype)

    @classmethod
    def read_mm(cls, ctx: Context, path: str, dtype=np.float64, sort: int = 0):
        """COO::input_mm(path) then CRS(COO&) (optionally sorted / merged),
        the conversion on the device."""
        h = _P()
        check(_lib.mgx_csr_read_mm(ctx.handle, os.fsencode(path), _dt(dtype), int(sort),
                                   C.byref(h)))
        return cls._adopt(ctx, h, dtype)

    def write_mm(self, path: str):
        """COO(A).output_mm(path): MatrixMarket text, max_digits10 scientific."""
        check(_lib.mgx_csr_write_mm(self._h, os.fsencode(path)))

    @property
    def handle(self):
        return self._h

    def download(self):
        rp = np.empty(self.rows + 1, np.int32)
        ci = np.empty(self.nnz, np.int32)
        v = np.empty(self.nnz, self.dtype)
        check(_lib.mgx_csr_download(self._h, _ptr(rp), _ptr(ci), _ptr(v)))
        return rp, ci, v

    def update_values(self, val):
        v = np.ascontiguousarray(val, dtype=self.dtype)
        if v.size != self.nnz:
            raise ValueError("size mismatch")
        check(_lib.mgx_csr_update_values(self._h, _ptr(v)))

    def diag(self, d: Vector):
        check(_lib.mgx_csr_diag(self._h, d.handle))

    def stats(self) -> dict:
        s = CsrStats()
        check(_lib.mgx_csr_stats_get(self._h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in CsrStats._fields_}

    def set_plan(self, rows_per_block=256, nnz_per_block=2048, xcd_remap=True, row_walk=0,
                 column_panels=0):
        """Fix the SpMV block plan (the autotuner's candidates).  row_walk: 0
        products staged in LDS in entry order, 1 row walk, 2 sliced layout, 3
        L2 panels.  xcd_remap: 0 launch order, 1 one range of blocks per XCD,
        c in [2, 1024] runs of c blocks dealt to the XCDs in turn."""
        check(_lib.mgx_csr_set_plan(self._h, rows_per_block, nnz_per_block, int(xcd_remap),
                                    int(row_walk), column_panels))

    def autotune(self, reps: int = 20) -> dict:
        """Time the SpMV under candidate block plans and keep the fastest."""
        t = SpmvTuning()
        check(_lib.mgx_csr_autotune(self._h, int(reps), C.byref(t)))
        return {f: getattr(t, f) for f, _ in SpmvTuning._fields_ if f != "pad"}

    def free(self):
        if getattr(self, "_h", None):
            if self.ctx.handle:
                check(_lib.mgx_csr_destroy(self._h))
            self._h = _P()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ------------------------------------------------------------- BLAS / VML --
def matvec(A: CRS, x: Vector, y: Vector):
    check(_lib.mgx_matvec(A.handle, x.handle, y.handle))


def matvec_ab(alpha: float, A: CRS, x: Vector, beta: float, y: Vector):
    check(_lib.mgx_matvec_ab(alpha, A.handle, x.handle, beta, y.handle))


def matvec_t(A: CRS, x: Vector, y: Vector):
    check(_lib.mgx_matvec_t(A.handle, x.handle, y.handle))


def dot(x: Vector, y: Vector) -> float:
    out = _D(0)
    check(_lib.mgx_dot(x.handle, y.handle, C.byref(out)))
    return out.value


def dot_device(x: Vector, y: Vector, out: Vector, index: int = 0):
    """<x, y> into out[index] (an FP64 vector), queued without a host round trip."""
    if out.dtype != np.float64 or not 0 <= index < out.size:
        raise ValueError("out must be an FP64 vector holding index")
    check(_lib.mgx_dot_device(x.handle, y.handle, out.device_ptr + 8 * index))


def dot_device_noalloc(x: Vector, y: Vector, out: Vector, index: int = 0):
    """dot_device without device allocations: the reduction's scratch is a
    static per-device array (at most 2**26 elements; calls overlapping on one
    device race)."""
    if out.dtype != np.float64 or not 0 <= index < out.size:
        raise ValueError("out must be an FP64 vector holding index")
    check(_lib.mgx_dot_device_noalloc(x.handle, y.handle, out.device_ptr + 8 * index))


def nrm2_device(x: Vector, out: Vector, index: int = 0):
    """||x|| into out[index] (an FP64 vector), queued without a host round trip."""
    if out.dtype != np.float64 or not 0 <= index < out.size:
        raise ValueError("out must be an FP64 vector holding index")
    check(_lib.mgx_nrm2_device(x.handle, out.device_ptr + 8 * index))


def set_stream(ctx: Context, stream: int | None):
    """Queue the context's later calls on a caller's hipStream_t (None: the
    library's own stream)."""
    check(_lib.mgx_set_stream(ctx.handle, C.c_void_p(stream) if stream else None))


def nrm2(x: Vector) -> float:
    out = _D(0)
    check(_lib.mgx_nrm2(x.handle, C.byref(out)))
    return out.value


def axpy(a, x, y):
    check(_lib.mgx_axpy(a, x.handle, y.handle))


def xpay(a, x, y):
    check(_lib.mgx_xpay(a, x.handle, y.handle))


def scal(a, x):
    check(_lib.mgx_scal(a, x.handle))


def copy(x, y):
    check(_lib.mgx_copy(x.handle, y.handle))


def axpyz(a, x, y, z):
    check(_lib.mgx_axpyz(a, x.handle, y.handle, z.handle))


def vadd(a, b, y):
    check(_lib.mgx_vadd(a.handle, b.handle, y.handle))


def vsub(a, b, y):
    check(_lib.mgx_vsub(a.handle, b.handle, y.handle))


def vmul(a, b, y):
    check(_lib.mgx_vmul(a.handle, b.handle, y.handle))


def vreciprocal(a, y):
    check(_lib.mgx_vreciprocal(a.handle, y.handle))


def residual_l2(A: CRS, x: Vector, b: Vector) -> float:
    out = _D(0)
    check(_lib.mgx_residual_l2(A.handle, x.handle, b.handle, C.byref(out)))
    return out.value


# ---------------------------------------------------------------- solvers --
@dataclass
class SolveResult:
    status: int
    final_iter: int
    final_resid: float
    seconds: float
    inner_iter: int = 0
    outer_iter: int = 0
    rhistory: np.ndarray | None = None


def _params(tol, maxiter, miniter, precond, mode, tol_is_relative, check_every, history,
            omega=1.9):
    p = SolveParams()
    _lib.mgx_solve_params_default(C.byref(p))
    p.omega = float(omega)
    p.tol = tol
    p.maxiter = int(maxiter)
    p.miniter = int(miniter)
    p.precond = precond
    p.mode = mode
    p.tol_is_relative = int(tol_is_relative)
    p.check_every = int(check_every)
    hist = None
    if history:
        hist = np.full(int(history), np.nan)
        p.rhistory = hist.ctypes.data_as(C.POINTER(_D))
        p.rhistory_len = hist.size
    return p, hist


def _solve(fn, A, x, b, tol, maxiter, miniter, precond, mode, tol_is_relative, check_every, history,
           omega=1.9):
    p, hist = _params(tol, maxiter, miniter, precond, mode, tol_is_relative, check_every, history,
                      omega)
    r = SolveResultC()
    rc = check(fn(A.handle, x.handle, b.handle, C.byref(p), C.byref(r)))
    if hist is not None:
        hist = hist[: min(r.final_iter, hist.size)]
    return SolveResult(rc, r.final_iter, r.final_resid, r.seconds, r.inner_iter, r.outer_iter, hist)


def cg(A, x, b, tol=1e-8, maxiter=0, miniter=0, precond=PRECOND_JACOBI, mode=MODE_FUSED,
       tol_is_relative=False, check_every=0, history=0, omega=1.9) -> SolveResult:
    """equation::CG<CRS<T>,T>::solve (src/equation/cg.cpp); omega: SOR precond."""
    return _solve(_lib.mgx_cg_solve, A, x, b, tol, maxiter, miniter, precond, mode,
                  tol_is_relative, check_every, history, omega)


def bicgstab(A, x, b, tol=1e-8, maxiter=0, miniter=0, precond=PRECOND_JACOBI, mode=MODE_REFERENCE,
             tol_is_relative=False, check_every=0, history=0, omega=1.9) -> SolveResult:
    """equation::BiCGSTAB<CRS<T>,T>::solve (src/equation/bicgstab.cpp)."""
    return _solve(_lib.mgx_bicgstab_solve, A, x, b, tol, maxiter, miniter, precond, mode,
                  tol_is_relative, check_every, history, omega)


def jacobi(A, x, b, tol=1e-8, maxiter=0, miniter=0, precond=PRECOND_DEFAULT, mode=MODE_REFERENCE,
           check_every=0, history=0, omega=1.9) -> SolveResult:
    """equation::Jacobi<CRS<T>,T>::solve (src/equation/jacobi.cpp); tol is relative."""
    return _solve(_lib.mgx_jacobi_solve, A, x, b, tol, maxiter, miniter, precond, mode, False,
                  check_every, history, omega)


def sor(A, x, b, tol=1e-8, maxiter=0, miniter=0, precond=PRECOND_DEFAULT, omega=1.9,
        history=0) -> SolveResult:
    """equation::SOR<CRS<T>,T>::solve (src/equation/sor.cpp); tol is relative."""
    return _solve(_lib.mgx_sor_solve, A, x, b, tol, maxiter, miniter, precond, MODE_REFERENCE,
                  False, 0, history, omega)


def ic_solve(A, x, b) -> SolveResult:
    """equation::IC<CRS<T>,T>::solve (src/equation/ic.cpp): x = (L L^T)^-1 b
    with the IC(0) factor of A."""
    return _solve(_lib.mgx_ic_solve, A, x, b, 1e-8, 0, 0, PRECOND_DEFAULT, MODE_REFERENCE,
                  False, 0, 0)


def ilu_solve(A, x, b) -> SolveResult:
    """equation::ILU<CRS<T>,T>::solve (src/equation/ilu.cpp): x = (L U)^-1 b
    with the ILU(0) factor of A."""
    return _solve(_lib.mgx_ilu_solve, A, x, b, 1e-8, 0, 0, PRECOND_DEFAULT, MODE_REFERENCE,
                  False, 0, 0)


def lobpcg(A, X, m, tol=1e-8, maxiter=0, miniter=0, precond=PRECOND_DEFAULT, omega=1.9,
           history=0):
    """standard_eigen::LOBPCG<CRS<T>,T>::solve (src/eigen/lobpcg_standard.cpp)
    with initvec_scheme::USER: X (a Vector of m * n entries, row i = vector
    i) holds the initial vectors and receives the eigenvectors.  Returns
    (SolveResult, eigenvalues as float64)."""
    p, hist = _params(tol, maxiter, miniter, precond, MODE_REFERENCE, False, 0, history, omega)
    lam = np.zeros(int(m))
    r = SolveResultC()
    rc = check(_lib.mgx_lobpcg_solve(A.handle, X.handle, int(m), lam.ctypes.data_as(C.c_void_p),
                                     C.byref(p), C.byref(r)))
    if hist is not None:
        hist = hist[: min(r.final_iter, hist.size)]
    return SolveResult(rc, r.final_iter, r.final_resid, r.seconds, 0, 0, hist), lam


def ir_cg(A, x, b, tol=1e-10, tol_is_relative=True, inner_tol=1e-4, max_outer=100, maxiter=0,
          precond=PRECOND_JACOBI, check_every=0) -> SolveResult:
    """Mixed-precision iterative refinement: FP32 inner Jacobi-PCG, FP64 outer."""
    p, _ = _params(tol, maxiter, 0, precond, MODE_FUSED, tol_is_relative, check_every, 0)
    r = SolveResultC()
    rc = check(_lib.mgx_ir_cg_solve(A.handle, x.handle, b.handle, C.byref(p), inner_tol,
                                    int(max_outer), C.byref(r)))
    return SolveResult(rc, r.final_iter, r.final_resid, r.seconds, r.inner_iter, r.outer_iter)


def mixed_cg(A, x, b, tol=1e-10, tol_is_relative=True, delta=0.1, max_updates=1000, maxiter=0,
             precond=PRECOND_JACOBI, check_every=0) -> SolveResult:
    """Mixed-precision Jacobi-PCG: FP32 iterations, FP64 residual replacement
    each time the FP32 residual has dropped by ``delta`` (search direction
    kept).  inner_iter = FP32 iterations, outer_iter = replacements."""
    p, _ = _params(tol, maxiter, 0, precond, MODE_FUSED, tol_is_relative, check_every, 0)
    r = SolveResultC()
    rc = check(_lib.mgx_mixed_cg_solve(A.handle, x.handle, b.handle, C.byref(p), delta,
                                       int(max_updates), C.byref(r)))
    return SolveResult(rc, r.final_iter, r.final_resid, r.seconds, r.inner_iter, r.outer_iter)


# ------------------------------------------------------------- generators --
def stencil3d(nx: int, ny: int, nz: int, points: int = 7, shift: float = 0.0):
    """3D 7- or 27-point stencil CRS arrays (row_ptr, col_ind, val)."""
    nnz = _I64(0)
    check(_lib.mgx_gen_stencil3d(nx, ny, nz, points, shift, C.byref(nnz), None, None, None))
    n = nx * ny * nz
    rp = np.empty(n + 1, np.int32)
    ci = np.empty(nnz.value, np.int32)
    v = np.empty(nnz.value, np.float64)
    check(_lib.mgx_gen_stencil3d(nx, ny, nz, points, shift, C.byref(nnz), _ptr(rp), _ptr(ci), _ptr(v)))
    return rp, ci, v


def random_lognormal(rows: int, cols: int, mean: float = 32.0, sigma: float = 1.0,
                     maxlen: int = 2048, seed: int = 2002):
    """Random CRS with lognormal row lengths (mean `mean`), distinct sorted
    uniform columns and U[-1,1) values."""
    rp = np.empty(rows + 1, np.int32)
    check(_lib.mgx_gen_random_lognormal(rows, cols, mean, sigma, maxlen, seed, _ptr(rp), None, None))
    nnz = int(rp[-1])
    ci = np.empty(nnz, np.int32)
    v = np.empty(nnz, np.float64)
    check(_lib.mgx_gen_random_lognormal(rows, cols, mean, sigma, maxlen, seed, _ptr(rp), _ptr(ci), _ptr(v)))
    return rp, ci, v


def laplacian2d(m: int, n: int):
    """util::laplacian_matrix_2D_5p(m, n) as CRS arrays (row_ptr, col_ind, val)."""
    return _gen2(_lib.mgx_gen_laplacian2d, m, n)


def convdiff2d(m: int, peclet: float):
    """2D upwind convection-diffusion on an m x m grid (nonsymmetric)."""
    return _gen2(_lib.mgx_gen_convdiff2d, m, float(peclet), rows=m * m)


def _gen2(fn, *args, rows=None):
    nnz = _I64(0)
    check(fn(*args, C.byref(nnz), None, None, None))
    n = rows if rows is not None else args[0] * args[1]
    rp = np.empty(n + 1, np.int32)
    ci = np.empty(nnz.value, np.int32)
    v = np.empty(nnz.value, np.float64)
    check(fn(*args, C.byref(nnz), _ptr(rp), _ptr(ci), _ptr(v)))
    return rp, ci, v


def mmatrix(n: int, kmin: int = 5, kmax: int = 15, shift: float = 0.5, seed: int = 4000):
    """SPD M-matrix with kmin..kmax random off-diagonals per row, symmetrised."""
    rp = np.empty(n + 1, np.int32)
    check(_lib.mgx_gen_mmatrix(n, kmin, kmax, shift, seed, _ptr(rp), None, None))
    nnz = int(rp[-1])
    ci = np.empty(nnz, np.int32)
    v = np.empty(nnz, np.float64)
    check(_lib.mgx_gen_mmatrix(n, kmin, kmax, shift, seed, _ptr(rp), _ptr(ci), _ptr(v)))
    return rp, ci, v


def random_vector(n: int, vmin: float, vmax: float, seed: int, dtype=np.float64) -> np.ndarray:
    """vector<T>(n, vmin, vmax, seed): std::mt19937 + uniform_real_distribution
    draws in order (FP32: rounded draws)."""
    out = np.empty(n, dtype)
    check(_lib.mgx_gen_random_vector(n, vmin, vmax, seed, _dt(dtype), _ptr(out)))
    return out


def set_graphs(ctx: Context, enable: bool):
    check(_lib.mgx_set_graphs(ctx.handle, int(enable)))


def set_persistent(ctx: Context, enable, max_blocks: int = 0):
    """Run single small fused CG solves in one cooperative launch (default off);
    max_blocks > 1 caps its grid."""
    check(_lib.mgx_set_persistent(ctx.handle, int(max_blocks) if enable and max_blocks > 1
                                  else int(bool(enable))))


def batch_solve(ctx: Context, method: int, As, xs, bs, tol=1e-8, maxiter=0, miniter=0,
                precond=PRECOND_JACOBI, tol_is_relative=False, check_every=0):
    """Solve many systems in shared fused launches; returns a SolveResult each."""
    ns = len(As)
    p, _ = _params(tol, maxiter, miniter, precond, MODE_FUSED, tol_is_relative, check_every, 0)
    res = (SolveResultC * max(ns, 1))()
    arrA = (_P * max(ns, 1))(*[a.handle for a in As])
    arrX = (_P * max(ns, 1))(*[x.handle for x in xs])
    arrB = (_P * max(ns, 1))(*[b.handle for b in bs])
    check(_lib.mgx_batch_solve(ctx.handle, method, ns, arrA, arrX, arrB, C.byref(p), res))
    return [SolveResult(r.status, r.final_iter, r.final_resid, r.seconds) for r in res[:ns]]


def batch_ir_cg(ctx: Context, As, xs, bs, tol=1e-10, tol_is_relative=True, inner_tol=1e-5,
                max_outer=100, maxiter=0, precond=PRECOND_JACOBI, check_every=0):
    """FP32-refined batched CG over FP64 systems; returns a SolveResult each."""
    ns = len(As)
    p, _ = _params(tol, maxiter, 0, precond, MODE_FUSED, tol_is_relative, check_every, 0)
    res = (SolveResultC * max(ns, 1))()
    arrA = (_P * max(ns, 1))(*[a.handle for a in As])
    arrX = (_P * max(ns, 1))(*[x.handle for x in xs])
    arrB = (_P * max(ns, 1))(*[b.handle for b in bs])
    check(_lib.mgx_batch_ir_cg_solve(ctx.handle, ns, arrA, arrX, arrB, C.byref(p), inner_tol,
                                     int(max_outer), res))
    return [SolveResult(r.status, r.final_iter, r.final_resid, r.seconds, r.inner_iter,
                        r.outer_iter) for r in res[:ns]]


class Server:
    """Continuous-batching solve server on one context (see mgx_server_*)."""

    def __init__(self, ctx: Context, hbm_budget: int = 0, max_active: int = 0, iters_per_poll: int = 0,
                 ir_inner_tol: float = 0.0, ir_max_outer: int = 0):
        cfg = ServerConfig(hbm_budget, max_active, iters_per_poll, ir_inner_tol, ir_max_outer, 0)
        self._h = _P()
        check(_lib.mgx_server_create(ctx.handle, C.byref(cfg), C.byref(self._h)))

    def submit(self, method, rp, ci, val, b, x0=None, tol=1e-10, tol_is_relative=True, maxiter=0,
               precond=PRECOND_JACOBI, mode=MODE_FUSED) -> int:
        """Queue a solve of a copied FP64 system; mode MODE_MIXED_IR serves a CG
        request by FP32-refined iterations."""
        rp = np.ascontiguousarray(rp, np.int32)
        ci = np.ascontiguousarray(ci, np.int32)
        val = np.ascontiguousarray(val, np.float64)
        b = np.ascontiguousarray(b, np.float64)
        x0p = None
        if x0 is not None:
            x0 = np.ascontiguousarray(x0, np.float64)
            x0p = _ptr(x0)
        p, _ = _params(tol, maxiter, 0, precond, mode, tol_is_relative, 0, 0)
        t = _I64(0)
        check(_lib.mgx_server_submit(self._h, method, len(b), len(val), _ptr(rp), _ptr(ci), _ptr(val),
                                     _ptr(b), x0p, C.byref(p), C.byref(t)))
        return t.value

    def submit_resident(self, method, A, b, x=None, tol=1e-10, tol_is_relative=True, maxiter=0,
                        precond=PRECOND_JACOBI, mode=MODE_FUSED) -> int:
        """Queue a solve of a device-resident system (A, b stay the caller's and may be
        shared by many requests); x, if given, is the start and receives the solution."""
        p, _ = _params(tol, maxiter, 0, precond, mode, tol_is_relative, 0, 0)
        t = _I64(0)
        check(_lib.mgx_server_submit_resident(self._h, method, A.handle, b.handle,
                                              x.handle if x is not None else None, C.byref(p),
                                              C.byref(t)))
        return t.value

    def wait(self, ticket: int, n: int | None = None, timeout: float = -1.0):
        r = SolveResultC()
        x = np.empty(n, np.float64) if n else None
        rc = _lib.mgx_server_wait(self._h, ticket, timeout, C.byref(r), _ptr(x) if x is not None else None)
        if rc == ERR_NOT_FOUND:
            return None
        check(rc)
        return SolveResult(r.status, r.final_iter, r.final_resid, r.seconds, r.inner_iter,
                           r.outer_iter), x

    def cancel(self, ticket: int):
        """End a queued or running request (status ERR_CANCELLED); a request
        that already finished keeps its result."""
        check(_lib.mgx_server_cancel(self._h, ticket))

    def set_deadline(self, ticket: int, seconds: float):
        """End the request with ERR_TIMEOUT unless it finishes within
        `seconds` of its submission (0: no deadline)."""
        check(_lib.mgx_server_set_deadline(self._h, ticket, float(seconds)))

    def poll(self, ticket: int) -> int:
        """REQ_QUEUED, REQ_ACTIVE or REQ_DONE."""
        st = _I32(0)
        check(_lib.mgx_server_poll(self._h, ticket, C.byref(st)))
        return int(st.value)

    def stats(self) -> dict:
        s = ServerStats()
        check(_lib.mgx_server_stats_get(self._h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in ServerStats._fields_}

    def close(self):
        if getattr(self, "_h", None):
            check(_lib.mgx_server_destroy(self._h))
            self._h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
