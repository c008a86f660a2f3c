"""bola_amd -- the reference crate's dense-solver path on MI355X (gfx950).

The functions here mirror the crate's public functions one for one (same names, same
argument order and meaning, column-major f64, 1-based int32 pivots, in-place results,
and the same error texts), on top of the C ABI of libbola_hip.so (include/bola_hip.h).

Operands may be numpy arrays (host memory: the call copies them to HBM and back, like a
drop-in replacement of the CPU function) or torch CUDA tensors (device memory: the call
runs on torch's current stream and copies nothing).  A matrix operand is a flat array
that starts at the operand's first element, exactly like the crate's raw pointers.

Functions that return ``Result<(), String>`` in the crate raise :class:`BolaError` with
the crate's message instead of returning ``Err``.

Crate sources mirrored (file:line): gemm.rs:88, getrf.rs:37, getrf2.rs:39, potrf.rs:101,
potrf2.rs:25, trsm.rs:33, syrk.rs:17, laswp.rs:22, idamax.rs:23, ger.rs:28, scal.rs:45,
utilities.rs:22 (dot_product), utilities.rs:107 (daxpy).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

__all__ = [
    "BolaError", "Handle", "build", "lib", "library_path",
    "gemm", "getrf", "getrf2", "potrf", "potrf2", "trsm", "syrk", "laswp", "idamax",
    "ger", "scal", "daxpy", "dot_product", "getrs", "potrs",
    "MixedStats", "PRECISIONS", "dgesv_mixed", "sgemm_lp",
    "BATCH_LU", "BATCH_CHOL", "BATCH_FP64", "BATCH_FP32_IR", "BATCH_DESC_DTYPE", "BatchConfig",
    "batch_solve", "batch_tune", "BatchGraph",
    "SERVE_LU", "SERVE_CHOL", "Server",
]

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "libbola_hip.so")
_lib = None

ERR_DEVICE = -1000


class BolaError(Exception):
    """The crate's ``Err(String)``; ``code`` is the C ABI status (negative = illegal
    argument, positive = 1-based column of a zero pivot / failing leading minor)."""

    def __init__(self, message: str, code: int = 0):
        super().__init__(message)
        self.message = message
        self.code = code


def library_path() -> str:
    return _SO


def build() -> None:
    """Compile libbola_hip.so for gfx950 (bola_amd/csrc/Makefile)."""
    subprocess.run(["make", "-s", "-j8", "-C", os.path.join(_HERE, "csrc")], check=True)


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_lp = C.POINTER(C.c_int64)
_sz = C.c_size_t
_vp = C.c_void_p
_ch = C.c_char
_d = C.c_double
_i32 = C.c_int32

# (name, restype, argtypes)
_SIGS = [
    ("bola_hip_create", C.c_int, [C.POINTER(_vp)]),
    ("bola_hip_destroy", C.c_int, [_vp]),
    ("bola_hip_set_stream", C.c_int, [_vp, _vp]),
    ("bola_hip_set_exclusive", C.c_int, [_vp, C.c_int]),
    ("bola_hip_get_stream", _vp, [_vp]),
    ("bola_hip_synchronize", C.c_int, [_vp]),
    ("bola_hip_error_string", C.c_char_p, [_vp]),
    ("bola_hip_last_error", C.c_char_p, []),
    ("bola_hip_profile_enable", C.c_int, [_vp, C.c_int]),
    ("bola_hip_profile_reset", C.c_int, [_vp]),
    ("bola_hip_profile_get", C.c_int, [_vp, C.c_char_p, _dp, C.POINTER(C.c_long), _dp]),
    ("bola_hip_profile_report", C.c_int, [_vp, _d, C.c_char_p, _sz]),
    ("bola_hip_idamax", _sz, [_sz, _vp]),
    ("bola_hip_idamax_dev", C.c_int, [_vp, _sz, _vp, _vp]),
    ("bola_hip_dlaswp", C.c_int, [_sz, _vp, _sz, _i32, _i32, _vp, _i32]),
    ("bola_hip_dlaswp_dev", C.c_int, [_vp, _sz, _vp, _sz, _i32, _i32, _vp, _i32]),
    ("bola_hip_dger", C.c_int, [_sz, _sz, _d, _vp, _vp, _sz, _vp, _sz]),
    ("bola_hip_dger_dev", C.c_int, [_vp, _sz, _sz, _d, _vp, _vp, _sz, _vp, _sz]),
    ("bola_hip_dscal", C.c_int, [_sz, _d, _vp, _sz]),
    ("bola_hip_dscal_dev", C.c_int, [_vp, _sz, _d, _vp, _sz]),
    ("bola_hip_daxpy", C.c_int, [_sz, _d, _vp, _sz, _vp, _sz]),
    ("bola_hip_daxpy_dev", C.c_int, [_vp, _sz, _d, _vp, _sz, _vp, _sz]),
    ("bola_hip_ddot", _d, [_sz, _vp, _sz, _vp, _sz]),
    ("bola_hip_ddot_dev", C.c_int, [_vp, _sz, _vp, _sz, _vp, _sz, _vp]),
    ("bola_hip_dgemm", C.c_int, [_ch, _ch, _sz, _sz, _sz, _d, _vp, _sz, _vp, _sz, _d, _vp, _sz]),
    ("bola_hip_dgemm_dev", C.c_int, [_vp, _ch, _ch, _sz, _sz, _sz, _d, _vp, _sz, _vp, _sz, _d, _vp, _sz]),
    ("bola_hip_dtrsm", C.c_int, [_ch, _ch, _ch, _ch, _sz, _sz, _d, _vp, _sz, _vp, _sz]),
    ("bola_hip_dtrsm_dev", C.c_int, [_vp, _ch, _ch, _ch, _ch, _sz, _sz, _d, _vp, _sz, _vp, _sz]),
    ("bola_hip_dsyrk", C.c_int, [_ch, _ch, _sz, _sz, _d, _vp, _sz, _d, _vp, _sz]),
    ("bola_hip_dsyrk_dev", C.c_int, [_vp, _ch, _ch, _sz, _sz, _d, _vp, _sz, _d, _vp, _sz]),
    ("bola_hip_dgetrf", C.c_int, [_sz, _sz, _vp, _sz, _vp]),
    ("bola_hip_dgetrf_dev", C.c_int, [_vp, _sz, _sz, _vp, _sz, _vp, _vp]),
    ("bola_hip_dgetrf_noalloc_dev", C.c_int, [_vp, _sz, _sz, _vp, _sz, _vp, _vp]),
    ("bola_hip_dgetrf2", C.c_int, [_sz, _sz, _vp, _sz, _vp]),
    ("bola_hip_dgetrf2_dev", C.c_int, [_vp, _sz, _sz, _vp, _sz, _vp, _vp]),
    ("bola_hip_dpotrf", C.c_int, [_ch, _sz, _vp, _sz]),
    ("bola_hip_dpotrf_dev", C.c_int, [_vp, _ch, _sz, _vp, _sz, _vp]),
    ("bola_hip_dpotrf2", C.c_int, [_ch, _sz, _vp, _sz]),
    ("bola_hip_dpotrf2_dev", C.c_int, [_vp, _ch, _sz, _vp, _sz, _vp]),
    ("bola_hip_dgetrs", C.c_int, [_sz, _sz, _vp, _sz, _vp, _vp, _sz]),
    ("bola_hip_dgetrs_dev", C.c_int, [_vp, _sz, _sz, _vp, _sz, _vp, _vp, _sz]),
    ("bola_hip_dpotrs", C.c_int, [_ch, _sz, _sz, _vp, _sz, _vp, _sz]),
    ("bola_hip_dpotrs_dev", C.c_int, [_vp, _ch, _sz, _sz, _vp, _sz, _vp, _sz]),
    ("bola_hip_dgesv_mixed", C.c_int, [C.c_int, _sz, _sz, _vp, _sz, _vp, _sz, _vp, _sz, _vp]),
    ("bola_hip_dgesv_mixed_dev", C.c_int, [_vp, C.c_int, _sz, _sz, _vp, _sz, _vp, _sz, _vp, _sz, _vp]),
    ("bola_hip_factor_create_dev", C.c_int, [_vp, C.c_int, C.c_int, _sz, _vp, _sz, _vp]),
    ("bola_hip_factor_solve_dev", C.c_int, [_vp, _vp, _sz, _vp, _sz, _vp, _sz, _vp]),
    ("bola_hip_factor_describe", C.c_int, [_vp, _vp]),
    ("bola_hip_factor_bytes", C.c_size_t, [C.c_int, C.c_int, _sz]),
    ("bola_hip_factor_destroy", C.c_int, [_vp]),
    ("bola_hip_sgemm_lp_dev", C.c_int, [_vp, C.c_int, _sz, _sz, _sz, _vp, _sz, _vp, _sz, _vp, _sz]),
    ("bola_hip_batch_solve_dev", C.c_int, [_vp, C.c_int, _sz, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("bola_hip_batch_tune_dev", C.c_int, [_vp, C.c_int, _sz, _vp, _vp, _vp]),
    ("bola_hip_batch_workspace_bytes", _sz, [_vp, _vp]),
    ("bola_hip_batch_graph_create", C.c_int, [_vp, C.c_int, _sz, _vp, _vp, _vp, _vp, _vp, _vp, C.POINTER(_vp)]),
    ("bola_hip_batch_graph_launch", C.c_int, [_vp]),
    ("bola_hip_batch_graph_destroy", C.c_int, [_vp]),
    ("bola_hip_server_create", C.c_int, [_sz, C.c_int, C.POINTER(_vp)]),
    ("bola_hip_server_submit", C.c_int, [_vp, _vp, C.POINTER(C.c_uint64)]),
    ("bola_hip_server_wait", C.c_int, [_vp, C.c_uint64, _vp]),
    ("bola_hip_server_poll", C.c_int, [_vp, C.c_uint64, _vp]),
    ("bola_hip_server_drain", C.c_int, [_vp]),
    ("bola_hip_server_get_stats", C.c_int, [_vp, _vp]),
    ("bola_hip_server_reset_stats", C.c_int, [_vp]),
    ("bola_hip_server_destroy", C.c_int, [_vp]),
]


def lib():
    """Load libbola_hip.so.  There is no fallback: without the built library every call
    raises."""
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            raise RuntimeError(f"{_SO} is not built; run bola_amd.build() (hipcc, gfx950)")
        L = C.CDLL(_SO)
        for name, res, args in _SIGS:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def declared_symbols():
    return [s[0] for s in _SIGS]


# ------------------------------------------------------------------ operand plumbing

def _is_dev(x) -> bool:
    return hasattr(x, "is_cuda") and x.is_cuda


def _ptr(x, dtype):
    if x is None:
        return None
    if _is_dev(x):
        import torch
        want = {np.float64: torch.float64, np.float32: torch.float32}.get(dtype, torch.int32)
        if x.dtype != want or not x.is_contiguous():
            raise TypeError(f"expected a contiguous {want} CUDA tensor")
        return C.c_void_p(x.data_ptr())
    if not isinstance(x, np.ndarray) or x.dtype != dtype or not x.flags.c_contiguous:
        raise TypeError(f"expected a contiguous numpy {np.dtype(dtype).name} array")
    return C.c_void_p(x.ctypes.data)


def _c(ch: str) -> bytes:
    return ch.encode("ascii")[:1]


class Handle:
    """A library handle (stream + device workspace).  By default it follows torch's
    current stream at every call."""

    def __init__(self, stream=None):
        h = C.c_void_p()
        rc = lib().bola_hip_create(C.byref(h))
        if rc != 0:
            raise BolaError(lib().bola_hip_last_error().decode(), rc)
        self._h = h
        self._fixed_stream = stream is not None
        if stream is not None:
            lib().bola_hip_set_stream(self._h, C.c_void_p(stream))

    @property
    def ptr(self):
        if not self._fixed_stream:
            import torch
            s = torch.cuda.current_stream().cuda_stream
            if lib().bola_hip_get_stream(self._h) != s:
                lib().bola_hip_set_stream(self._h, C.c_void_p(s))
        return self._h

    def error(self) -> str:
        return lib().bola_hip_error_string(self._h).decode()

    def set_exclusive(self, on: bool):
        """False: this handle's work shares the GPU (no CU-partitioned LU streams, no one-launch
        triangular solves)."""
        lib().bola_hip_set_exclusive(self._h, 1 if on else 0)

    def synchronize(self):
        lib().bola_hip_synchronize(self.ptr)

    def profile(self, on: bool = True):
        lib().bola_hip_profile_enable(self._h, 1 if on else 0)

    def profile_reset(self):
        lib().bola_hip_profile_reset(self._h)

    def profile_get(self, name: str):
        ms = C.c_double()
        calls = C.c_long()
        fl = C.c_double()
        lib().bola_hip_profile_get(self._h, name.encode(), C.byref(ms), C.byref(calls), C.byref(fl))
        return ms.value, calls.value, fl.value

    def profile_report(self, iterations: float = 1.0) -> str:
        buf = C.create_string_buffer(1 << 16)
        lib().bola_hip_profile_report(self._h, iterations, buf, len(buf))
        return buf.value.decode()

    def close(self):
        if self._h:
            lib().bola_hip_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_dev_handle = None


def default_handle() -> Handle:
    global _default_dev_handle
    if _default_dev_handle is None:
        _default_dev_handle = Handle()
    return _default_dev_handle


def _check(rc, dev):
    if rc != 0:
        msg = default_handle().error() if dev else lib().bola_hip_last_error().decode()
        raise BolaError(msg, rc)


# ------------------------------------------------------------------ BLAS-1/2

def idamax(n, x) -> int:
    """idamax.rs:23 -- 0-based index of the largest |x_i| (the crate's tie/NaN rule)."""
    if _is_dev(x):
        import torch
        out = torch.zeros(1, dtype=torch.int64, device=x.device)
        rc = lib().bola_hip_idamax_dev(default_handle().ptr, n, _ptr(x, np.float64), C.c_void_p(out.data_ptr()))
        _check(rc, True)
        return int(out.item())
    r = lib().bola_hip_idamax(n, _ptr(x, np.float64))
    if r == C.c_size_t(-1).value:
        raise BolaError(lib().bola_hip_last_error().decode(), ERR_DEVICE)
    return int(r)


def laswp(n, a, lda, k1, k2, ipiv, incx) -> None:
    """laswp.rs:22 -- row interchanges k1..k2 (1-based) applied to n columns."""
    if _is_dev(a):
        rc = lib().bola_hip_dlaswp_dev(default_handle().ptr, n, _ptr(a, np.float64), lda, k1, k2,
                                       _ptr(ipiv, np.int32), incx)
        _check(rc, True)
    else:
        _check(lib().bola_hip_dlaswp(n, _ptr(a, np.float64), lda, k1, k2, _ptr(ipiv, np.int32), incx), False)


def ger(m, n, alpha, x, y, ldy, a, lda) -> None:
    """ger.rs:28 -- A += alpha * x * y^T."""
    if _is_dev(a):
        _check(lib().bola_hip_dger_dev(default_handle().ptr, m, n, alpha, _ptr(x, np.float64), _ptr(y, np.float64),
                                       ldy, _ptr(a, np.float64), lda), True)
    else:
        _check(lib().bola_hip_dger(m, n, alpha, _ptr(x, np.float64), _ptr(y, np.float64), ldy,
                                   _ptr(a, np.float64), lda), False)


def scal(n, alpha, x, incx) -> None:
    """scal.rs:45 -- x *= alpha."""
    if _is_dev(x):
        _check(lib().bola_hip_dscal_dev(default_handle().ptr, n, alpha, _ptr(x, np.float64), incx), True)
    else:
        _check(lib().bola_hip_dscal(n, alpha, _ptr(x, np.float64), incx), False)


def daxpy(n, alpha, x, incx, y, incy) -> None:
    """utilities.rs:107 -- y += alpha * x."""
    if _is_dev(y):
        _check(lib().bola_hip_daxpy_dev(default_handle().ptr, n, alpha, _ptr(x, np.float64), incx,
                                        _ptr(y, np.float64), incy), True)
    else:
        _check(lib().bola_hip_daxpy(n, alpha, _ptr(x, np.float64), incx, _ptr(y, np.float64), incy), False)


def dot_product(n, x, incx, y, incy) -> float:
    """utilities.rs:22 -- sum_i x_i * y_i."""
    if _is_dev(x):
        import torch
        out = torch.zeros(1, dtype=torch.float64, device=x.device)
        _check(lib().bola_hip_ddot_dev(default_handle().ptr, n, _ptr(x, np.float64), incx, _ptr(y, np.float64), incy,
                                       C.c_void_p(out.data_ptr())), True)
        return float(out.item())
    return float(lib().bola_hip_ddot(n, _ptr(x, np.float64), incx, _ptr(y, np.float64), incy))


# ------------------------------------------------------------------ BLAS-3

def gemm(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc) -> None:
    """gemm.rs:88 -- C = alpha*op(A)*op(B) + beta*C."""
    if _is_dev(c):
        _check(lib().bola_hip_dgemm_dev(default_handle().ptr, _c(transa), _c(transb), m, n, k, alpha,
                                        _ptr(a, np.float64), lda, _ptr(b, np.float64), ldb, beta,
                                        _ptr(c, np.float64), ldc), True)
    else:
        _check(lib().bola_hip_dgemm(_c(transa), _c(transb), m, n, k, alpha, _ptr(a, np.float64), lda,
                                    _ptr(b, np.float64), ldb, beta, _ptr(c, np.float64), ldc), False)


def trsm(side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb) -> None:
    """trsm.rs:33 -- op(A) X = alpha B ('L') or X op(A) = alpha B ('R'); X overwrites B."""
    if _is_dev(b):
        _check(lib().bola_hip_dtrsm_dev(default_handle().ptr, _c(side), _c(uplo), _c(transa), _c(diag), m, n, alpha,
                                        _ptr(a, np.float64), lda, _ptr(b, np.float64), ldb), True)
    else:
        _check(lib().bola_hip_dtrsm(_c(side), _c(uplo), _c(transa), _c(diag), m, n, alpha, _ptr(a, np.float64), lda,
                                    _ptr(b, np.float64), ldb), False)


def syrk(uplo, trans, n, k, alpha, a, lda, beta, c, ldc) -> None:
    """syrk.rs:17 -- C = alpha*A*A^T + beta*C ('N') or alpha*A^T*A + beta*C ('T'), one triangle."""
    if _is_dev(c):
        _check(lib().bola_hip_dsyrk_dev(default_handle().ptr, _c(uplo), _c(trans), n, k, alpha, _ptr(a, np.float64),
                                        lda, beta, _ptr(c, np.float64), ldc), True)
    else:
        _check(lib().bola_hip_dsyrk(_c(uplo), _c(trans), n, k, alpha, _ptr(a, np.float64), lda, beta,
                                    _ptr(c, np.float64), ldc), False)


# ------------------------------------------------------------------ LAPACK

def getrf(m, n, a, lda, ipiv) -> None:
    """getrf.rs:37 -- P*A = L*U in place; raises BolaError like the crate's Err."""
    if _is_dev(a):
        _check(lib().bola_hip_dgetrf_dev(default_handle().ptr, m, n, _ptr(a, np.float64), lda,
                                         _ptr(ipiv, np.int32), None), True)
    else:
        _check(lib().bola_hip_dgetrf(m, n, _ptr(a, np.float64), lda, _ptr(ipiv, np.int32)), False)


def getrf_noalloc(m, n, a, lda, ipiv, handle=None) -> None:
    """getrf.rs:37 on device tensors with no device allocation: the library's static
    device scratch instead of the handle's workspaces (bit-identical to getrf)."""
    h = handle or default_handle()
    _check(lib().bola_hip_dgetrf_noalloc_dev(h.ptr, m, n, _ptr(a, np.float64), lda, _ptr(ipiv, np.int32), None),
           True)


def getrf2(m, n, a, lda, ipiv) -> None:
    """getrf2.rs:39 -- the crate's recursive/unblocked LU, callable directly."""
    if _is_dev(a):
        _check(lib().bola_hip_dgetrf2_dev(default_handle().ptr, m, n, _ptr(a, np.float64), lda,
                                          _ptr(ipiv, np.int32), None), True)
    else:
        _check(lib().bola_hip_dgetrf2(m, n, _ptr(a, np.float64), lda, _ptr(ipiv, np.int32)), False)


def potrf(uplo, n, a, lda) -> None:
    """potrf.rs:101 -- Cholesky in place ('L': A = L L^T, 'U': A = U^T U)."""
    if _is_dev(a):
        _check(lib().bola_hip_dpotrf_dev(default_handle().ptr, _c(uplo), n, _ptr(a, np.float64), lda, None), True)
    else:
        _check(lib().bola_hip_dpotrf(_c(uplo), n, _ptr(a, np.float64), lda), False)


def potrf2(uplo, n, a, lda) -> None:
    """potrf2.rs:25 -- the crate's recursive Cholesky, callable directly."""
    if _is_dev(a):
        _check(lib().bola_hip_dpotrf2_dev(default_handle().ptr, _c(uplo), n, _ptr(a, np.float64), lda, None), True)
    else:
        _check(lib().bola_hip_dpotrf2(_c(uplo), n, _ptr(a, np.float64), lda), False)


def getrs(n, nrhs, lu, lda, ipiv, b, ldb) -> None:
    """LU solve composed from the crate's parts: laswp, trsm(L,L,N,U), trsm(L,U,N,N)."""
    if _is_dev(b):
        _check(lib().bola_hip_dgetrs_dev(default_handle().ptr, n, nrhs, _ptr(lu, np.float64), lda,
                                         _ptr(ipiv, np.int32), _ptr(b, np.float64), ldb), True)
    else:
        _check(lib().bola_hip_dgetrs(n, nrhs, _ptr(lu, np.float64), lda, _ptr(ipiv, np.int32),
                                     _ptr(b, np.float64), ldb), False)


def potrs(uplo, n, nrhs, f, lda, b, ldb) -> None:
    """Cholesky solve composed from the crate's trsm: 'L' -> (L,L,N,N), (L,L,T,N)."""
    if _is_dev(b):
        _check(lib().bola_hip_dpotrs_dev(default_handle().ptr, _c(uplo), n, nrhs, _ptr(f, np.float64), lda,
                                         _ptr(b, np.float64), ldb), True)
    else:
        _check(lib().bola_hip_dpotrs(_c(uplo), n, nrhs, _ptr(f, np.float64), lda, _ptr(b, np.float64), ldb), False)


# ------------------------------------------------------------------ mixed precision
# No crate counterpart (the crate solves in FP64 only); see include/bola_hip.h.

PRECISIONS = {"fp32": 0, "fp16": 1, "bf16": 2, "fp16x3": 3}


class MixedStats(C.Structure):
    _fields_ = [("ir_iterations", C.c_int), ("gmres_cycles", C.c_int), ("gmres_iterations", C.c_int),
                ("fallback", C.c_int), ("converged", C.c_int), ("pad", C.c_int),
                ("first_berr", C.c_double), ("final_berr", C.c_double), ("tolerance", C.c_double),
                ("factor_ms", C.c_double), ("refine_ms", C.c_double), ("fallback_ms", C.c_double)]

    def as_dict(self):
        return {f[0]: getattr(self, f[0]) for f in self._fields_ if f[0] != "pad"}


MIX_RBT = 0x100


def dgesv_mixed(n, nrhs, a, lda, b, ldb, x, ldx, precision="fp32", rbt=False) -> dict:
    """A X = B to FP64 accuracy from a low-precision LU (fp32 / fp16 / bf16 products),
    FP64 refinement, GMRES-IR and the FP64 fallback.  ``rbt``: randomise with random
    butterflies and factor without interchanges instead of tournament pivoting.  Returns
    the solver statistics."""
    st = MixedStats()
    p = PRECISIONS[precision] | (MIX_RBT if rbt else 0)
    if _is_dev(a):
        rc = lib().bola_hip_dgesv_mixed_dev(default_handle().ptr, p, n, nrhs, _ptr(a, np.float64), lda,
                                            _ptr(b, np.float64), ldb, _ptr(x, np.float64), ldx, C.byref(st))
        _check(rc, True)
    else:
        rc = lib().bola_hip_dgesv_mixed(p, n, nrhs, _ptr(a, np.float64), lda, _ptr(b, np.float64), ldb,
                                        _ptr(x, np.float64), ldx, C.byref(st))
        _check(rc, False)
    return st.as_dict()


FACTOR_LU, FACTOR_CHOL_L, FACTOR_CHOL_U, FACTOR_MIXED = 0, 1, 2, 3


class FactorDesc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("precision", C.c_int32), ("n", C.c_size_t), ("bytes", C.c_size_t),
                ("factor", C.c_void_p), ("ld", C.c_size_t), ("ipiv", C.c_void_p), ("row_scale", C.c_void_p),
                ("col_scale", C.c_void_p), ("matrix", C.c_void_p), ("butterfly", C.c_int32), ("pad", C.c_int32)]

    def as_dict(self):
        return {f[0]: getattr(self, f[0]) for f in self._fields_ if f[0] != "pad"}


class Factor:
    """A factorization kept on the device for repeated right-hand sides (the C ABI's
    bola_hip_factor_*): kind FACTOR_LU / FACTOR_CHOL_L / FACTOR_CHOL_U (the reference's
    FP64 factorizations, bit-identical) or FACTOR_MIXED (the mixed-precision solver's
    factorization, `precision` and `rbt` as in dgesv_mixed).  `a`: numpy array or CUDA
    tensor, not modified."""

    def __init__(self, kind, n, a, lda=None, precision="fp32", rbt=False):
        p = (PRECISIONS[precision] | (MIX_RBT if rbt else 0)) if kind == FACTOR_MIXED else 0
        f = C.c_void_p()
        rc = lib().bola_hip_factor_create_dev(default_handle().ptr, kind, p, n, _ptr(a, np.float64), lda or n,
                                              C.byref(f))
        _check(rc, True)
        self._f = f
        self.n = n
        self.kind = kind

    def solve(self, nrhs, b, x, ldb=None, ldx=None) -> dict:
        """x = A^-1 b on CUDA tensors (x may be b for the FP64 kinds); solver statistics."""
        st = MixedStats()
        rc = lib().bola_hip_factor_solve_dev(default_handle().ptr, self._f, nrhs, _ptr(b, np.float64),
                                             ldb or self.n, _ptr(x, np.float64), ldx or self.n, C.byref(st))
        _check(rc, True)
        return st.as_dict()

    def describe(self) -> dict:
        d = FactorDesc()
        lib().bola_hip_factor_describe(self._f, C.byref(d))
        return d.as_dict()

    def close(self) -> None:
        if getattr(self, "_f", None):
            lib().bola_hip_factor_destroy(self._f)
            self._f = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def factor_bytes(kind, n, precision="fp32", rbt=False) -> int:
    p = (PRECISIONS[precision] | (MIX_RBT if rbt else 0)) if kind == FACTOR_MIXED else 0
    return int(lib().bola_hip_factor_bytes(kind, p, n))


def serve_mixed(precision="fp32", rbt=False) -> int:
    """request precision code of a mixed-precision served LU solve"""
    return 1 + (PRECISIONS[precision] | (MIX_RBT if rbt else 0))


def sgemm_lp(m, n, k, a, lda, b, ldb, c, ldc, precision="fp32") -> None:
    """C -= A B on fp32 device tensors, products in `precision` (the mixed LU's update)."""
    _check(lib().bola_hip_sgemm_lp_dev(default_handle().ptr, PRECISIONS[precision], m, n, k, _ptr(a, np.float32),
                                       lda, _ptr(b, np.float32), ldb, _ptr(c, np.float32), ldc), True)


# ------------------------------------------------------------------ batched solves

BATCH_LU, BATCH_CHOL = 0, 1
BATCH_FP64, BATCH_FP32_IR = 0, 1
# the C struct bola_hip_batch_desc, field for field
BATCH_DESC_DTYPE = np.dtype([("n", "<i4"), ("nrhs", "<i4"), ("kind", "<i4"), ("pad", "<i4"),
                             ("a", "<i8"), ("b", "<i8"), ("piv", "<i8")])


class BatchConfig(C.Structure):
    _fields_ = [("nb", C.c_int32), ("wg_per_cu", C.c_int32)]

    def as_tuple(self):
        return (self.nb, self.wg_per_cu)


def _descs(descs):
    d = np.ascontiguousarray(descs, dtype=BATCH_DESC_DTYPE)
    return d, C.c_void_p(d.ctypes.data)


def batch_solve(variant, descs, arena, piv, info, iters=None, config=None) -> None:
    """Solve every system of the batch in one persistent launch (see include/bola_hip.h).
    descs: numpy array of BATCH_DESC_DTYPE (host); arena (float64), piv and info (int32)
    and the optional iters (int32): torch CUDA tensors.  config: (nb, wg_per_cu) or None
    for the tuned / default choice."""
    d, dp = _descs(descs)
    cfg = None
    if config is not None:
        cfg = C.byref(BatchConfig(*config))
    rc = lib().bola_hip_batch_solve_dev(default_handle().ptr, variant, len(d), dp, _ptr(arena, np.float64),
                                        _ptr(piv, np.int32), _ptr(info, np.int32), _ptr(iters, np.int32), cfg)
    _check(rc, True)


def batch_tune(variant, descs, arena):
    """Time every launch configuration on a sample of the batch; the fastest is kept for
    batches like this one and returned as (nb, wg_per_cu).  The arena is not modified."""
    d, dp = _descs(descs)
    out = BatchConfig()
    rc = lib().bola_hip_batch_tune_dev(default_handle().ptr, variant, len(d), dp, _ptr(arena, np.float64),
                                       C.byref(out))
    _check(rc, True)
    return out.as_tuple()


class BatchGraph:
    """A captured batched solve (hipGraph) of fixed descriptors and buffers; launch() replays
    it on the handle's stream."""

    def __init__(self, variant, descs, arena, piv, info, iters=None, config=None):
        d, dp = _descs(descs)
        cfg = C.byref(BatchConfig(*config)) if config is not None else None
        g = C.c_void_p()
        self._keep = (arena, piv, info, iters)
        rc = lib().bola_hip_batch_graph_create(default_handle().ptr, variant, len(d), dp, _ptr(arena, np.float64),
                                               _ptr(piv, np.int32), _ptr(info, np.int32), _ptr(iters, np.int32),
                                               cfg, C.byref(g))
        _check(rc, True)
        self._g = g

    def launch(self) -> None:
        _check(lib().bola_hip_batch_graph_launch(self._g), True)

    def close(self) -> None:
        if getattr(self, "_g", None):
            lib().bola_hip_batch_graph_destroy(self._g)
            self._g = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------------ serving

SERVE_LU, SERVE_CHOL = 0, 1
ERR_CANCELLED = -1001


class _Request(C.Structure):
    _fields_ = [("kind", C.c_int32), ("precision", C.c_int32), ("n", C.c_size_t), ("nrhs", C.c_size_t),
                ("a", C.c_void_p), ("b", C.c_void_p), ("x", C.c_void_p), ("factor_key", C.c_uint64)]


class _Result(C.Structure):
    _fields_ = [("status", C.c_int32), ("info", C.c_int32), ("cache_hit", C.c_int32), ("slot", C.c_int32),
                ("queue_us", C.c_double), ("latency_us", C.c_double)]

    def as_dict(self):
        return {f[0]: getattr(self, f[0]) for f in self._fields_}


class _ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("rejected", C.c_uint64),
                ("cache_hits", C.c_uint64), ("evictions", C.c_uint64), ("p50_us", C.c_double),
                ("p99_us", C.c_double), ("mean_us", C.c_double), ("solves_per_s", C.c_double),
                ("hbm_in_use", C.c_size_t), ("hbm_peak", C.c_size_t), ("hbm_budget", C.c_size_t),
                ("hbm_cached", C.c_size_t)]

    def as_dict(self):
        return {f[0]: getattr(self, f[0]) for f in self._fields_}


class Server:
    """The C ABI's solve server: `slots` workers with their own streams admit requests as
    slots and the HBM budget allow, overlap them on the device, and keep factorizations
    with a nonzero key resident for later right-hand sides.  Operands are numpy arrays
    or torch CUDA tensors (float64, column-major flat); they must stay alive until the
    ticket is waited for."""

    def __init__(self, hbm_budget: int, slots: int = 4):
        lib()
        s = C.c_void_p()
        rc = lib().bola_hip_server_create(hbm_budget, slots, C.byref(s))
        if rc != 0:
            raise BolaError(lib().bola_hip_last_error().decode(), rc)
        self._s = s
        self._live = {}

    def submit(self, kind, n, nrhs, a, b, x, key=0, precision=0) -> int:
        """precision: 0 (FP64, the reference's factorization) or serve_mixed(...) (LU only)."""
        r = _Request(kind, precision, n, nrhs, _ptr(a, np.float64) if a is not None else None, _ptr(b, np.float64),
                     _ptr(x, np.float64), key)
        t = C.c_uint64()
        rc = lib().bola_hip_server_submit(self._s, C.byref(r), C.byref(t))
        if rc != 0:
            raise BolaError(lib().bola_hip_last_error().decode(), rc)
        self._live[t.value] = (a, b, x)
        return t.value

    def wait(self, ticket) -> dict:
        res = _Result()
        rc = lib().bola_hip_server_wait(self._s, ticket, C.byref(res))
        self._live.pop(ticket, None)
        if rc != 0:
            raise BolaError("unknown ticket", rc)
        return res.as_dict()

    def poll(self, ticket):
        res = _Result()
        rc = lib().bola_hip_server_poll(self._s, ticket, C.byref(res))
        if rc == 1:
            return None
        self._live.pop(ticket, None)
        if rc != 0:
            raise BolaError("unknown ticket", rc)
        return res.as_dict()

    def drain(self) -> None:
        lib().bola_hip_server_drain(self._s)

    def stats(self) -> dict:
        st = _ServerStats()
        lib().bola_hip_server_get_stats(self._s, C.byref(st))
        return st.as_dict()

    def reset_stats(self) -> None:
        lib().bola_hip_server_reset_stats(self._s)

    def close(self) -> None:
        if getattr(self, "_s", None):
            lib().bola_hip_server_destroy(self._s)
            self._s = None
            self._live.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
