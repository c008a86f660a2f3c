"""ctypes binding of the CPU oracle (libbola_oracle.so, built from bola_oracle.c).

TEST INFRASTRUCTURE ONLY: the oracle restates the reference crate's CPU algorithm so
that tests can check the GPU library against it, and so that the benchmark can time the
reference algorithm on the host beside the GPU.  The product package (bola_amd) never
imports this module.

Arrays are flat float64 / int32 numpy arrays in column-major order; a matrix operand may
be passed as a flat slice that starts at the operand's first element (``a[j + j*lda:]``).
Routines that return ``Result<(), String>`` in the reference return ``(code, message)``
here: ``code == 0`` is Ok.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "libbola_oracle.so")
_lib = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_sz = C.c_size_t


def build() -> None:
    """Compile libbola_oracle.so with the recipe in oracle/Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = C.CDLL(_SO)
        L.bola_oracle_last_error.restype = C.c_char_p
        L.bola_oracle_dot_product.restype = C.c_double
        L.bola_oracle_dot_product.argtypes = [_sz, _dp, _sz, _dp, _sz]
        L.bola_oracle_daxpy.argtypes = [_sz, C.c_double, _dp, _sz, _dp, _sz]
        L.bola_oracle_scal_kernel.restype = _sz
        L.bola_oracle_scal_kernel.argtypes = [_sz, C.c_double, _dp]
        L.bola_oracle_scal.argtypes = [_sz, C.c_double, _dp, _sz]
        L.bola_oracle_idamax.restype = _sz
        L.bola_oracle_idamax.argtypes = [_sz, _dp]
        L.bola_oracle_laswp.argtypes = [_sz, _dp, _sz, C.c_int32, C.c_int32, _ip, C.c_int32]
        L.bola_oracle_ger.argtypes = [_sz, _sz, C.c_double, _dp, _dp, _sz, _dp, _sz]
        L.bola_oracle_gemm.argtypes = [C.c_char, C.c_char, _sz, _sz, _sz, C.c_double, _dp, _sz, _dp, _sz,
                                       C.c_double, _dp, _sz]
        L.bola_oracle_trsm.argtypes = [C.c_char, C.c_char, C.c_char, C.c_char, _sz, _sz, C.c_double, _dp, _sz,
                                       _dp, _sz]
        L.bola_oracle_syrk.argtypes = [C.c_char, C.c_char, _sz, _sz, C.c_double, _dp, _sz, C.c_double, _dp, _sz]
        L.bola_oracle_getrf2.restype = C.c_int
        L.bola_oracle_getrf2.argtypes = [_sz, _sz, _dp, _sz, _ip]
        L.bola_oracle_getrf.restype = C.c_int
        L.bola_oracle_getrf.argtypes = [_sz, _sz, _dp, _sz, _ip]
        L.bola_oracle_potrf2.restype = C.c_int
        L.bola_oracle_potrf2.argtypes = [C.c_char, _sz, _dp, _sz]
        L.bola_oracle_potrf.restype = C.c_int
        L.bola_oracle_potrf.argtypes = [C.c_char, _sz, _dp, _sz]
        L.bola_oracle_num_threads.restype = C.c_int
        L.bola_oracle_set_num_threads.argtypes = [C.c_int]
        _lib = L
    return _lib


def _d(x):
    assert x.dtype == np.float64 and x.flags.c_contiguous, "flat contiguous float64 array expected"
    return x.ctypes.data_as(_dp)


def _i(x):
    assert x.dtype == np.int32 and x.flags.c_contiguous, "flat contiguous int32 array expected"
    return x.ctypes.data_as(_ip)


def _c(ch: str) -> bytes:
    return ch.encode("ascii")


def _result(code: int):
    if code == 0:
        return 0, ""
    return code, lib().bola_oracle_last_error().decode()


def dot_product(n, x, incx, y, incy) -> float:
    return lib().bola_oracle_dot_product(n, _d(x), incx, _d(y), incy)


def daxpy(n, alpha, x, incx, y, incy) -> None:
    lib().bola_oracle_daxpy(n, alpha, _d(x), incx, _d(y), incy)


def scal_kernel(m, inv_diag, col) -> int:
    return lib().bola_oracle_scal_kernel(m, inv_diag, _d(col))


def scal(n, alpha, x, incx) -> None:
    lib().bola_oracle_scal(n, alpha, _d(x), incx)


def idamax(n, x) -> int:
    return lib().bola_oracle_idamax(n, _d(x))


def laswp(n, a, lda, k1, k2, ipiv, incx) -> None:
    lib().bola_oracle_laswp(n, _d(a), lda, k1, k2, _i(ipiv), incx)


def ger(m, n, alpha, x, y, ldy, a, lda) -> None:
    lib().bola_oracle_ger(m, n, alpha, _d(x), _d(y), ldy, _d(a), lda)


def gemm(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc) -> None:
    lib().bola_oracle_gemm(_c(transa), _c(transb), m, n, k, alpha, _d(a), lda, _d(b), ldb, beta, _d(c), ldc)


def trsm(side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb) -> None:
    lib().bola_oracle_trsm(_c(side), _c(uplo), _c(transa), _c(diag), m, n, alpha, _d(a), lda, _d(b), ldb)


def syrk(uplo, trans, n, k, alpha, a, lda, beta, c, ldc) -> None:
    lib().bola_oracle_syrk(_c(uplo), _c(trans), n, k, alpha, _d(a), lda, beta, _d(c), ldc)


def getrf2(m, n, a, lda, ipiv):
    return _result(lib().bola_oracle_getrf2(m, n, _d(a), lda, _i(ipiv)))


def getrf(m, n, a, lda, ipiv):
    return _result(lib().bola_oracle_getrf(m, n, _d(a), lda, _i(ipiv)))


def potrf2(uplo, n, a, lda):
    return _result(lib().bola_oracle_potrf2(_c(uplo), n, _d(a), lda))


def potrf(uplo, n, a, lda):
    return _result(lib().bola_oracle_potrf(_c(uplo), n, _d(a), lda))


def num_threads() -> int:
    return lib().bola_oracle_num_threads()


def set_num_threads(n: int) -> None:
    lib().bola_oracle_set_num_threads(n)


# Solve compositions built from the reference's parts (the crate has no getrs/potrs):
# LU: laswp(nrhs, B, 1, n, ipiv, 1) -> trsm(L,L,N,U) -> trsm(L,U,N,N).
# Cholesky: 'L' -> trsm(L,L,N,N), trsm(L,L,T,N); 'U' -> trsm(L,U,T,N), trsm(L,U,N,N).

def getrs(n, nrhs, lu, lda, ipiv, b, ldb) -> None:
    laswp(nrhs, b, ldb, 1, n, ipiv, 1)
    trsm("L", "L", "N", "U", n, nrhs, 1.0, lu, lda, b, ldb)
    trsm("L", "U", "N", "N", n, nrhs, 1.0, lu, lda, b, ldb)


def potrs(uplo, n, nrhs, f, lda, b, ldb) -> None:
    if uplo == "L":
        trsm("L", "L", "N", "N", n, nrhs, 1.0, f, lda, b, ldb)
        trsm("L", "L", "T", "N", n, nrhs, 1.0, f, lda, b, ldb)
    else:
        trsm("L", "U", "T", "N", n, nrhs, 1.0, f, lda, b, ldb)
        trsm("L", "U", "N", "N", n, nrhs, 1.0, f, lda, b, ldb)
