"""CORAL's call surface over host numpy buffers (the drop-in, synchronous form).

Names, argument order and meaning follow the reference crate ``coral-aarch64``:
column-major storage, ``lda``/``incx`` in elements, complex data interleaved
(a complex64 array, or a float32 array of [re, im] pairs), complex scalars as
two-element sequences, ``i?amax`` returning a 0-based index.  Output buffers are
updated in place, as CORAL's ``&mut [T]`` arguments are.

Every function calls the ``coralh_*`` entry point of libcoralhip.so; errors come
back as ``CoralError`` (CORAL itself panics in debug builds and has undefined
behaviour in release builds on a bad length/stride).

Reference signatures: level3/sgemm.rs:106, dgemm.rs:107, cgemm.rs:127, zgemm.rs:123;
level2/dgemv.rs:61, cgemv.rs:67, dsymv.rs:99, chemv.rs:96, dtrsv.rs:62, dtrmv.rs:58,
sger.rs:68, cgeru.rs:63, cgerc.rs:62, dsyr.rs:61, dsyr2.rs:63, cher.rs:62, cher2.rs:65;
level1/sdot.rs:42, cdotc.rs:46, daxpy.rs:42, snrm2.rs:45, isamax.rs:43, sscal.rs:30,
csscal.rs:39, sasum.rs:40, scopy.rs:28, sswap.rs:35, srot.rs:51, srotm.rs:47, srotg.rs:40,
srotmg.rs:42 (all under coral-aarch64/src/).
"""
from __future__ import annotations

import ctypes
from enum import IntEnum

import numpy as np

from ._lib import call


class CoralTranspose(IntEnum):  # coral-aarch64/src/enums.rs:2-6
    NoTranspose = 0
    Transpose = 1
    ConjugateTranspose = 2


class CoralTriangular(IntEnum):  # coral-aarch64/src/enums.rs:9-12
    UpperTriangular = 0
    LowerTriangular = 1


class CoralDiagonal(IntEnum):  # coral-aarch64/src/enums.rs:15-18
    UnitDiagonal = 0
    NonUnitDiagonal = 1


_REAL = {"s": np.float32, "d": np.float64, "c": np.float32, "z": np.float64}


def _buf(a, prefix: str, writable: bool = False):
    """Pointer to a host buffer of the routine's element type (or None)."""
    if a is None:
        return None
    a = np.asarray(a)
    real = _REAL[prefix]
    if a.dtype.kind == "c":
        want = np.complex64 if real == np.float32 else np.complex128
        if a.dtype != want:
            raise TypeError(f"expected {np.dtype(want)}, got {a.dtype}")
    elif a.dtype != real:
        raise TypeError(f"expected {np.dtype(real)}, got {a.dtype}")
    if not a.flags["C_CONTIGUOUS"] and not a.flags["F_CONTIGUOUS"]:
        raise ValueError("buffers must be contiguous (strides are passed as lda/inc)")
    if writable and not a.flags["WRITEABLE"]:
        raise ValueError("output buffer is read-only")
    return a.ctypes.data


def _scalar2(v, real):
    arr = (ctypes.c_float * 2)() if real == np.float32 else (ctypes.c_double * 2)()
    v = complex(v[0], v[1]) if isinstance(v, (list, tuple, np.ndarray)) else complex(v)
    arr[0], arr[1] = v.real, v.imag
    return arr


# ---------------------------------------------------------------- level 3
def _gemm(p, name):
    cplx = p in "cz"

    def f(op_a, op_b, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
        al = _scalar2(alpha, _REAL[p]) if cplx else alpha
        be = _scalar2(beta, _REAL[p]) if cplx else beta
        call(name, int(op_a), int(op_b), m, n, k, al, _buf(a, p), lda, _buf(b, p), ldb, be,
             _buf(c, p, True), ldc)

    f.__name__ = name.replace("coralh_", "")
    return f


sgemm = _gemm("s", "coralh_sgemm")
dgemm = _gemm("d", "coralh_dgemm")
cgemm = _gemm("c", "coralh_cgemm")
zgemm = _gemm("z", "coralh_zgemm")
sgemm_bf16x3 = _gemm("s", "coralh_sgemm_bf16x3")
cgemm_3m_bf16x3 = _gemm("c", "coralh_cgemm_3m_bf16x3")
sgemm_fp16x3 = _gemm("s", "coralh_sgemm_fp16x3")
cgemm_3m_fp16x3 = _gemm("c", "coralh_cgemm_3m_fp16x3")


# ---------------------------------------------------------------- level 2
def _gemv(p):
    cplx = p in "cz"

    def f(trans, n_rows, n_cols, alpha, matrix, lda, x, incx, beta, y, incy):
        al = _scalar2(alpha, _REAL[p]) if cplx else alpha
        be = _scalar2(beta, _REAL[p]) if cplx else beta
        call(f"coralh_{p}gemv", int(trans), n_rows, n_cols, al, _buf(matrix, p), lda, _buf(x, p), incx, be,
             _buf(y, p, True), incy)

    f.__name__ = f"{p}gemv"
    return f


sgemv, dgemv, cgemv, zgemv = (_gemv(p) for p in "sdcz")


def _symv(p, name):
    cplx = p in "cz"

    def f(uplo, n, alpha, matrix, lda, x, incx, beta, y, incy):
        al = _scalar2(alpha, _REAL[p]) if cplx else alpha
        be = _scalar2(beta, _REAL[p]) if cplx else beta
        call(f"coralh_{name}", int(uplo), n, al, _buf(matrix, p), lda, _buf(x, p), incx, be, _buf(y, p, True), incy)

    f.__name__ = name
    return f


ssymv = _symv("s", "ssymv")
dsymv = _symv("d", "dsymv")
chemv = _symv("c", "chemv")
zhemv = _symv("z", "zhemv")


def _tr(p, what):
    def f(uplo, transpose, diagonal, n, matrix, lda, x, incx):
        call(f"coralh_{p}{what}", int(uplo), int(transpose), int(diagonal), n, _buf(matrix, p), lda,
             _buf(x, p, True), incx)

    f.__name__ = f"{p}{what}"
    return f


strsv, dtrsv, ctrsv, ztrsv = (_tr(p, "trsv") for p in "sdcz")
strmv, dtrmv, ctrmv, ztrmv = (_tr(p, "trmv") for p in "sdcz")


def _ger(p, name):
    cplx = p in "cz"

    def f(n_rows, n_cols, alpha, x, incx, y, incy, matrix, lda):
        al = _scalar2(alpha, _REAL[p]) if cplx else alpha
        call(f"coralh_{name}", n_rows, n_cols, al, _buf(x, p), incx, _buf(y, p), incy, _buf(matrix, p, True), lda)

    f.__name__ = name
    return f


sger, dger = _ger("s", "sger"), _ger("d", "dger")
cgeru, cgerc = _ger("c", "cgeru"), _ger("c", "cgerc")
zgeru, zgerc = _ger("z", "zgeru"), _ger("z", "zgerc")


def _syr(p, name):
    def f(uplo, n, alpha, x, incx, matrix, lda):  # alpha is real, also for cher/zher
        call(f"coralh_{name}", int(uplo), n, alpha, _buf(x, p), incx, _buf(matrix, p, True), lda)

    f.__name__ = name
    return f


ssyr, dsyr, cher, zher = _syr("s", "ssyr"), _syr("d", "dsyr"), _syr("c", "cher"), _syr("z", "zher")


def _syr2(p, name):
    cplx = p in "cz"

    def f(uplo, n, alpha, x, incx, y, incy, matrix, lda):
        al = _scalar2(alpha, _REAL[p]) if cplx else alpha
        call(f"coralh_{name}", int(uplo), n, al, _buf(x, p), incx, _buf(y, p), incy, _buf(matrix, p, True), lda)

    f.__name__ = name
    return f


ssyr2, dsyr2 = _syr2("s", "ssyr2"), _syr2("d", "dsyr2")
cher2, zher2 = _syr2("c", "cher2"), _syr2("z", "zher2")

# ---------------------------------------------------------------- level 1
def _dot(p, name):
    cplx = p in "cz"
    real = _REAL[p]

    def f(n, x, incx, y, incy):
        out = np.zeros(2 if cplx else 1, real)
        call(f"coralh_{name}", n, _buf(x, p), incx, _buf(y, p), incy, out.ctypes.data)
        return [out[0], out[1]] if cplx else out[0]

    f.__name__ = name
    return f


sdot = _dot("s", "sdot")
ddot = _dot("d", "ddot")
cdotu = _dot("c", "cdotu")
cdotc = _dot("c", "cdotc")
zdotu = _dot("z", "zdotu")
zdotc = _dot("z", "zdotc")


def _axpy(p):
    cplx = p in "cz"

    def f(n, alpha, x, incx, y, incy):
        al = _scalar2(alpha, _REAL[p]) if cplx else alpha
        call(f"coralh_{p}axpy", n, al, _buf(x, p), incx, _buf(y, p, True), incy)

    f.__name__ = f"{p}axpy"
    return f


saxpy, daxpy, caxpy, zaxpy = (_axpy(p) for p in "sdcz")


def _nrm2(p, name):
    real = _REAL[p]

    def f(n, x, incx):
        out = np.zeros(1, real)
        call(f"coralh_{name}", n, _buf(x, p), incx, out.ctypes.data)
        return out[0]

    f.__name__ = name
    return f


snrm2 = _nrm2("s", "snrm2")
dnrm2 = _nrm2("d", "dnrm2")
scnrm2 = _nrm2("c", "scnrm2")
dznrm2 = _nrm2("z", "dznrm2")


def _iamax(p, name):
    def f(n, x, incx):
        out = np.zeros(1, np.int64)
        call(f"coralh_{name}", n, _buf(x, p), incx, out.ctypes.data)
        return int(out[0])

    f.__name__ = name
    return f


isamax = _iamax("s", "isamax")
idamax = _iamax("d", "idamax")
icamax = _iamax("c", "icamax")
izamax = _iamax("z", "izamax")


def _scal(p, name, real_alpha):
    cplx_alpha = p in "cz" and not real_alpha

    def f(n, alpha, x, incx):
        al = _scalar2(alpha, _REAL[p]) if cplx_alpha else alpha
        call(f"coralh_{name}", n, al, _buf(x, p, True), incx)

    f.__name__ = name
    return f


sscal, dscal = _scal("s", "sscal", True), _scal("d", "dscal", True)
cscal, zscal = _scal("c", "cscal", False), _scal("z", "zscal", False)
csscal, zdscal = _scal("c", "csscal", True), _scal("z", "zdscal", True)

sasum = _nrm2("s", "sasum")
dasum = _nrm2("d", "dasum")
scasum = _nrm2("c", "scasum")
dzasum = _nrm2("z", "dzasum")


def _xy(p, name, x_out):
    def f(n, x, incx, y, incy):
        call(f"coralh_{name}", n, _buf(x, p, x_out), incx, _buf(y, p, True), incy)

    f.__name__ = name
    return f


scopy, dcopy, ccopy, zcopy = (_xy(p, f"{p}copy", False) for p in "sdcz")
sswap, dswap, cswap, zswap = (_xy(p, f"{p}swap", True) for p in "sdcz")


def _rot(p, name):
    def f(n, x, incx, y, incy, c, s):
        call(f"coralh_{name}", n, _buf(x, p, True), incx, _buf(y, p, True), incy, c, s)

    f.__name__ = name
    return f


srot, drot, csrot, zdrot = _rot("s", "srot"), _rot("d", "drot"), _rot("c", "csrot"), _rot("z", "zdrot")


def _rotm(p, name):
    def f(n, x, incx, y, incy, param):
        pa = np.ascontiguousarray(param, _REAL[p])
        call(f"coralh_{name}", n, _buf(x, p, True), incx, _buf(y, p, True), incy, pa.ctypes.data)

    f.__name__ = name
    return f


srotm, drotm = _rotm("s", "srotm"), _rotm("d", "drotm")


def _rotg(p, name):
    def f(a, b):
        """CORAL's rotg on scalars: returns (r, z, c, s), the values CORAL leaves in a, b, c, s."""
        v = np.array([a, b, 0, 0], _REAL[p])
        base, sz = v.ctypes.data, v.itemsize
        call(f"coralh_{name}", base, base + sz, base + 2 * sz, base + 3 * sz)
        return tuple(v)

    f.__name__ = name
    return f


srotg, drotg = _rotg("s", "srotg"), _rotg("d", "drotg")


def _rotmg(p, name):
    def f(d1, d2, x1, y1):
        """CORAL's rotmg: returns (d1, d2, x1, param) as CORAL leaves them."""
        v = np.array([d1, d2, x1], _REAL[p])
        param = np.zeros(5, _REAL[p])
        base, sz = v.ctypes.data, v.itemsize
        call(f"coralh_{name}", base, base + sz, base + 2 * sz, y1, param.ctypes.data)
        return v[0], v[1], v[2], param

    f.__name__ = name
    return f


srotmg, drotmg = _rotmg("s", "srotmg"), _rotmg("d", "drotmg")
