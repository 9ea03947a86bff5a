"""ctypes front end of oracle/libcoral_port.so, the CPU restatement of CORAL.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker or the timed CPU baseline.  The
product path (coral_amd) never imports it.

Functions take the same arguments as ``coral_amd.blas`` (CORAL's argument order,
host numpy buffers updated in place), so a test can run one call on both.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcoral_port.so")

i32, i64, f32, f64, P = ctypes.c_int, ctypes.c_int64, ctypes.c_float, ctypes.c_double, ctypes.c_void_p

_SIGS = {
    "port_sgemm": (None, [i32, i32, i64, i64, i64, f32, P, i64, P, i64, f32, P, i64]),
    "port_dgemm": (None, [i32, i32, i64, i64, i64, f64, P, i64, P, i64, f64, P, i64]),
    "port_cgemm": (None, [i32, i32, i64, i64, i64, P, P, i64, P, i64, P, P, i64]),
    "port_zgemm": (None, [i32, i32, i64, i64, i64, P, P, i64, P, i64, P, P, i64]),
    "port_sgemv": (None, [i32, i64, i64, f32, P, i64, P, i64, f32, P, i64]),
    "port_dgemv": (None, [i32, i64, i64, f64, P, i64, P, i64, f64, P, i64]),
    "port_cgemv": (None, [i32, i64, i64, P, P, i64, P, i64, P, P, i64]),
    "port_zgemv": (None, [i32, i64, i64, P, P, i64, P, i64, P, P, i64]),
    "port_ssymv": (None, [i32, i64, f32, P, i64, P, i64, f32, P, i64]),
    "port_dsymv": (None, [i32, i64, f64, P, i64, P, i64, f64, P, i64]),
    "port_chemv": (None, [i32, i64, P, P, i64, P, i64, P, P, i64]),
    "port_zhemv": (None, [i32, i64, P, P, i64, P, i64, P, P, i64]),
    "port_sdot": (f32, [i64, P, i64, P, i64]),
    "port_ddot": (f64, [i64, P, i64, P, i64]),
    "port_snrm2": (f32, [i64, P, i64]),
    "port_dnrm2": (f64, [i64, P, i64]),
    "port_scnrm2": (f32, [i64, P, i64]),
    "port_dznrm2": (f64, [i64, P, i64]),
    "port_saxpy": (None, [i64, f32, P, i64, P, i64]),
    "port_daxpy": (None, [i64, f64, P, i64, P, i64]),
    "port_caxpy": (None, [i64, P, P, i64, P, i64]),
    "port_zaxpy": (None, [i64, P, P, i64, P, i64]),
}
for _p in "sdcz":
    for _w in ("trsv", "trmv"):
        _SIGS[f"port_{_p}{_w}"] = (None, [i32, i32, i32, i64, P, i64, P, i64])
    _SIGS[f"port_i{_p}amax"] = (i64, [i64, P, i64])
for _n in ("cdotu", "cdotc", "zdotu", "zdotc"):
    _SIGS[f"port_{_n}"] = (None, [i64, P, i64, P, i64, P])
# level-2 rank updates
_SIGS["port_sger"] = (None, [i64, i64, f32, P, i64, P, i64, P, i64])
_SIGS["port_dger"] = (None, [i64, i64, f64, P, i64, P, i64, P, i64])
for _n in ("cgeru", "cgerc", "zgeru", "zgerc"):
    _SIGS[f"port_{_n}"] = (None, [i64, i64, P, P, i64, P, i64, P, i64])
_SIGS["port_ssyr"] = (None, [i32, i64, f32, P, i64, P, i64])
_SIGS["port_dsyr"] = (None, [i32, i64, f64, P, i64, P, i64])
_SIGS["port_cher"] = (None, [i32, i64, f32, P, i64, P, i64])
_SIGS["port_zher"] = (None, [i32, i64, f64, P, i64, P, i64])
_SIGS["port_ssyr2"] = (None, [i32, i64, f32, P, i64, P, i64, P, i64])
_SIGS["port_dsyr2"] = (None, [i32, i64, f64, P, i64, P, i64, P, i64])
_SIGS["port_cher2"] = (None, [i32, i64, P, P, i64, P, i64, P, i64])
_SIGS["port_zher2"] = (None, [i32, i64, P, P, i64, P, i64, P, i64])
# rest of level 1
_SIGS["port_sasum"] = (f32, [i64, P, i64])
_SIGS["port_dasum"] = (f64, [i64, P, i64])
_SIGS["port_scasum"] = (f32, [i64, P, i64])
_SIGS["port_dzasum"] = (f64, [i64, P, i64])
for _p in "sdcz":
    _SIGS[f"port_{_p}copy"] = (None, [i64, P, i64, P, i64])
    _SIGS[f"port_{_p}swap"] = (None, [i64, P, i64, P, i64])
_SIGS["port_srot"] = (None, [i64, P, i64, P, i64, f32, f32])
_SIGS["port_drot"] = (None, [i64, P, i64, P, i64, f64, f64])
_SIGS["port_csrot"] = (None, [i64, P, i64, P, i64, f32, f32])
_SIGS["port_zdrot"] = (None, [i64, P, i64, P, i64, f64, f64])
_SIGS["port_srotm"] = (None, [i64, P, i64, P, i64, P])
_SIGS["port_drotm"] = (None, [i64, P, i64, P, i64, P])
_SIGS["port_srotg"] = (None, [P, P, P, P])
_SIGS["port_drotg"] = (None, [P, P, P, P])
_SIGS["port_srotmg"] = (None, [P, P, P, f32, P])
_SIGS["port_drotmg"] = (None, [P, P, P, f64, P])
_SIGS["port_sscal"] = (None, [i64, f32, P, i64])
_SIGS["port_dscal"] = (None, [i64, f64, P, i64])
_SIGS["port_csscal"] = (None, [i64, f32, P, i64])
_SIGS["port_zdscal"] = (None, [i64, f64, P, i64])
_SIGS["port_cscal"] = (None, [i64, P, P, i64])
_SIGS["port_zscal"] = (None, [i64, P, P, i64])

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


_REAL = {"s": np.float32, "d": np.float64, "c": np.float32, "z": np.float64}


def _ptr(a):
    return None if a is None else np.asarray(a).ctypes.data


def _s2(v, p):
    real = _REAL[p]
    arr = np.zeros(2, real)
    v = complex(v[0], v[1]) if isinstance(v, (list, tuple, np.ndarray)) else complex(v)
    arr[0], arr[1] = v.real, v.imag
    return arr


def _mk_gemm(p):
    def f(op_a, op_b, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
        fn = getattr(load(), f"port_{p}gemm")
        if p in "cz":
            al, be = _s2(alpha, p), _s2(beta, p)
            fn(int(op_a), int(op_b), m, n, k, al.ctypes.data, _ptr(a), lda, _ptr(b), ldb, be.ctypes.data, _ptr(c), ldc)
        else:
            fn(int(op_a), int(op_b), m, n, k, alpha, _ptr(a), lda, _ptr(b), ldb, beta, _ptr(c), ldc)
    return f


sgemm, dgemm, cgemm, zgemm = (_mk_gemm(p) for p in "sdcz")


def _mk_gemv(p):
    def f(trans, m, n, alpha, a, lda, x, incx, beta, y, incy):
        fn = getattr(load(), f"port_{p}gemv")
        if p in "cz":
            al, be = _s2(alpha, p), _s2(beta, p)
            fn(int(trans), m, n, al.ctypes.data, _ptr(a), lda, _ptr(x), incx, be.ctypes.data, _ptr(y), incy)
        else:
            fn(int(trans), m, n, alpha, _ptr(a), lda, _ptr(x), incx, beta, _ptr(y), incy)
    return f


sgemv, dgemv, cgemv, zgemv = (_mk_gemv(p) for p in "sdcz")


def _mk_symv(p, name):
    def f(uplo, n, alpha, a, lda, x, incx, beta, y, incy):
        fn = getattr(load(), f"port_{name}")
        if p in "cz":
            al, be = _s2(alpha, p), _s2(beta, p)
            fn(int(uplo), n, al.ctypes.data, _ptr(a), lda, _ptr(x), incx, be.ctypes.data, _ptr(y), incy)
        else:
            fn(int(uplo), n, alpha, _ptr(a), lda, _ptr(x), incx, beta, _ptr(y), incy)
    return f


ssymv, dsymv = _mk_symv("s", "ssymv"), _mk_symv("d", "dsymv")
chemv, zhemv = _mk_symv("c", "chemv"), _mk_symv("z", "zhemv")


def _mk_tr(p, w):
    def f(uplo, trans, diag, n, a, lda, x, incx):
        getattr(load(), f"port_{p}{w}")(int(uplo), int(trans), int(diag), n, _ptr(a), lda, _ptr(x), incx)
    return f


strsv, dtrsv, ctrsv, ztrsv = (_mk_tr(p, "trsv") for p in "sdcz")
strmv, dtrmv, ctrmv, ztrmv = (_mk_tr(p, "trmv") for p in "sdcz")


def sdot(n, x, incx, y, incy):
    return load().port_sdot(n, _ptr(x), incx, _ptr(y), incy)


def ddot(n, x, incx, y, incy):
    return load().port_ddot(n, _ptr(x), incx, _ptr(y), incy)


def _mk_cdot(name, real):
    def f(n, x, incx, y, incy):
        out = np.zeros(2, real)
        getattr(load(), f"port_{name}")(n, _ptr(x), incx, _ptr(y), incy, out.ctypes.data)
        return [out[0], out[1]]
    return f


cdotu, cdotc = _mk_cdot("cdotu", np.float32), _mk_cdot("cdotc", np.float32)
zdotu, zdotc = _mk_cdot("zdotu", np.float64), _mk_cdot("zdotc", np.float64)


def _mk_axpy(p):
    def f(n, alpha, x, incx, y, incy):
        fn = getattr(load(), f"port_{p}axpy")
        if p in "cz":
            al = _s2(alpha, p)  # kept alive across the call
            fn(n, al.ctypes.data, _ptr(x), incx, _ptr(y), incy)
        else:
            fn(n, alpha, _ptr(x), incx, _ptr(y), incy)
    return f


saxpy, daxpy, caxpy, zaxpy = (_mk_axpy(p) for p in "sdcz")


def _mk_ger(p, name):
    def f(m, n, alpha, x, incx, y, incy, a, lda):
        keep = _s2(alpha, p) if p in "cz" else None  # kept alive across the call
        al = keep.ctypes.data if p in "cz" else alpha
        getattr(load(), f"port_{name}")(m, n, al, _ptr(x), incx, _ptr(y), incy, _ptr(a), lda)
    return f


sger, dger = _mk_ger("s", "sger"), _mk_ger("d", "dger")
cgeru, cgerc = _mk_ger("c", "cgeru"), _mk_ger("c", "cgerc")
zgeru, zgerc = _mk_ger("z", "zgeru"), _mk_ger("z", "zgerc")


def _mk_syr(name):
    def f(uplo, n, alpha, x, incx, a, lda):
        getattr(load(), f"port_{name}")(int(uplo), n, alpha, _ptr(x), incx, _ptr(a), lda)
    return f


ssyr, dsyr, cher, zher = (_mk_syr(nm) for nm in ("ssyr", "dsyr", "cher", "zher"))


def _mk_syr2(p, name):
    def f(uplo, n, alpha, x, incx, y, incy, a, lda):
        keep = _s2(alpha, p) if p in "cz" else None  # kept alive across the call
        al = keep.ctypes.data if p in "cz" else alpha
        getattr(load(), f"port_{name}")(int(uplo), n, al, _ptr(x), incx, _ptr(y), incy, _ptr(a), lda)
    return f


ssyr2, dsyr2 = _mk_syr2("s", "ssyr2"), _mk_syr2("d", "dsyr2")
cher2, zher2 = _mk_syr2("c", "cher2"), _mk_syr2("z", "zher2")


def snrm2(n, x, incx):
    return load().port_snrm2(n, _ptr(x), incx)


def dnrm2(n, x, incx):
    return load().port_dnrm2(n, _ptr(x), incx)


def scnrm2(n, x, incx):
    return load().port_scnrm2(n, _ptr(x), incx)


def dznrm2(n, x, incx):
    return load().port_dznrm2(n, _ptr(x), incx)


def isamax(n, x, incx):
    return int(load().port_isamax(n, _ptr(x), incx))


def idamax(n, x, incx):
    return int(load().port_idamax(n, _ptr(x), incx))


def icamax(n, x, incx):
    return int(load().port_icamax(n, _ptr(x), incx))


def izamax(n, x, incx):
    return int(load().port_izamax(n, _ptr(x), incx))


# ---------------------------------------------------------------- rest of level 1
def _mk_asum(name):
    def f(n, x, incx):
        return getattr(load(), f"port_{name}")(n, _ptr(x), incx)
    return f


sasum, dasum, scasum, dzasum = (_mk_asum(nm) for nm in ("sasum", "dasum", "scasum", "dzasum"))


def _mk_xy(name):
    def f(n, x, incx, y, incy):
        getattr(load(), f"port_{name}")(n, _ptr(x), incx, _ptr(y), incy)
    return f


scopy, dcopy, ccopy, zcopy = (_mk_xy(f"{p}copy") for p in "sdcz")
sswap, dswap, cswap, zswap = (_mk_xy(f"{p}swap") for p in "sdcz")


def _mk_rot(name):
    def f(n, x, incx, y, incy, c, s):
        getattr(load(), f"port_{name}")(n, _ptr(x), incx, _ptr(y), incy, c, s)
    return f


srot, drot, csrot, zdrot = (_mk_rot(nm) for nm in ("srot", "drot", "csrot", "zdrot"))


def _mk_rotm(name, real):
    def f(n, x, incx, y, incy, param):
        pa = np.ascontiguousarray(param, real)
        getattr(load(), f"port_{name}")(n, _ptr(x), incx, _ptr(y), incy, pa.ctypes.data)
    return f


srotm, drotm = _mk_rotm("srotm", np.float32), _mk_rotm("drotm", np.float64)


def _mk_rotg(name, real):
    def f(a, b):
        """Returns (r, z, c, s): CORAL's a and b on return, then c and s."""
        v = np.array([a, b, 0, 0], real)
        p = v.ctypes.data
        sz = v.itemsize
        getattr(load(), f"port_{name}")(p, p + sz, p + 2 * sz, p + 3 * sz)
        return tuple(v)
    return f


srotg, drotg = _mk_rotg("srotg", np.float32), _mk_rotg("drotg", np.float64)


def _mk_rotmg(name, real):
    def f(d1, d2, x1, y1):
        """Returns (d1, d2, x1, param[5]) after CORAL's rotmg."""
        v = np.array([d1, d2, x1], real)
        param = np.zeros(5, real)
        p, sz = v.ctypes.data, v.itemsize
        getattr(load(), f"port_{name}")(p, p + sz, p + 2 * sz, y1, param.ctypes.data)
        return v[0], v[1], v[2], param
    return f


srotmg, drotmg = _mk_rotmg("srotmg", np.float32), _mk_rotmg("drotmg", np.float64)


def _mk_scal(name, p):
    def f(n, alpha, x, incx):
        keep = _s2(alpha, p) if name in ("cscal", "zscal") else None  # kept alive across the call
        al = keep.ctypes.data if keep is not None else alpha
        getattr(load(), f"port_{name}")(n, al, _ptr(x), incx)
    return f


sscal, dscal = _mk_scal("sscal", "s"), _mk_scal("dscal", "d")
cscal, zscal = _mk_scal("cscal", "c"), _mk_scal("zscal", "z")
csscal, zdscal = _mk_scal("csscal", "c"), _mk_scal("zdscal", "z")
