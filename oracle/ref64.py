"""FP64 reference formulas for the BLAS call surface (test infrastructure).

Each function computes the mathematical result in float64 / complex128 (numpy)
from CORAL-style column-major buffers, reading only what the routine may read
(the stored triangle for SYMV/HEMV/TRSV/TRMV).  These are the "FP64 oracle"
against which error bounds are stated; ``oracle.port`` is the restatement of
CORAL's own CPU algorithms.
"""
from __future__ import annotations

import math

import numpy as np


def mat(buf, rows, cols, ld, dtype=None):
    """Column-major (rows x cols, leading dimension ld) view of a flat buffer."""
    buf = np.asarray(buf)
    out = np.empty((rows, cols), buf.dtype)
    for j in range(cols):
        out[:, j] = buf[j * ld: j * ld + rows]
    return out if dtype is None else out.astype(dtype)


def put(buf, m_, ld):
    """Write matrix m_ back into a column-major buffer copy (padding untouched)."""
    out = np.array(buf, copy=True)
    rows, cols = m_.shape
    for j in range(cols):
        out[j * ld: j * ld + rows] = m_[:, j]
    return out


def vec(buf, n, inc):
    return np.asarray(buf)[: (n - 1) * inc + 1: inc] if n > 0 else np.asarray(buf)[:0]


def _wide(a):
    return a.astype(np.complex128) if np.iscomplexobj(a) else a.astype(np.float64)


def opm(a, op):
    return a if op == 0 else (a.T if op == 1 else a.conj().T)


def gemm(op_a, op_b, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    """alpha*op(A)*op(B) + beta*C in FP64; beta == 0 ignores C (even NaN)."""
    ar, ac = (m, k) if op_a == 0 else (k, m)
    br, bc = (k, n) if op_b == 0 else (n, k)
    A = _wide(mat(a, ar, ac, lda)) if k and m else np.zeros((ar, ac))
    B = _wide(mat(b, br, bc, ldb)) if k and n else np.zeros((br, bc))
    C = _wide(mat(c, m, n, ldc))
    al = complex(*alpha) if isinstance(alpha, (list, tuple)) else alpha
    be = complex(*beta) if isinstance(beta, (list, tuple)) else beta
    prod = opm(A, op_a) @ opm(B, op_b) if k else np.zeros((m, n), C.dtype)
    res = al * prod + (0 if be == 0 else be * C)
    return put(_wide(np.asarray(c)), res, ldc)


def gemv(trans, m, n, alpha, a, lda, x, incx, beta, y, incy):
    """alpha*op(A)*x + beta*y in FP64 (beta == 0 overwrites y)."""
    A = _wide(mat(a, m, n, lda))
    xl, yl = (n, m) if trans == 0 else (m, n)
    xv = _wide(vec(x, xl, incx))
    yv = _wide(vec(y, yl, incy))
    al = complex(*alpha) if isinstance(alpha, (list, tuple)) else alpha
    be = complex(*beta) if isinstance(beta, (list, tuple)) else beta
    r = al * (opm(A, trans) @ xv) + (0 if be == 0 else be * yv)
    out = _wide(np.asarray(y)).copy()
    out[: (yl - 1) * incy + 1: incy] = r
    return out


def hermitian_from(a, n, lda, uplo, herm):
    """Full matrix from the stored triangle only (other triangle never read)."""
    A = _wide(mat(a, n, n, lda))
    keep = np.triu(np.ones((n, n), bool)) if uplo == 0 else np.tril(np.ones((n, n), bool))
    T = np.where(keep, A, 0)
    D = np.diag(np.diag(T))
    if herm:
        D = np.diag(np.diag(T).real).astype(T.dtype)
    strict = T - np.diag(np.diag(T))
    other = strict.conj().T if herm else strict.T
    return strict + other + D


def symv(uplo, n, alpha, a, lda, x, incx, beta, y, incy, herm=False):
    H = hermitian_from(a, n, lda, uplo, herm)
    xv = _wide(vec(x, n, incx))
    yv = _wide(vec(y, n, incy))
    al = complex(*alpha) if isinstance(alpha, (list, tuple)) else alpha
    be = complex(*beta) if isinstance(beta, (list, tuple)) else beta
    r = al * (H @ xv) + (0 if be == 0 else be * yv)
    out = _wide(np.asarray(y)).copy()
    out[: (n - 1) * incy + 1: incy] = r
    return out


def triangular(a, n, lda, uplo, trans, diag):
    """op(A) restricted to the stored triangle, unit diagonal substituted."""
    A = _wide(mat(a, n, n, lda))
    keep = np.triu(np.ones((n, n), bool)) if uplo == 0 else np.tril(np.ones((n, n), bool))
    T = np.where(keep, A, 0)
    if diag == 0:  # UnitDiagonal: the stored diagonal is never read
        T = np.array(T, copy=True)
        T[np.diag_indices(n)] = 1
    if trans == 2 and not np.iscomplexobj(T):
        trans = 1
    return opm(T, trans)


def trsv(uplo, trans, diag, n, a, lda, x, incx):
    import scipy.linalg

    M = triangular(a, n, lda, uplo, trans, diag)
    lower = bool(np.allclose(M, np.tril(M)))
    xv = _wide(vec(x, n, incx))
    r = scipy.linalg.solve_triangular(M, xv, lower=lower) if n else xv
    out = _wide(np.asarray(x)).copy()
    if n:
        out[: (n - 1) * incx + 1: incx] = r
    return out


def trmv(uplo, trans, diag, n, a, lda, x, incx):
    M = triangular(a, n, lda, uplo, trans, diag)
    xv = _wide(vec(x, n, incx))
    out = _wide(np.asarray(x)).copy()
    if n:
        out[: (n - 1) * incx + 1: incx] = M @ xv
    return out


def dot_exact(x, y, conj=False):
    """Correctly rounded real dot (math.fsum of exact products); complex via parts."""
    x = np.asarray(x)
    y = np.asarray(y)
    if np.iscomplexobj(x):
        xr, xi = x.real.astype(np.float64), x.imag.astype(np.float64)
        if conj:
            xi = -xi
        yr, yi = y.real.astype(np.float64), y.imag.astype(np.float64)
        re = math.fsum(list(xr * yr) + list(-xi * yi))
        im = math.fsum(list(xr * yi) + list(xi * yr))
        return complex(re, im)
    return math.fsum((x.astype(np.float64) * y.astype(np.float64)).tolist())


def nrm2(x):
    """Scaled FP64 two-norm (no overflow for any binary64 input)."""
    v = np.abs(np.asarray(x).astype(np.complex128 if np.iscomplexobj(x) else np.float64))
    if np.iscomplexobj(x):
        v = np.concatenate([np.abs(np.asarray(x).real), np.abs(np.asarray(x).imag)]).astype(np.float64)
    if v.size == 0:
        return 0.0
    if np.any(np.isnan(v)):
        return float("nan")
    mx = v.max()
    if not np.isfinite(mx):
        return float("inf")
    if mx == 0:
        return 0.0
    s = 2.0 ** -math.frexp(mx)[1]
    return float(math.sqrt(math.fsum(((v * s) ** 2).tolist())) / s)


def iamax(x, n, incx, cplx=False):
    """Brute-force first index of the largest |x_i| (|re|+|im|), NaN skipped."""
    if n == 0 or incx <= 0:
        return 0
    v = vec(x, n, incx)
    best, bi = -math.inf, 0
    for i, e in enumerate(v):
        m = (abs(e.real) + abs(e.imag)) if cplx else abs(e)
        m = np.float32(m) if np.asarray(v).dtype in (np.float32, np.complex64) else m
        if m > best:
            best, bi = m, i
    return bi
