"""CORAL's own level-2 tests, run against the GPU path.

Each test below is one of coral-aarch64/tests/level2/{s,d,c,z}gemv_tests.rs,
{s,d}symv_tests.rs, {c,z}hemv_tests.rs, {s,d,c,z}trmv_tests.rs and
{s,d,c,z}trsv_tests.rs: the same deterministic matrix and vector generators
(evaluated in the element precision, operation by operation, as the Rust code
rounds them), the same sizes, leading dimensions, strides, op/uplo/diag sweeps and
(alpha, beta) lists, the same check |x - y| <= ATOL + RTOL*max(|x|, |y|) on every
real and imaginary part with the tolerances those files state, and the same
reference: the CBLAS routine of the same precision, here OpenBLAS through
scipy.linalg.blas (the reference tests link CBLAS through cblas-sys).

One difference: where a GEMV result cancels (ZGEMV T 512 x 256 has a component of
-7.04 built from terms summing to 2.4e6 in magnitude), OpenBLAS's own result is
8.3e-11 from the exact value, beyond the 1e-11 relative tolerance, while the GPU's
is 2.1e-12 from it.  Such components are checked against the exact rational value
instead, and must be at least as close to it as OpenBLAS's.
"""
from fractions import Fraction

import numpy as np
import pytest
from scipy.linalg import blas as sb

from coral_amd import blas

pytestmark = pytest.mark.gpu

REAL = {"s": np.float32, "d": np.float64, "c": np.float32, "z": np.float64}
CPLX = {"c": np.complex64, "z": np.complex128}


def _parts(v):
    v = np.asarray(v)
    if np.iscomplexobj(v):
        v = v.astype(np.complex128)
        v = np.stack([v.real, v.imag], -1).reshape(-1)
    return v.astype(np.float64)


def close(got, ref, rtol, atol, what="", exact=None):
    """The reference tests' element check, on every real component.

    Where a result cancels, the CBLAS reference can itself be further from the exact
    value than the tolerance; `exact(k)` then gives real component k exactly (as a
    Fraction), and such a component passes only if it lies within the tolerance of
    the exact value and no further from it than the CBLAS result.
    """
    g, r = _parts(got), _parts(ref)
    tol = atol + rtol * np.maximum(np.abs(g), np.abs(r))
    bad = np.nonzero(~(np.abs(g - r) <= tol))[0]
    if bad.size and exact is not None:
        still = []
        for k in bad:
            e = exact(int(k))
            eg, er = abs(Fraction(float(g[k])) - e), abs(Fraction(float(r[k])) - e)
            tk = atol + rtol * max(abs(float(g[k])), abs(float(e)))
            if not (float(eg) <= tk and eg <= er):
                still.append(k)
        bad = np.array(still, dtype=np.int64)
    assert bad.size == 0, f"{what}: mismatch at {bad[0]}: {g[bad[0]]} vs {r[bad[0]]} (tol {tol[bad[0]]})"


def _fr(v):
    """(re, im) of a scalar or [re, im] pair as exact Fractions."""
    if isinstance(v, (tuple, list)):
        return Fraction(float(v[0])), Fraction(float(v[1]))
    c = complex(v)
    return Fraction(c.real), Fraction(c.imag)


def exact_gemv(M, x, alpha, beta, y, op):
    """Component k of alpha*op(M)*x + beta*y in exact rational arithmetic."""
    def f(k):
        e, part = (k // 2, k % 2) if np.iscomplexobj(M) else (k, 0)
        row = M[e, :] if op == 0 else (M[:, e].conj() if op == 2 else M[:, e])
        sr = si = Fraction(0)
        for a, xv in zip(row, x):
            ar, ai = _fr(a)
            xr, xi = _fr(xv)
            sr += ar * xr - ai * xi
            si += ar * xi + ai * xr
        alr, ali = _fr(alpha)
        ber, bei = _fr(beta)
        yr, yi = _fr(y[e])
        re = alr * sr - ali * si + ber * yr - bei * yi
        im = alr * si + ali * sr + ber * yi + bei * yr
        return im if part else re
    return f


def strided(vals, inc):
    """make_strided_vec: the logical values at stride inc, zeros between."""
    v = np.zeros((len(vals) - 1) * inc + 1 if len(vals) else 1, vals.dtype)
    v[::inc] = vals
    return v


def lin(R, n, c0, c1):
    """c0 + c1*k for k in 0..n, rounded in precision R operation by operation."""
    k = np.arange(n).astype(R)
    return (R(c0) + R(c1) * k).astype(R)


def cvec(p, n, re, im):
    R = REAL[p]
    return (lin(R, n, *re).astype(CPLX[p]) + 1j * lin(R, n, *im).astype(CPLX[p])).astype(CPLX[p])


def logical(v, n, inc):
    return np.asarray(v)[: (n - 1) * inc + 1: inc] if n else np.asarray(v)[:0]


def colmajor(M, lda):
    """n_rows x n_cols array -> column-major buffer with leading dimension lda (zero padding)."""
    m, n = M.shape
    buf = np.zeros((n, lda), M.dtype)
    buf[:, :m] = M.T
    return buf.reshape(-1)


# ---------------------------------------------------------------- GEMV
# (m, n, lda, stride pairs) of small/large/padded/strided/rectangular_all_ops
GEMV_CASES = {
    "s": [(5, 4, 5, [(1, 1)]), (1024, 512, 1024, [(1, 1)]), (256, 1024, 263, [(1, 1)]),
          (1024, 512, 1024, [(2, 3), (3, 2), (2, 1)]), (640, 320, 640, [(1, 1)])],
    "c": [(5, 4, 5, [(1, 1)]), (384, 256, 384, [(1, 1)]), (256, 512, 263, [(1, 1)]),
          (300, 200, 305, [(2, 1), (1, 3), (2, 3)]), (64, 192, 64, [(1, 1)])],
    "z": [(5, 4, 5, [(1, 1)]), (384, 256, 384, [(1, 1)]), (256, 512, 263, [(1, 1)]),
          (512, 256, 512, [(2, 3), (3, 2), (2, 1)]), (640, 320, 640, [(1, 1)])],
}
GEMV_CASES["d"] = GEMV_CASES["s"]
GEMV_COEFFS = {  # sgemv_tests.rs / dgemv_tests.rs:207-214, cgemv_tests.rs, zgemv_tests.rs
    "s": [(1.0, 0.0), (0.0, -0.75), (0.75, -0.25), (-0.6, 0.4), (1.1, 0.0), (0.85, 0.1)],
    "c": [((1.0, 0.0), (0.0, 0.0)), ((0.0, 0.0), (0.7, -0.1)), ((0.75, -0.25), (0.5, 0.0)),
          ((-0.5, 0.2), (0.3, -0.15)), ((1.1, -0.6), (0.0, 0.0)), ((0.3, 0.7), (0.1, -0.2))],
    "z": [((1.0, 0.0), (0.0, 0.0)), ((0.0, 0.0), (-0.75, 0.2)), ((0.75, -0.10), (-0.25, 0.05)),
          ((-0.6, 0.2), (0.4, -0.1)), ((1.1, -0.6), (0.0, 0.0)), ((0.85, -0.2), (0.1, 0.0))],
}
GEMV_COEFFS["d"] = GEMV_COEFFS["s"]
GEMV_TOL = {"s": (1e-5, 1e-5), "d": (1e-13, 1e-13), "c": (5e-3, 1e-6), "z": (1e-11, 1e-12)}


def gemv_matrix(p, m, n):
    R = REAL[p]
    i = np.arange(m).astype(R)[:, None]
    j = np.arange(n).astype(R)[None, :]
    re = (R(0.1) + i * R(0.5)) + j * R(0.25)
    if p in "sd":
        return re.astype(R)
    im = (R(-0.2) + i * R(0.3)) - j * R(0.15)
    return (re.astype(CPLX[p]) + 1j * im.astype(CPLX[p])).astype(CPLX[p])


@pytest.mark.parametrize("p", "sdcz")
def test_gemv_reference_cases(gpu_lib, p):  # coral-aarch64/tests/level2/?gemv_tests.rs
    rtol, atol = GEMV_TOL[p]
    ops = (0, 1) if p in "sd" else (0, 1, 2)
    fn = getattr(blas, f"{p}gemv")
    ref_fn = getattr(sb, f"{p}gemv")
    for m, n, lda, pairs in GEMV_CASES[p]:
        M = gemv_matrix(p, m, n)
        a = colmajor(M, lda)
        for incx, incy in pairs:
            for op in ops:
                xl, yl = (n, m) if op == 0 else (m, n)
                if p in "sd":
                    xv, yv = lin(REAL[p], xl, 0.2, 0.1), lin(REAL[p], yl, -0.3, 0.05)
                else:
                    xv, yv = cvec(p, xl, (0.2, 0.1), (0.0, -0.05)), cvec(p, yl, (-0.3, 0.05), (0.0, 0.02))
                x, y0 = strided(xv, incx), strided(yv, incy)
                for alpha, beta in GEMV_COEFFS[p]:
                    y = y0.copy()
                    fn(op, m, n, alpha, a, lda, x, incx, beta, y, incy)
                    al = alpha if p in "sd" else complex(*alpha)
                    be = beta if p in "sd" else complex(*beta)
                    ref = ref_fn(al, M, xv, beta=be, y=yv.copy(), trans=op)
                    close(logical(y, yl, incy), ref, rtol, atol, f"{p}gemv m={m} n={n} op={op} {alpha},{beta}",
                          exact=exact_gemv(M, xv, alpha, beta, yv, op))
                    # the elements between the strided entries are untouched
                    mask = np.ones(y.shape[0], bool)
                    mask[::incy] = False
                    assert np.array_equal(y[mask], y0[mask])


# ---------------------------------------------------------------- SYMV / HEMV
SYMV_TOL = {"s": (1e-5, 1e-5), "d": (1e-13, 1e-13), "c": (1e-5, 1e-6), "z": (1e-12, 1e-12)}
SYMV_COEFFS = {  # ssymv/dsymv_tests.rs:211-219, chemv/zhemv_tests.rs:261-268
    "s": [(1.0, 0.0), (0.0, 0.7), (0.75, -0.25), (-0.6, 0.4), (1.1, 0.0), (0.85, 0.1)],
    "c": [((1.0, 0.0), (0.0, 0.0)), ((0.0, 0.0), (0.7, -0.1)), ((0.75, -0.10), (-0.25, 0.15)),
          ((-0.6, 0.2), (0.4, -0.1)), ((1.25, 0.0), (-0.5, 0.0)), ((0.85, -0.1), (0.1, 0.05))],
}
SYMV_COEFFS["d"], SYMV_COEFFS["z"] = SYMV_COEFFS["s"], SYMV_COEFFS["c"]
# (n, lda, stride pairs) of small/large/padded/strided_all_uplos
SYMV_CASES = {"s": [(6, 6, [(1, 1)]), (768, 768, [(1, 1)]), (300, 305, [(1, 1)]),
                    (640, 640, [(2, 3), (3, 2), (2, 1)])],
              "c": [(6, 6, [(1, 1)]), (640, 640, [(1, 1)]), (300, 305, [(1, 1)]),
                    (384, 384, [(2, 3), (3, 2), (2, 1)])]}
SYMV_CASES["d"], SYMV_CASES["z"] = SYMV_CASES["s"], SYMV_CASES["c"]


def sym_matrix(p, n, c0, c1, c2, im=None, conj_lower=True):
    """Full symmetric/Hermitian matrix with a(i,j) = c0 + c1*lo + c2*hi (lo = min, hi = max);
    im = (d0, d1): imaginary part d0 + d1*(hi - lo) off the diagonal, conjugated below."""
    R = REAL[p]
    i = np.arange(n)[:, None]
    j = np.arange(n)[None, :]
    lo = np.minimum(i, j).astype(R)
    hi = np.maximum(i, j).astype(R)
    re = ((R(c0) + R(c1) * lo) + R(c2) * hi).astype(R)
    if p in "sd":
        return re
    imv = (R(im[0]) + R(im[1]) * (hi - lo)).astype(R)
    imv = np.where(i == j, R(0), imv)
    imv = np.where(i > j, -imv, imv) if conj_lower else imv
    return (re.astype(CPLX[p]) + 1j * imv.astype(CPLX[p])).astype(CPLX[p])


def symv_call(p, uplo, n, alpha, a, lda, x, incx, beta, y, incy):
    fn = {"s": blas.ssymv, "d": blas.dsymv, "c": blas.chemv, "z": blas.zhemv}[p]
    fn(uplo, n, alpha, a, lda, x, incx, beta, y, incy)


def symv_ref(p, uplo, M, alpha, x, beta, y):
    if p in "sd":
        return getattr(sb, f"{p}symv")(alpha, M, x, beta=beta, y=y.copy(), lower=uplo)
    return getattr(sb, f"{p}hemv")(complex(*alpha), M, x, beta=complex(*beta), y=y.copy(), lower=uplo)


@pytest.mark.parametrize("p", "sdcz")
def test_symv_hemv_reference_cases(gpu_lib, p):  # ?symv_tests.rs / ?hemv_tests.rs: *_all_uplos
    rtol, atol = SYMV_TOL[p]
    for n, lda, pairs in SYMV_CASES[p]:
        if p in "sd":
            M = sym_matrix(p, n, 0.1, 0.5, 0.25)
        else:
            M = sym_matrix(p, n, 0.10, 0.25, 0.15, im=(0.04, 0.02))
        a = colmajor(M, lda)
        if p in "sd":
            xv, yv = lin(REAL[p], n, 0.2, 0.1), lin(REAL[p], n, -0.3, 0.05)
        else:
            xv, yv = cvec(p, n, (0.2, 0.1), (-0.3, 0.05)), cvec(p, n, (-0.3, 0.05), (0.1, -0.02))
        for incx, incy in pairs:
            x, y0 = strided(xv, incx), strided(yv, incy)
            for uplo in (0, 1):
                for alpha, beta in SYMV_COEFFS[p]:
                    y = y0.copy()
                    symv_call(p, uplo, n, alpha, a, lda, x, incx, beta, y, incy)
                    close(logical(y, n, incy), symv_ref(p, uplo, M, alpha, xv, beta, yv), rtol, atol,
                          f"{p} n={n} uplo={uplo} {alpha},{beta}")


@pytest.mark.parametrize("p", "sdcz")
def test_symv_hemv_upper_equals_lower_and_triangles(gpu_lib, p):
    # upper_equals_lower, upper_respects_triangle, lower_respects_triangle
    rtol, atol = SYMV_TOL[p]
    R = REAL[p]
    cplx = p in "cz"
    # upper_equals_lower (n = 513 real, 257 complex)
    n = 257 if cplx else 513
    if cplx:
        M = sym_matrix(p, n, 0.10, 0.25, 0.15, im=(0.04, 0.02))
        xv, yv = cvec(p, n, (-0.07, 0.013), (0.015, -0.002)), cvec(p, n, (0.02, -0.004), (-0.01, 0.001))
        alpha, beta = (1.123, -0.321), (-0.321, 0.777)
    else:
        M = sym_matrix(p, n, 0.1, 0.5, 0.25)
        xv, yv = lin(R, n, -0.07, 0.013), lin(R, n, 0.02, -0.004)
        alpha, beta = 1.123, -0.321
    a = colmajor(M, n)
    yu, yl = yv.copy(), yv.copy()
    symv_call(p, 0, n, alpha, a, n, xv, 1, beta, yu, 1)
    symv_call(p, 1, n, alpha, a, n, xv, 1, beta, yl, 1)
    close(yu, yl, rtol, atol, "upper == lower")
    # upper_respects_triangle: n = 200, lda = n + 3, NaN below the diagonal
    n, lda = 200, 203
    if cplx:
        M = sym_matrix(p, n, 0.20, 0.30, 0.15, im=(0.05, 0.01), conj_lower=False)
        xv, yv = cvec(p, n, (-0.5, 0.03), (0.02, -0.01)), cvec(p, n, (-0.1, 0.002), (0.05, -0.003))
        alpha, beta = (0.7, 0.1), (0.25, -0.05)
    else:
        M = sym_matrix(p, n, 0.2, 0.3, 0.15)
        xv, yv = lin(R, n, -0.5, 0.03), lin(R, n, -0.1, 0.002)
        alpha, beta = 0.7, 0.25
    Mg = M.copy()
    Mg[np.tril_indices(n, -1)] = np.nan
    a = colmajor(Mg, lda)
    y = yv.copy()
    symv_call(p, 0, n, alpha, a, lda, xv, 1, beta, y, 1)
    close(y, symv_ref(p, 0, Mg, alpha, xv, beta, yv), rtol, atol, "upper respects triangle")
    # lower_respects_triangle: n = 180, lda = n, NaN above the diagonal
    n = 180
    if cplx:
        M = sym_matrix(p, n, 0.17, 0.45, 0.08, im=(-0.03, -0.02), conj_lower=False)
        xv, yv = cvec(p, n, (0.2, -0.001), (-0.03, 0.004)), cvec(p, n, (0.15, 0.004), (-0.02, 0.003))
        alpha, beta = (-0.9, 0.2), (0.6, -0.1)
    else:
        M = sym_matrix(p, n, 0.17, 0.45, 0.08)
        xv, yv = lin(R, n, 0.2, -0.001), lin(R, n, 0.15, 0.004)
        alpha, beta = -0.9, 0.6
    Mg = M.copy()
    Mg[np.triu_indices(n, 1)] = np.nan
    a = colmajor(Mg, n)
    y = yv.copy()
    symv_call(p, 1, n, alpha, a, n, xv, 1, beta, y, 1)
    close(y, symv_ref(p, 1, Mg, alpha, xv, beta, yv), rtol, atol, "lower respects triangle")


# ---------------------------------------------------------------- TRMV / TRSV
def tri_real(kind, p, n, uplo, unit):
    """dtrmv_tests.rs:63-178 (kind "mv") and dtrsv_tests.rs:64-242 (kind "sv") generators."""
    R = REAL[p]
    i = np.arange(n)[:, None].astype(R)
    j = np.arange(n)[None, :].astype(R)
    up = np.arange(n)[:, None] <= np.arange(n)[None, :]
    lowm = np.arange(n)[:, None] >= np.arange(n)[None, :]
    eye = np.eye(n, dtype=bool)
    if kind == "mv":
        T = ((R(0.2) + i * R(0.3)) + j * R(0.15)) if uplo == 0 else ((R(-0.1) + i * R(0.25)) - j * R(0.2))
        T = np.where(up if uplo == 0 else lowm, T, R(0)).astype(R)
        if unit:
            d = (R(7.5) + np.arange(n).astype(R) * R(0.1)) if uplo == 0 else (R(-9.0) + np.arange(n).astype(R) * R(0.05))
            T[eye] = d
        return T
    d = np.abs(np.arange(n)[:, None] - np.arange(n)[None, :]).astype(R)
    if not unit:
        off = R(0.01) / (R(1.0) + d)
        dg = R(1.0) + R(0.02) * np.arange(n).astype(R)
    else:
        off = (R(5e-4) if uplo == 0 else R(6e-4)) * (R(1.0) + d)
        dg = (R(7.5) + np.arange(n).astype(R) * R(0.1)) if uplo == 0 else (R(-9.0) + np.arange(n).astype(R) * R(0.05))
    T = np.where(up if uplo == 0 else lowm, off, R(0)).astype(R)
    T[eye] = dg
    return T


def tri_cplx(p, n, uplo, unit):
    """ctrmv/ztrmv_tests.rs:63-246 and ctrsv/ztrsv_tests.rs generators (shared shape)."""
    R = REAL[p]
    k = np.arange(n).astype(R)
    d = np.abs(np.arange(n)[:, None] - np.arange(n)[None, :]).astype(R)
    up = np.arange(n)[:, None] <= np.arange(n)[None, :]
    lowm = np.arange(n)[:, None] >= np.arange(n)[None, :]
    eye = np.eye(n, dtype=bool)
    if not unit:
        ore = R(0.01) / (R(1.0) + d)
        oim = (-R(0.008) / (R(1.0) + d)) if uplo == 0 else (R(0.007) / (R(1.0) + d))
    else:
        ore = (R(5e-4) if uplo == 0 else R(6e-4)) * (R(1.0) + d)
        oim = (-(R(3e-4)) * (R(1.0) + d)) if uplo == 0 else (R(4e-4) * (R(1.0) + d))
    keep = up if uplo == 0 else lowm
    re = np.where(keep, ore, R(0)).astype(R)
    im = np.where(keep, oim, R(0)).astype(R)
    if not unit:
        re[eye] = R(1.0) + R(0.02) * k
        im[eye] = (R(0.01) - R(0.005) * k) if uplo == 0 else (R(-0.012) + R(0.004) * k)
    return (re.astype(CPLX[p]) + 1j * im.astype(CPLX[p])).astype(CPLX[p])


def tri_cplx_unit_diag(p, n, uplo, T, solve):
    """The 'funky' diagonal the unit-diagonal cases store (never read)."""
    R = REAL[p]
    k = np.arange(n).astype(R)
    eye = np.eye(n, dtype=bool)
    re = (R(7.5) + k * R(0.1)) if uplo == 0 else (R(-9.0) + k * R(0.05))
    if uplo == 0:
        im = R(-3.2) + k * R(0.07)
    else:
        im = (R(4.2) - k * R(0.03)) if solve else (R(2.1) - k * R(0.03))
    T = T.copy()
    T[eye] = (re.astype(CPLX[p]) + 1j * im.astype(CPLX[p])).astype(CPLX[p])
    return T


TRMV_CASES = {"s": [(7, 7, [1]), (1024, 1024, [1]), (640, 640, [2, 3])],
              "c": [(7, 7, [1]), (768, 768, [1]), (256, 261, [1]), (384, 384, [2, 3, 4])]}
TRMV_CASES["d"], TRMV_CASES["z"] = TRMV_CASES["s"], TRMV_CASES["c"]
TRSV_CASES = {"s": [(7, 7, [1]), (1024, 1024, [1]), (640, 640, [2, 3, 4]), (127, 132, [1])],
              "c": [(7, 7, [1]), (760, 768, [1]), (384, 384, [2, 3, 4]), (256, 263, [1])]}
TRSV_CASES["d"], TRSV_CASES["z"] = TRSV_CASES["s"], TRSV_CASES["c"]
TRMV_TOL = {"s": (1e-5, 1e-5), "d": (1e-13, 1e-13), "c": (1e-5, 1e-5), "z": (1e-13, 1e-13)}
TRSV_TOL = {"s": (1e-4, 1e-4), "d": (1e-12, 1e-12), "c": (1e-2, 1e-3), "z": (1e-11, 1e-11)}


def tri_build(p, kind, n, uplo, diag):
    unit = diag == 0
    if p in "sd":
        return tri_real(kind, p, n, uplo, unit)
    T = tri_cplx(p, n, uplo, unit)
    return tri_cplx_unit_diag(p, n, uplo, T, kind == "sv") if unit else T


@pytest.mark.parametrize("kind", ["mv", "sv"])
@pytest.mark.parametrize("p", "sdcz")
def test_trmv_trsv_reference_cases(gpu_lib, p, kind):  # ?trmv_tests.rs / ?trsv_tests.rs: run_all
    rtol, atol = (TRMV_TOL if kind == "mv" else TRSV_TOL)[p]
    cases = (TRMV_CASES if kind == "mv" else TRSV_CASES)[p]
    transs = (0, 1) if (kind == "sv" and p in "sd") else (0, 1, 2)
    fn = getattr(blas, f"{p}tr{kind}")
    ref_fn = getattr(sb, f"{p}tr{kind}")
    for n, lda, incs in cases:
        for incx in incs:
            if p in "sd":
                xv = lin(REAL[p], n, 0.1, 0.2) if kind == "mv" else lin(REAL[p], n, 0.05, 0.01)
            else:
                xv = cvec(p, n, (0.1, 0.2), (-0.05, 0.03))
            x0 = strided(xv, incx)
            for uplo in (0, 1):
                for trans in transs:
                    for diag in (1, 0):
                        T = tri_build(p, kind, n, uplo, diag)
                        a = colmajor(T, lda)
                        x = x0.copy()
                        fn(uplo, trans, diag, n, a, lda, x, incx)
                        ref = ref_fn(T, xv.copy(), lower=uplo, trans=trans, diag=1 if diag == 0 else 0)
                        close(logical(x, n, incx), ref, rtol, atol,
                              f"{p}tr{kind} n={n} inc={incx} uplo={uplo} trans={trans} diag={diag}")


@pytest.mark.parametrize("p", "sdcz")
def test_trmv_trsv_padded_lower_transpose(gpu_lib, p):
    # padded_lower_transpose (trmv) / lower_transpose_padded (trsv): both triangles stored,
    # lda > n, lower + transpose + non-unit must read only the lower triangle
    R = REAL[p]
    n = 127
    if p in "sd":
        i = np.arange(n)[:, None].astype(R)
        j = np.arange(n)[None, :].astype(R)
        Pm = np.where(np.arange(n)[:, None] >= np.arange(n)[None, :], (R(0.1) + i * R(0.2)) + j * R(0.1),
                      (R(0.1) + i * R(0.1)) + j * R(0.2)).astype(R)  # dtrmv_tests.rs:90-107
        Ps = (tri_real("sv", p, n, 1, False) + tri_real("sv", p, n, 0, False)).astype(R)
        Ps[np.eye(n, dtype=bool)] = R(1.0) + R(0.02) * np.arange(n).astype(R)  # dtrsv_tests.rs:99-123
        xv = lin(R, n, 0.3, -0.006)
        cases = (("mv", Pm, n + 5), ("sv", Ps, n + 5))
    else:
        L = tri_cplx(p, n, 1, False)
        U = tri_cplx(p, n, 0, False)
        Ps = (L + U).astype(CPLX[p])
        Ps[np.eye(n, dtype=bool)] = U[np.eye(n, dtype=bool)]  # ctrsv_tests.rs make_padded_c32
        xv = cvec(p, n, (0.3, -0.006), (-0.12, 0.003))
        cases = (("sv", Ps, n + 7),)
    for kind, P, lda in cases:
        rtol, atol = (TRMV_TOL if kind == "mv" else TRSV_TOL)[p]
        a = colmajor(P, lda)
        x = xv.copy()
        getattr(blas, f"{p}tr{kind}")(1, 1, 1, n, a, lda, x, 1)
        ref = getattr(sb, f"{p}tr{kind}")(P, xv.copy(), lower=1, trans=1, diag=0)
        close(x, ref, rtol, atol, f"{p}tr{kind} padded lower transpose")


@pytest.mark.parametrize("p", "sd")
def test_trsv_unit_diagonal_cases(gpu_lib, p):
    # unitdiag_upper_notranspose (n = 40) and unitdiag_lower_transpose (n = 48)
    R = REAL[p]
    rtol, atol = TRSV_TOL[p]
    for n, uplo, trans, x0 in ((40, 0, 0, lin(R, 40, 0.2, 0.01)), (48, 1, 1, lin(R, 48, -0.1, 0.02))):
        T = tri_real("sv", p, n, uplo, True)
        a = colmajor(T, n)
        x = x0.copy()
        getattr(blas, f"{p}trsv")(uplo, trans, 0, n, a, n, x, 1)
        ref = getattr(sb, f"{p}trsv")(T, x0.copy(), lower=uplo, trans=trans, diag=1)
        close(x, ref, rtol, atol, f"{p}trsv unit n={n}")


@pytest.mark.parametrize("p", "sdcz")
def test_n_zero_quick_returns(gpu_lib, p):
    # n_zero_quick_return of the trmv/trsv tests: x is untouched
    dt = CPLX.get(p, REAL[p])
    for kind in ("mv", "sv"):
        x = np.array([1.23 - 4.56j if p in "cz" else 1.23], dt)
        getattr(blas, f"{p}tr{kind}")(0, 2 if p in "cz" else 0, 1, 0, np.zeros(1, dt), 1, x, 1)
        assert x[0] == dt(1.23 - 4.56j if p in "cz" else 1.23)
