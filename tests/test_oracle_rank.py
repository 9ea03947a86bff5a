"""Pins the oracle's level-2 rank updates and axpys (oracle/coral_port.c) before
they serve as the checker of the GPU kernels.  CPU only.

* The reference tests' own cases (coral-aarch64/tests/level2/{sger,dger,cger,zger,
  ssyr,dsyr,ssyr2,dsyr2,cher,zher,cher2,zher2}_tests.rs): their matrix and vector
  generators, sizes, strides, padding, alphas and tolerances, checked against
  OpenBLAS through scipy.linalg.blas, the cblas those tests call.
* Exact arithmetic: with dyadic inputs every product and sum is exact, so the
  result must equal the textbook float64 formula bit for bit.  This pins the
  indices, the conjugates, the stored triangle, the diagonal rule of HER/HER2 and
  CORAL's skip of zero coefficients (visible on -0.0 and non-finite entries).
* Rounding: inputs for which a fused multiply-add gives 2^-26 (f32) or 2^-56
  (f64) and a multiply-then-add gives 0.  The result shows which elements CORAL's
  NEON axpy body fuses (the first floor(len/Q)*Q elements of a unit-stride
  segment: Q = 8 saxpy, 4 daxpy/caxpy, 2 zaxpy) and which it rounds twice (the
  tail, the strided axpy, the strided complex GER loop), and that the strided
  real GER/SYR/SYR2 loops fuse everywhere (`mul_add`).
"""
from fractions import Fraction

import numpy as np
import pytest
import scipy.linalg.blas as sblas

from oracle import port
from tests.conftest import allclose_ref

REAL = {"s": np.float32, "d": np.float64, "c": np.float32, "z": np.float64}
CPLX = {"c": np.complex64, "z": np.complex128}
DT = {"s": np.float32, "d": np.float64, "c": np.complex64, "z": np.complex128}

# (rtol, atol) of each reference test file
TOL = {"sger": (1e-6, 1e-6), "dger": (1e-13, 1e-13), "cgeru": (1e-6, 1e-6), "cgerc": (1e-6, 1e-6),
       "zgeru": (1e-12, 1e-12), "zgerc": (1e-12, 1e-12), "ssyr": (1e-6, 1e-5), "dsyr": (1e-6, 1e-5),
       "ssyr2": (1e-6, 1e-5), "dsyr2": (1e-6, 1e-5), "cher": (1e-5, 1e-6), "zher": (1e-13, 1e-13),
       "cher2": (1e-5, 1e-6), "zher2": (1e-13, 1e-13)}
GER = ["sger", "dger", "cgeru", "cgerc", "zgeru", "zgerc"]
SYR = ["ssyr", "dsyr", "cher", "zher"]
SYR2 = ["ssyr2", "dsyr2", "cher2", "zher2"]


def strided(vals, inc, dt):
    """make_strided_vec of the reference tests: logical element k at k*inc, zeros between."""
    n = len(vals)
    v = np.zeros((n - 1) * inc + 1 if n else 0, dt)
    v[::inc] = vals
    return v


def gen_ger_matrix(m, n, lda, dt):
    """sger_tests.rs:33-48: a[i + j*lda] = 0.1 + 0.5 i + 0.25 j, zero padding (an
    imaginary part 0.05 - 0.01 i + 0.02 j for the complex cases)."""
    a = np.zeros(lda * n, dt)
    i, j = np.meshgrid(np.arange(m), np.arange(n), indexing="ij")
    val = 0.1 + 0.5 * i + 0.25 * j
    if np.issubdtype(dt, np.complexfloating):
        val = val + 1j * (0.05 - 0.01 * i + 0.02 * j)
    a.reshape(n, lda)[:, :m] = val.T
    return a


def gen_sym_matrix(n, lda, dt, herm):
    """ssyr_tests.rs:55-70 (0.1 + 0.5 lo + 0.25 hi, symmetric) and
    cher2_tests.rs:33-55 (re 0.10 + 0.25 lo + 0.15 hi, im 0.04 + 0.02 (hi-lo) off
    the diagonal, Hermitian)."""
    a = np.zeros(lda * n, dt)
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    lo, hi = np.minimum(i, j), np.maximum(i, j)
    if herm:
        im = np.where(i == j, 0.0, 0.04 + 0.02 * (hi - lo))
        val = (0.10 + 0.25 * lo + 0.15 * hi) + 1j * np.where(i <= j, im, -im)
    else:
        val = 0.1 + 0.5 * lo + 0.25 * hi
    a.reshape(n, lda)[:, :n] = val.T
    return a


def logical(a, m, n, lda):
    return a.reshape(n, lda)[:, :m].T


def vec_gen(kind, cplx):
    """The reference tests' vector generators (linear in the index)."""
    gens = {
        "x_small": (0.2, 0.1, -0.05, 0.02), "y_small": (-0.3, 0.05, 0.07, -0.01),
        "x_large": (0.05, 0.002, -0.02, 0.0007), "y_large": (0.40, -0.003, 0.01, -0.0005),
        "x_pad": (0.05, 0.03, -0.02, 0.01), "y_pad": (0.40, -0.02, 0.03, -0.01),
    }
    r0, r1, i0, i1 = gens[kind]
    return lambda k: (r0 + r1 * k) + (1j * (i0 + i1 * k) if cplx else 0)


# (name-family, uplo, m, n, lda, incx, incy, alpha, gen suffix): the reference tests'
# run_contiguous_sets / run_small_and_large / run_strided_padded / lower_padded_strided
GER_CASES = [(7, 5, 7, 1, 1, 1.25, "small"), (1024, 768, 1024, 1, 1, -0.37, "large"),
             (512, 1024, 512, 1, 1, 0.93, "large"), (9, 4, 12, 2, 3, -0.85, "pad")]
TRI_CASES = [(0, 7, 7, 1, 1, 1.25, "small"), (1, 7, 7, 1, 1, -0.8, "small"),
             (0, 1024, 1024, 1, 1, -0.37, "large"), (1, 512, 512, 1, 1, 0.93, "large"),
             (1, 9, 12, 2, 3, -0.85, "pad")]
CALPHA = {1.25: (1.10, -0.35), -0.8: (-0.80, 0.50), -0.37: (-0.37, 0.20), 0.93: (0.93, -0.10),
          -0.85: (-0.85, 0.25)}


def _vals(kind, cplx, n):
    g = vec_gen(kind, cplx)
    return np.array([g(k) for k in range(n)])


@pytest.mark.parametrize("name", GER)
def test_ger_reference_cases_vs_blas(name):
    p = name[0]
    cplx = p in "cz"
    dt = DT[p]
    for m, n, lda, incx, incy, alpha, g in GER_CASES:
        al = CALPHA[alpha] if cplx else alpha
        xv, yv = _vals("x_" + g, cplx, m).astype(dt), _vals("y_" + g, cplx, n).astype(dt)
        x, y = strided(xv, incx, dt), strided(yv, incy, dt)
        a0 = gen_ger_matrix(m, n, lda, dt)
        a = a0.copy()
        getattr(port, name)(m, n, al, x, incx, y, incy, a, lda)
        blas_name = name if not cplx else name
        want = getattr(sblas, blas_name)(complex(*al) if cplx else al, xv, yv,
                                         a=np.asfortranarray(logical(a0, m, n, lda)))
        rtol, atol = TOL[name]
        assert allclose_ref(logical(a, m, n, lda), want, rtol, atol), (name, m, n)
        if lda > m:  # padding rows untouched
            assert np.array_equal(a.reshape(n, lda)[:, m:], a0.reshape(n, lda)[:, m:])


@pytest.mark.parametrize("name", SYR + SYR2)
def test_syr_her_reference_cases_vs_blas(name):
    p = name[0]
    cplx = p in "cz"
    two = name in SYR2
    dt = DT[p]
    for uplo, n, lda, incx, incy, alpha, g in TRI_CASES:
        al = (CALPHA[alpha] if (cplx and two) else alpha)
        xv = _vals("x_" + g, cplx, n).astype(dt)
        yv = _vals("y_" + g, cplx, n).astype(dt)
        x, y = strided(xv, incx, dt), strided(yv, incy, dt)
        a0 = gen_sym_matrix(n, lda, dt, herm=cplx)
        a = a0.copy()
        if two:
            getattr(port, name)(uplo, n, al, x, incx, y, incy, a, lda)
            want = getattr(sblas, name)(complex(*al) if cplx else al, xv, yv, lower=uplo,
                                        a=np.asfortranarray(logical(a0, n, n, lda)))
        else:
            getattr(port, name)(uplo, n, al, x, incx, a, lda)
            want = getattr(sblas, name)(al, xv, lower=uplo, a=np.asfortranarray(logical(a0, n, n, lda)))
        got = logical(a, n, n, lda)
        rtol, atol = TOL[name]
        assert allclose_ref(got, want, rtol, atol), (name, uplo, n)
        # only the stored triangle is touched (assert_only_triangle_touched of the tests)
        i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
        other = (i > j) if uplo == 0 else (i < j)
        assert np.array_equal(got[other], logical(a0, n, n, lda)[other]), (name, uplo)


@pytest.mark.parametrize("name", GER + SYR + SYR2)
def test_alpha_zero_and_empty_are_no_ops(name):
    p = name[0]
    cplx = p in "cz"
    dt = DT[p]
    n = 300
    a0 = gen_sym_matrix(n, n + 5, dt, herm=cplx)
    x = _vals("x_small", cplx, n).astype(dt)
    zero = (0.0, 0.0) if cplx and name not in SYR else 0.0
    for nn, al in ((n, zero), (0, (1.0, 0.5) if cplx and name not in SYR else 1.0)):
        a = a0.copy()
        if name in GER:
            getattr(port, name)(nn, n, al, x, 1, x, 1, a, n + 5)
            getattr(port, name)(n, nn, al, x, 1, x, 1, a, n + 5)
        elif name in SYR:
            getattr(port, name)(0, nn, al, x, 1, a, n + 5)
        else:
            getattr(port, name)(0, nn, al, x, 1, x, 1, a, n + 5)
        assert np.array_equal(a.view(np.uint8), a0.view(np.uint8)), name


def _dyadic(rng, shape, cplx, lim=8):
    v = rng.integers(-lim * 8, lim * 8 + 1, shape) / 8.0
    if cplx:
        v = v + 1j * (rng.integers(-lim * 8, lim * 8 + 1, shape) / 8.0)
    return v


@pytest.mark.parametrize("name", GER + SYR + SYR2)
@pytest.mark.parametrize("strides", [(1, 1), (2, 3)])
def test_exact_dyadic_against_formula(name, strides):
    """Dyadic inputs: CORAL's result is the exact value of the BLAS definition."""
    p = name[0]
    cplx = p in "cz"
    dt = DT[p]
    incx, incy = strides
    rng = np.random.default_rng(hash(name) % 1000 + incx)
    m, n = (13, 11) if name in GER else (12, 12)
    lda = m + 2
    xv = _dyadic(rng, m if name in GER else n, cplx)
    yv = _dyadic(rng, n, cplx)
    xv[3] = 0  # zero coefficient columns (alpha*x_3 == 0 for SYR/HER, ...)
    yv[2] = 0  # ... alpha*y_2 == 0 for GER/SYR2/HER2
    a0v = _dyadic(rng, (m, n), cplx)
    if name in SYR2:  # both coefficients of column 5 are zero
        xv[5] = yv[5] = 0
    zero_col = 2 if name in GER else (3 if name in SYR else 5)
    a0v[:, zero_col] = complex(-0.0, -0.0) if cplx else -0.0  # a skipped column keeps its -0.0
    a0 = np.zeros(lda * n, dt)
    a0.reshape(n, lda)[:, :m] = a0v.T.astype(dt)
    x, y = strided(xv.astype(dt), incx, dt), strided(yv.astype(dt), incy, dt)
    a = a0.copy()
    A = a0v.astype(np.complex128 if cplx else np.float64)
    X, Y = xv.astype(A.dtype), yv.astype(A.dtype)
    if name in GER:
        al = (0.5, -1.25) if cplx else -1.25
        alc = complex(*al) if cplx else al
        getattr(port, name)(m, n, al, x, incx, y, incy, a, lda)
        Yc = np.conj(Y) if name.endswith("c") else Y
        want = A + alc * np.outer(X, Yc)
        want[:, 2] = A[:, 2]  # skipped (coefficient zero): untouched, -0.0 kept
    else:
        uplo = 1 if incx == 2 else 0
        stored = np.triu(np.ones((n, n), bool)) if uplo == 0 else np.tril(np.ones((n, n), bool))
        if name in SYR:
            al = 0.75
            getattr(port, name)(uplo, n, al, x, incx, a, lda)
            upd = al * np.outer(X, np.conj(X) if cplx else X)
            skip = [3]
        else:
            al = (0.5, -1.25) if cplx else -1.25
            alc = complex(*al) if cplx else al
            getattr(port, name)(uplo, n, al, x, incx, y, incy, a, lda)
            if cplx:
                upd = alc * np.outer(X, np.conj(Y)) + np.conj(alc) * np.outer(Y, np.conj(X))
            else:
                upd = alc * (np.outer(X, Y) + np.outer(Y, X))
            skip = [5]  # column 2 has alpha*y_2 == 0 but alpha*x_2 != 0: updated
        want = np.where(stored, A + upd, A)
        for j in skip:
            want[:, j] = A[:, j]
        if cplx:
            # Im(a_jj) = 0 for updated columns; HER2 with unit strides: every column
            for j in range(n):
                updated = j not in skip
                if name in SYR2 and strides == (1, 1):
                    updated = True  # HER2 with unit strides: every column
                if updated:
                    want[j, j] = want[j, j].real
    got = logical(a, m, n, lda).astype(A.dtype)
    assert np.array_equal(got, want), name
    # -0.0 survives in skipped columns
    if cplx and name in SYR2:  # Im(a_55) = -0.0: cleared (+0.0) only with unit strides
        assert np.signbit(logical(a, m, n, lda)[5, 5].imag) == (strides != (1, 1)), name
    for j in [zero_col]:
        col = logical(a, m, n, lda)[:, j]
        part = col.real if cplx else col
        rows = np.arange(m) if name in GER else (np.arange(j + 1) if uplo == 0 else np.arange(j, n))
        rows = [r for r in rows if not (cplx and r == j)]
        assert np.all(np.signbit(part[rows])), name


def test_skipped_column_ignores_non_finite_x():
    """A zero coefficient skips the column, so inf in x never reaches it."""
    x = np.array([1.0, np.inf, 2.0])
    y = np.array([1.0, 0.0])
    a = np.zeros(6)
    port.dger(3, 2, 1.0, x, 1, y, 1, a, 3)
    assert np.isinf(a[1]) and np.all(np.isfinite(a[3:]))


# ---------------------------------------------------------------- rounding modes
F_TINY = {np.float32: (1 + 2.0 ** -13, 2.0 ** -26), np.float64: (1 + 2.0 ** -28, 2.0 ** -56)}
Q = {"s": 8, "d": 4, "c": 4, "z": 2}


def _fusion_inputs(p, length):
    """x_i = y_i = u, alpha = 1, a = -(u*u rounded): fused -> tiny, two roundings -> 0."""
    real = REAL[p]
    u, tiny = F_TINY[real]
    uu = real(real(u) * real(u))
    # the rounded product misses the exact one by exactly tiny
    assert Fraction(float(u)) ** 2 - Fraction(float(uu)) == Fraction(tiny)
    return real(u), uu, tiny


def _expect_fused(mode, t, length, q):
    if mode == "fused":
        return True
    if mode == "plain":
        return False
    return t < length // q * q


@pytest.mark.parametrize("p", "sdcz")
@pytest.mark.parametrize("n", [1, 5, 13, 17])
def test_axpy_rounding_modes(p, n):
    u, uu, tiny = _fusion_inputs(p, n)
    dt = DT[p]
    for inc, mode in ((1, "unit"), (2, "plain")):
        x = strided(np.full(n, u, dt), inc, dt)
        y = strided(np.full(n, -uu, dt), inc, dt)
        getattr(port, f"{p}axpy")(n, (float(u), 0.0) if p in "cz" else u, x, inc, y, inc)
        want = [tiny if _expect_fused(mode, t, n, Q[p]) else 0.0 for t in range(n)]
        assert np.array_equal(np.real(y[::inc]), np.array(want, REAL[p])), (p, inc)


@pytest.mark.parametrize("name", GER)
def test_ger_rounding_modes(name):
    p = name[0]
    dt = DT[p]
    m, n = 13, 3
    u, uu, tiny = _fusion_inputs(p, m)
    for incx, incy in ((1, 1), (2, 1), (1, 3)):
        mode = "unit" if (incx, incy) == (1, 1) else ("fused" if p in "sd" else "plain")
        x = strided(np.full(m, u, dt), incx, dt)
        y = strided(np.full(n, u, dt), incy, dt)
        a = np.full(m * n, -uu, dt)
        getattr(port, name)(m, n, (1.0, 0.0) if p in "cz" else 1.0, x, incx, y, incy, a, m)
        want = np.array([tiny if _expect_fused(mode, i, m, Q[p]) else 0.0 for i in range(m)], REAL[p])
        for j in range(n):
            assert np.array_equal(np.real(a[j * m:(j + 1) * m]), want), (name, incx, incy, j)
            if p in "cz":
                assert np.all(np.imag(a[j * m:(j + 1) * m]) == 0)


@pytest.mark.parametrize("name", SYR + SYR2)
@pytest.mark.parametrize("uplo", [0, 1])
def test_syr_her_rounding_modes(name, uplo):
    p = name[0]
    dt = DT[p]
    two = name in SYR2
    n = 11
    u, uu, tiny = _fusion_inputs(p, n)
    # SYR2/HER2: a = -2*fl(u*u); first update -> -fl(u*u) either way (see below),
    # second update -> tiny when fused, 0 when rounded twice.
    start = -2 * uu if two else -uu
    strides = [(1, 1), (2, 1), (1, 3)] if two else [(1, 1), (2, 1)]
    for incx, incy in strides:
        if p in "sd":
            mode2 = "unit" if (incx == 1 and (not two or incy == 1)) else "fused"
        else:
            # complex: each caxpy fuses its body only for a unit source stride
            mode2 = "unit" if (incy if two else incx) == 1 else "plain"
        x = strided(np.full(n, u, dt), incx, dt)
        y = strided(np.full(n, u, dt), incy, dt)
        a = np.full(n * n, start, dt)
        if two:
            getattr(port, name)(uplo, n, (1.0, 0.0) if p in "cz" else 1.0, x, incx, y, incy, a, n)
        else:
            getattr(port, name)(uplo, n, 1.0, x, incx, a, n)
        A = a.reshape(n, n).T
        for j in range(n):
            lo, hi = (0, j + 1) if uplo == 0 else (j, n)
            for i in range(n):
                got = np.real(A[i, j])
                if not (lo <= i < hi):
                    assert got == start, (name, i, j)
                    continue
                fused = _expect_fused(mode2, i - lo, hi - lo, Q[p])
                assert got == (tiny if fused else 0.0), (name, uplo, incx, incy, i, j)
