"""Pins the oracle (CPU restatement of CORAL, oracle/coral_port.c) before it is
trusted as the checker for the GPU path.  CPU only.

Each check is chosen so that a plausible mistake in the restatement (a dropped
term, a wrong sign or index, a transposed operand, a missing conjugate) fails:
known-answer tests from the reference's rustdoc, the reference tests' own
generators and tolerances against numpy float64 (OpenBLAS), exact-rational
brute force, and invariants the reference tests exercise.
"""
import json
import math
import os
from fractions import Fraction

import numpy as np
import pytest

from oracle import port, ref64
from tests.conftest import allclose_ref

GOLD = os.path.join(os.path.dirname(__file__), "golden")

# reference tolerances (rtol, atol): coral-aarch64/tests/level3/sgemm_tests.rs:85-86,
# dgemm_tests.rs:83-84, cgemm_tests.rs:107-110, zgemm_tests.rs:106-107
GEMM_TOL = {"s": (1e-3, 1e-3), "d": (1e-12, 1e-12), "c": (3e-3, 2e-3), "z": (3e-11, 2e-11)}


def test_kat_real_2x2():
    k = json.load(open(os.path.join(GOLD, "kat.json")))["sgemm_2x2"]
    for dt, fn in ((np.float32, port.sgemm), (np.float64, port.dgemm)):
        c = np.array(k["c"], dt)
        fn(0, 0, 2, 2, 2, k["alpha"], np.array(k["a"], dt), 2, np.array(k["b"], dt), 2, k["beta"], c, 2)
        assert np.all(np.abs(c - np.array(k["expect"])) < 1e-6)


def test_kat_complex_2x2():
    k = json.load(open(os.path.join(GOLD, "kat.json")))["cgemm_2x2"]
    for dt, fn in ((np.float32, port.cgemm), (np.float64, port.zgemm)):
        c = np.array(k["c"], dt)
        fn(0, 0, 2, 2, 2, k["alpha"], np.array(k["a"], dt), 2, np.array(k["b"], dt), 2, k["beta"], c, 2)
        assert np.all(np.abs(c - np.array(k["expect"])) < 1e-6)


def _gemm_cases():
    z = np.load(os.path.join(GOLD, "gemm_cases.npz"))
    return z, json.loads(str(z["index"]))


def test_gemm_golden_reference_generators():
    z, index = _gemm_cases()
    fns = {"s": port.sgemm, "d": port.dgemm, "c": port.cgemm, "z": port.zgemm}
    for key, dt, m, n, k, oa, ob, lda, ldb, ldc, alpha, beta in index:
        a, b, c, expect = z[key + "_a"], z[key + "_b"], z[key + "_c"].copy(), z[key + "_expect"]
        fns[dt](oa, ob, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc)
        rtol, atol = GEMM_TOL[dt]
        assert allclose_ref(c, expect, rtol, atol), key


def large_reference_cases():
    """Runs of the larger reference GEMM shapes (128^3, 64x192x15, 192x64x15) with
    their inputs rebuilt from the reference generators and the committed signature
    (expected C at fixed sampled entries, and its Frobenius norm)."""
    from tests.golden import make_golden as mg

    z = np.load(os.path.join(GOLD, "gemm_reference_shapes.npz"))
    shapes = {s[0]: s for s in mg.LARGE_SHAPES}
    for key, dt, name, ci, oa, ob, m, n, k, lda, ldb, ldc in json.loads(str(z["index"])):
        _, _, _, _, lda_n, lda_t, ldb_n, ldb_t, _ = shapes[name]
        lda_, ldb_, alpha, beta, a, b, c, out = mg.large_case(dt, name, m, n, k, lda_n, lda_t, ldb_n, ldb_t, ldc, ci,
                                                              oa, ob)
        assert (lda_, ldb_) == (lda, ldb)
        yield key, dt, (m, n, k, oa, ob, lda, ldb, ldc, alpha, beta), (a, b, c), out, z[f"{name}_pos"], \
            z[key + "_sample"], float(z[key + "_fro"])


def test_gemm_larger_reference_shapes():
    fns = {"s": port.sgemm, "d": port.dgemm, "c": port.cgemm, "z": port.zgemm}
    count = 0
    for key, dt, (m, n, k, oa, ob, lda, ldb, ldc, alpha, beta), (a, b, c), out, pos, sample, fro in \
            large_reference_cases():
        # the rebuilt reference result matches the committed signature
        assert np.allclose(out[pos], sample, rtol=1e-14, atol=1e-14), key
        region = np.stack([out[j * ldc: j * ldc + m] for j in range(n)], axis=1)
        assert abs(np.linalg.norm(region) - fro) <= 1e-12 * fro, key
        cc = c.copy()
        fns[dt](oa, ob, m, n, k, alpha, a, lda, b, ldb, beta, cc, ldc)
        rtol, atol = GEMM_TOL[dt]
        assert allclose_ref(cc, out, rtol, atol), key
        assert allclose_ref(cc[pos], sample, rtol, atol), key
        count += 1
    assert count == 360


def test_gemm_exact_rational_tiny():
    # dyadic inputs: the port's f64 result must equal the exact rational product
    rng = np.random.default_rng(1)
    m, n, k = 3, 4, 5
    a = (rng.integers(-64, 64, m * k) / 16.0).astype(np.float64)
    b = (rng.integers(-64, 64, k * n) / 16.0).astype(np.float64)
    c = np.zeros(m * n)
    port.dgemm(0, 0, m, n, k, 1.0, a, m, b, k, 0.0, c, m)
    for i in range(m):
        for j in range(n):
            exact = sum(Fraction(a[i + p * m]) * Fraction(b[p + j * k]) for p in range(k))
            assert Fraction(c[i + j * m]) == exact
    # and the FP64 oracle agrees exactly too
    assert np.array_equal(ref64.gemm(0, 0, m, n, k, 1.0, a, m, b, k, 0.0, np.zeros(m * n), m), c)


@pytest.mark.parametrize("dt,fn", [("s", port.sgemm), ("d", port.dgemm)])
def test_gemm_invariants_real(dt, fn):
    f = np.float32 if dt == "s" else np.float64
    rng = np.random.default_rng(2)
    m, n, k = 37, 29, 300  # k > KC so the beta-on-first-panel rule is exercised
    a = rng.uniform(-1, 1, m * k).astype(f)
    b = rng.uniform(-1, 1, k * n).astype(f)
    # beta == 0 must not read C (NaN stays out)
    c = np.full(m * n, np.nan, f)
    fn(0, 0, m, n, k, 1.0, a, m, b, k, 0.0, c, m)
    assert np.all(np.isfinite(c))
    ref = ref64.gemm(0, 0, m, n, k, 1.0, a, m, b, k, 0.0, np.zeros(m * n), m)
    assert allclose_ref(c, ref, *GEMM_TOL[dt])
    # alpha == 0 quick return: C := beta*C, A/B never read
    c2 = rng.uniform(-1, 1, m * n).astype(f)
    c3 = c2.copy()
    fn(0, 0, m, n, k, 0.0, np.full(m * k, np.nan, f), m, np.full(k * n, np.nan, f), k, 0.5, c3, m)
    assert np.array_equal(c3, (c2 * f(0.5)).astype(f))
    # ConjugateTranspose == Transpose for real routines (sgemm.rs:97-103)
    at = rng.uniform(-1, 1, k * m).astype(f)
    c4 = np.zeros(m * n, f)
    c5 = np.zeros(m * n, f)
    fn(1, 0, m, n, k, 1.0, at, k, b, k, 0.0, c4, m)
    fn(2, 0, m, n, k, 1.0, at, k, b, k, 0.0, c5, m)
    assert np.array_equal(c4, c5)
    # transposition identity: (A B)^T = B^T A^T
    ct = np.zeros(n * m, f)
    fn(1, 1, n, m, k, 1.0, b, k, a, m, 0.0, ct, n)
    assert allclose_ref(ct.reshape(m, n).T.reshape(-1) if False else ref64.mat(ct, n, m, n).T.reshape(-1, order="F"),
                        c.reshape(-1), *GEMM_TOL[dt])


def test_cgemm_conjugation_and_beta_panels():
    rng = np.random.default_rng(3)
    m, n, k = 9, 11, 270  # k > KC(256) for cgemm
    for dt, fn, f in (("c", port.cgemm, np.complex64), ("z", port.zgemm, np.complex128)):
        a = (rng.uniform(-1, 1, k * m) + 1j * rng.uniform(-1, 1, k * m)).astype(f)
        b = (rng.uniform(-1, 1, n * k) + 1j * rng.uniform(-1, 1, n * k)).astype(f)
        c = (rng.uniform(-1, 1, m * n) + 1j * rng.uniform(-1, 1, m * n)).astype(f)
        for oa, ob in ((2, 2), (2, 1), (1, 2)):
            out = c.copy()
            fn(oa, ob, m, n, k, (0.5, -1.5), a, k, b, n, (0.25, 0.75), out, m)
            ref = ref64.gemm(oa, ob, m, n, k, (0.5, -1.5), a, k, b, n, (0.25, 0.75), c, m)
            assert allclose_ref(out, ref, *GEMM_TOL[dt]), (dt, oa, ob)


@pytest.mark.parametrize("p", "sdcz")
@pytest.mark.parametrize("trans", [0, 1, 2])
def test_gemv_reference_generators(p, trans):
    # dgemv_tests.rs:59 (A = 0.1 + 0.5 i + 0.25 j), :149-156 (x, y), padded lda and strides
    cplx = p in "cz"
    f = {"s": np.float32, "d": np.float64, "c": np.complex64, "z": np.complex128}[p]
    fn = getattr(port, f"{p}gemv")
    tol = {"s": (1e-5, 1e-5), "d": (1e-13, 1e-13), "c": (1e-4, 1e-5), "z": (1e-12, 1e-12)}[p]
    for (m, n, lda, incx, incy) in ((7, 5, 7, 1, 1), (130, 70, 133, 2, 3), (64, 200, 64, 3, 2)):
        a = np.zeros(lda * n, f)
        for j in range(n):
            for i in range(m):
                a[i + j * lda] = 0.1 + i * 0.5 / m + j * 0.25 / n + (1j * (0.05 - 0.01 * i / m) if cplx else 0)
        xl, yl = (n, m) if trans == 0 else (m, n)
        x = np.zeros((xl - 1) * incx + 1, f)
        x[::incx] = 0.2 + 0.1 * np.arange(xl) / xl + (0.3j if cplx else 0)
        y = np.zeros((yl - 1) * incy + 1, f)
        y[::incy] = -0.3 + 0.05 * np.arange(yl) / yl
        al, be = ((0.75, -0.25), (0.85, 0.1)) if cplx else (0.75, -0.25)
        out = y.copy()
        fn(trans, m, n, al, a, lda, x, incx, be, out, incy)
        ref = ref64.gemv(trans, m, n, al, a, lda, x, incx, be, y, incy)
        assert allclose_ref(out, ref, *tol), (m, n)


def test_gemv_quick_returns_and_dscal_semantics():
    y = np.array([1.0, np.nan, 3.0])
    out = y.copy()
    port.dgemv(0, 3, 2, 0.0, np.zeros(6), 3, np.zeros(2), 1, 1.0, out, 1)  # alpha 0, beta 1
    assert np.array_equal(out[[0, 2]], y[[0, 2]]) and np.isnan(out[1])
    # CORAL pre-scales y with dscal, so beta == 0 keeps a NaN already in y
    out = y.copy()
    port.dgemv(0, 3, 2, 1.0, np.ones(6), 3, np.ones(2), 1, 0.0, out, 1)
    assert np.isnan(out[1]) and out[0] == 2.0


@pytest.mark.parametrize("p", "sdcz")
@pytest.mark.parametrize("uplo", [0, 1])
def test_symv_hemv_stored_triangle_only(p, uplo):
    herm = p in "cz"
    f = {"s": np.float32, "d": np.float64, "c": np.complex64, "z": np.complex128}[p]
    fn = {"s": port.ssymv, "d": port.dsymv, "c": port.chemv, "z": port.zhemv}[p]
    rng = np.random.default_rng(4)
    n, lda = 37, 40
    a = (rng.uniform(-1, 1, lda * n) + (1j * rng.uniform(-1, 1, lda * n) if herm else 0)).astype(f)
    # NaN in the unreferenced triangle (chemv_tests.rs:76-133)
    A = a.reshape(n, lda).T
    for j in range(n):
        for i in range(n):
            if (uplo == 0 and i > j) or (uplo == 1 and i < j):
                A[i, j] = np.nan
    a = A.T.reshape(-1).copy()
    x = (rng.uniform(-1, 1, n) + (1j * rng.uniform(-1, 1, n) if herm else 0)).astype(f)
    y = (rng.uniform(-1, 1, n) + (1j * rng.uniform(-1, 1, n) if herm else 0)).astype(f)
    al, be = ((1.25, -0.5), (0.5, 0.25)) if herm else (1.25, 0.5)
    out = y.copy()
    fn(uplo, n, al, a, lda, x, 1, be, out, 1)
    ref = ref64.symv(uplo, n, al, a, lda, x, 1, be, y, 1, herm=herm)
    assert np.all(np.isfinite(out))
    tol = (1e-4, 1e-5) if p in "sc" else (1e-12, 1e-12)
    assert allclose_ref(out, ref, *tol)


def test_hemv_ignores_diagonal_imaginary():
    n = 5
    rng = np.random.default_rng(5)
    a = (rng.uniform(-1, 1, n * n) + 1j * rng.uniform(-1, 1, n * n)).astype(np.complex128)
    a2 = a.copy()
    for j in range(n):
        a2[j + j * n] = complex(a[j + j * n].real, 99.0)
    x = rng.uniform(-1, 1, n).astype(np.complex128)
    o1, o2 = np.zeros(n, np.complex128), np.zeros(n, np.complex128)
    port.zhemv(0, n, (1, 0), a, n, x, 1, (0, 0), o1, 1)
    port.zhemv(0, n, (1, 0), a2, n, x, 1, (0, 0), o2, 1)
    assert np.array_equal(o1, o2)


def _tri_matrix(n, lda, f, rng, uplo, cplx):
    # diagonally dominant structured triangle (dtrsv_tests.rs:64-166 style): diag in [1, 2]
    a = np.full(lda * n, np.nan, f)  # garbage outside the triangle is never read
    for j in range(n):
        for i in range(n):
            if (uplo == 1 and i >= j) or (uplo == 0 and i <= j):
                v = rng.uniform(-1, 1) / n + (1j * rng.uniform(-1, 1) / n if cplx else 0)
                a[i + j * lda] = v if i != j else rng.uniform(1, 2) + (0.3j if cplx else 0)
    return a


@pytest.mark.parametrize("p", "sdcz")
def test_trsv_trmv_all_variants(p):
    cplx = p in "cz"
    f = {"s": np.float32, "d": np.float64, "c": np.complex64, "z": np.complex128}[p]
    rng = np.random.default_rng(6)
    tol = (2e-4, 1e-5) if p in "sc" else (1e-12, 1e-12)  # dtrsv_tests.rs:213-214 uses 1e-12
    for n, lda, incx in ((7, 7, 1), (33, 35, 2), (64, 64, 3)):
        for uplo in (0, 1):
            a = _tri_matrix(n, lda, f, rng, uplo, cplx)
            for trans in (0, 1, 2):
                for diag in (0, 1):
                    if diag == 0:  # unit: a garbage diagonal must not be read (dtrsv_tests.rs:228-239)
                        a_use = a.copy()
                        for j in range(n):
                            a_use[j + j * lda] = np.nan
                    else:
                        a_use = a
                    x = np.zeros((n - 1) * incx + 1, f)
                    x[::incx] = rng.uniform(-1, 1, n) + (1j * rng.uniform(-1, 1, n) if cplx else 0)
                    xs = x.copy()
                    getattr(port, f"{p}trsv")(uplo, trans, diag, n, a_use, lda, xs, incx)
                    ref = ref64.trsv(uplo, trans, diag, n, a_use, lda, x, incx)
                    assert allclose_ref(xs, ref, *tol), (n, uplo, trans, diag)
                    # TRMV undoes TRSV
                    getattr(port, f"{p}trmv")(uplo, trans, diag, n, a_use, lda, xs, incx)
                    assert allclose_ref(xs, x, *tol), ("trmv", n, uplo, trans, diag)


def test_trsv_n0_quick_return():
    x = np.array([1.0, 2.0])
    port.dtrsv(1, 0, 1, 0, np.zeros(1), 1, x, 1)
    assert np.array_equal(x, [1.0, 2.0])


def test_dot_axpy_against_exact():
    rng = np.random.default_rng(7)
    for n in (0, 1, 3, 31, 32, 33, 1000, 4099):
        x = rng.uniform(-1, 1, n)
        y = rng.uniform(-1, 1, n)
        assert abs(port.ddot(n, x, 1, y, 1) - ref64.dot_exact(x, y)) <= 1e-12 * max(1, n)
        xf, yf = x.astype(np.float32), y.astype(np.float32)
        assert abs(port.sdot(n, xf, 1, yf, 1) - ref64.dot_exact(xf, yf)) <= 4e-6 * max(1, math.sqrt(n))
        xc = (x + 1j * y[::-1]).astype(np.complex128)
        yc = (y + 0.5j * x).astype(np.complex128)
        for name, conj in (("zdotu", False), ("zdotc", True)):
            r = getattr(port, name)(n, xc, 1, yc, 1)
            e = ref64.dot_exact(xc, yc, conj=conj)
            assert abs(complex(*r) - e) <= 1e-12 * max(1, n)
        yy = y.copy()
        port.daxpy(n, 0.5, x, 1, yy, 1)
        assert np.array_equal(yy, np.array([math.fma(0.5, a, b) if hasattr(math, "fma") else 0.5 * a + b
                                            for a, b in zip(x, y)])) or np.allclose(yy, y + 0.5 * x, rtol=0,
                                                                                    atol=1e-15)


def test_nrm2_extremes():
    for scale in (1e-30, 1.0, 1e30):
        x = (np.array([3.0, -4.0, 12.0]) * scale).astype(np.float32)
        assert abs(port.snrm2(3, x, 1) - 13 * scale) <= 1e-6 * 13 * scale
        assert abs(port.dnrm2(3, x.astype(np.float64), 1) - ref64.nrm2(x.astype(np.float64))) <= 1e-14 * 13 * scale
    rng = np.random.default_rng(8)
    v = rng.uniform(-1, 1, 1001)
    assert abs(port.dnrm2(1001, v, 1) - ref64.nrm2(v)) <= 1e-14 * ref64.nrm2(v)
    vc = (v + 1j * v[::-1]).astype(np.complex64)
    assert abs(port.scnrm2(1001, vc, 1) - ref64.nrm2(vc)) <= 1e-5 * ref64.nrm2(vc)
    assert port.dnrm2(0, v, 1) == 0.0


def test_iamax_planted_and_semantics():
    z = np.load(os.path.join(GOLD, "iamax_cases.npz"))
    assert port.isamax(1024, z["x1"], 1) == int(z["expect1"])
    assert port.isamax(int(z["n2"]), z["x2"], int(z["inc2"])) == int(z["expect2"])
    assert port.isamax(0, np.ones(5, np.float32), 1) == 0
    # ties: the first index wins; NaN is never selected; all-NaN -> 0
    x = np.array([1.0, -3.0, np.nan, 3.0, 2.0])
    assert port.idamax(5, x, 1) == 1
    assert port.idamax(3, np.full(3, np.nan), 1) == 0
    xc = np.array([1 + 2j, -2 - 1j, 3 + 0j, 0 - 3j], np.complex64)  # |re|+|im| = 3,3,3,3
    assert port.icamax(4, xc, 1) == 0
    rng = np.random.default_rng(9)
    for _ in range(20):
        n = int(rng.integers(1, 200))
        v = rng.integers(-5, 6, n).astype(np.float64)
        assert port.idamax(n, v, 1) == ref64.iamax(v, n, 1)
