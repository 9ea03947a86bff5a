"""GPU parity for the checked front end (coral_amd.safe), mirroring coral-safe's tests.

Each test is one of coral-safe's (coral-safe/tests/level1/*.rs, level2/*.rs, fused/*.rs):
the same sizes, strides, lda padding and scalars, its buffer generators
(tests/common.rs:11-120) and its check, |x - y| <= ATOL + RTOL*max(|x|, |y|) with
RTOL 1e-5, ATOL 1e-6 (1e-4 where a test loosens it), against the CBLAS those tests
call: OpenBLAS through scipy.linalg.blas.  On top of that:

- element-wise routines are bit-identical to the oracle's restatement of CORAL;
- every routine, called through views with non-zero offsets, gives exactly what
  the plain host-pointer routine gives on the window that starts at the offset,
  and leaves every element outside the view untouched.
"""
import numpy as np
import pytest
from scipy.linalg import blas as sb
from scipy.linalg import solve_triangular

from coral_amd import blas, safe
from coral_amd.safe import CoralDiagonal, CoralTranspose, CoralTriangular, MatrixMut, MatrixRef, VectorMut, VectorRef
from oracle import port

pytestmark = pytest.mark.gpu
RTOL, ATOL = 1e-5, 1e-6  # coral-safe/tests/common.rs:8-9


def assert_close(a, b, rtol=RTOL, atol=ATOL):  # coral-safe/tests/common.rs:122-136
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    assert a.shape == b.shape
    tol = atol + rtol * np.maximum(np.abs(a), np.abs(b))
    bad = np.nonzero(np.abs(a - b) > tol)[0]
    assert bad.size == 0, f"mismatch at idx {bad[0]}: {a[bad[0]]} vs. {b[bad[0]]}"


def make_strided_vec(rng, n, inc):  # coral-safe/tests/common.rs:11-29
    if n == 0:
        return np.ones(1, np.float32)
    buf = np.zeros((n - 1) * inc + 1, np.float32)
    buf[::inc] = rng.random(n, dtype=np.float32)
    return buf


def make_strided_mat(rng, m, n, lda):  # coral-safe/tests/common.rs:31-54
    buf = np.zeros(lda * n, np.float32)
    buf.reshape(n, lda)[:, :m] = rng.random((n, m), dtype=np.float32)
    return buf


def make_triangular_mat(rng, uplo, diag, n, lda):  # coral-safe/tests/common.rs:56-120
    buf = np.zeros(lda * n, np.float32)
    a = buf.reshape(n, lda).T  # a[i, j] = buf[i + j*lda]
    eps = np.float32(1e-3 / max(n, 1))
    off = rng.uniform(-eps, eps, (n, n)).astype(np.float32)
    tri = np.triu(off, 1) if uplo is CoralTriangular.Upper else np.tril(off, -1)
    a[:n, :n] = tri
    d = np.ones(n, np.float32) if diag.is_unit() else rng.uniform(1.0, 2.0, n).astype(np.float32)
    a[np.arange(n), np.arange(n)] = d
    return buf


def _f64(v, n, inc, off=0):
    return np.asarray(v[off:off + (n - 1) * inc + 1:inc] if n else v[:0], np.float64)


# ---------------------------------------------------------------- level 1
@pytest.mark.parametrize("n,incx,incy", [(1024, 1, 1), (1024, 3, 5), (0, 1, 1)])
def test_saxpy(gpu_lib, n, incx, incy):  # coral-safe/tests/level1/saxpy.rs
    rng = np.random.default_rng(100 + incx)
    alpha = np.float32(3.1415926)
    x, y = make_strided_vec(rng, n, incx), make_strided_vec(rng, n, incy)
    yc = y.copy()
    po = y.copy()
    safe.saxpy(alpha, VectorRef(x, n, incx, 0), VectorMut(y, n, incy, 0))
    if n:
        ref = sb.saxpy(x, yc.copy(), n=n, a=alpha, incx=incx, incy=incy)
        assert_close(y, ref)
    else:
        assert np.array_equal(y, yc)
    port.saxpy(n, alpha, x, incx, po, incy)
    assert np.array_equal(y, po)


@pytest.mark.parametrize("n,incx", [(1024, 1), (1024, 3), (0, 1)])
def test_sscal(gpu_lib, n, incx):  # coral-safe/tests/level1/sscal.rs
    rng = np.random.default_rng(110 + incx)
    x = make_strided_vec(rng, n, incx)
    ref, po = x.copy(), x.copy()
    safe.sscal(np.float32(3.1415926), VectorMut(x, n, incx, 0))
    if n:
        ref = sb.sscal(np.float32(3.1415926), ref, n=n, incx=incx)
    assert_close(x, ref)
    port.sscal(n, np.float32(3.1415926), po, incx)
    assert np.array_equal(x, po)


@pytest.mark.parametrize("n,incx", [(1024, 1), (1024, 3), (0, 1)])
def test_sasum_snrm2_isamax(gpu_lib, n, incx):  # coral-safe/tests/level1/{sasum,snrm2,isamax}.rs
    rng = np.random.default_rng(120 + incx)
    x = make_strided_vec(rng, n, incx)
    xv = VectorRef(x, n, incx, 0)
    if n:
        assert_close([safe.sasum(xv)], [sb.sasum(x, n=n, incx=incx)])
        assert_close([safe.snrm2(xv)], [sb.snrm2(x, n=n, incx=incx)])
        vals = np.abs(_f64(x, n, incx))
        assert safe.isamax(xv) == int(np.argmax(vals))  # first index of the maximum
    else:
        assert safe.sasum(xv) == 0.0 and safe.snrm2(xv) == 0.0 and safe.isamax(xv) == 0


@pytest.mark.parametrize("n,incx,incy", [(1024, 1, 1), (1024, 3, 5), (0, 1, 1)])
def test_sdot(gpu_lib, n, incx, incy):  # coral-safe/tests/level1/sdot.rs
    rng = np.random.default_rng(130 + incx)
    x, y = make_strided_vec(rng, n, incx), make_strided_vec(rng, n, incy)
    r = safe.sdot(VectorRef(x, n, incx, 0), VectorRef(y, n, incy, 0))
    if n:
        assert_close([r], [sb.sdot(x, y, n=n, incx=incx, incy=incy)], rtol=1e-4)  # sdot.rs:36 uses 1e-4
    else:
        assert r == 0.0


@pytest.mark.parametrize("n,incx,incy", [(1024, 1, 1), (1024, 3, 5), (0, 1, 1)])
def test_scopy_sswap(gpu_lib, n, incx, incy):  # coral-safe/tests/level1/{scopy,sswap}.rs
    rng = np.random.default_rng(140 + incx)
    x, y = make_strided_vec(rng, n, incx), make_strided_vec(rng, n, incy)
    x0, y0 = x.copy(), y.copy()
    safe.scopy(VectorRef(x, n, incx, 0), VectorMut(y, n, incy, 0))
    assert np.array_equal(_f64(y, n, incy), _f64(x0, n, incx)) and np.array_equal(x, x0)
    y = y0.copy()
    safe.sswap(VectorMut(x, n, incx, 0), VectorMut(y, n, incy, 0))
    assert np.array_equal(_f64(x, n, incx), _f64(y0, n, incy))
    assert np.array_equal(_f64(y, n, incy), _f64(x0, n, incx))


@pytest.mark.parametrize("n,incx,incy", [(1024, 1, 1), (1024, 3, 5), (0, 1, 1)])
def test_srot(gpu_lib, n, incx, incy):  # coral-safe/tests/level1/srot.rs
    rng = np.random.default_rng(150 + incx)
    x, y = make_strided_vec(rng, n, incx), make_strided_vec(rng, n, incy)
    c, s = np.float32(0.8), np.float32(0.6)
    px, py = x.copy(), y.copy()
    rx, ry = x.copy(), y.copy()
    safe.srot(VectorMut(x, n, incx, 0), VectorMut(y, n, incy, 0), c, s)
    if n:
        rx, ry = sb.srot(rx, ry, c, s, n=n, incx=incx, incy=incy)
    assert_close(x, rx)
    assert_close(y, ry)
    port.srot(n, px, incx, py, incy, c, s)
    assert np.array_equal(x, px) and np.array_equal(y, py)


@pytest.mark.parametrize("flag", [-1.0, 0.0, 1.0, -2.0])
@pytest.mark.parametrize("n,incx,incy", [(1024, 1, 1), (1024, 3, 5), (0, 1, 1)])
def test_srotm(gpu_lib, n, incx, incy, flag):  # coral-safe/tests/level1/srotm.rs
    rng = np.random.default_rng(160 + incx)
    x, y = make_strided_vec(rng, n, incx), make_strided_vec(rng, n, incy)
    param = np.array([flag, 0.9, -0.2, 0.3, 1.1], np.float32)
    px, py = x.copy(), y.copy()
    rx, ry = x.copy(), y.copy()
    safe.srotm(VectorMut(x, n, incx, 0), VectorMut(y, n, incy, 0), param)
    if n:
        rx, ry = sb.srotm(rx, ry, param, n=n, incx=incx, incy=incy)
    assert_close(x, rx)
    assert_close(y, ry)
    port.srotm(n, px, incx, py, incy, param)
    assert np.array_equal(x, px) and np.array_equal(y, py)


def test_srotg(gpu_lib):  # coral-safe/tests/level1/srotg.rs:13-40
    a, b, c, s = (np.array([v], np.float32) for v in (0.5, -1.25, 0.0, 0.0))
    safe.srotg(a, b, c, s)
    rc, rs = sb.srotg(np.float32(0.5), np.float32(-1.25))
    assert_close([c[0], s[0]], [rc, rs])
    assert_close([a[0], b[0]], port.srotg(0.5, -1.25)[:2])  # r and z as CORAL leaves them


def test_srotmg(gpu_lib):  # coral-safe/tests/level1/srotmg.rs:13-42
    d1, d2, x1 = (np.array([v], np.float32) for v in (1.5, 2.0, -0.75))
    param = np.zeros(5, np.float32)
    safe.srotmg(d1, d2, x1, np.float32(0.5), param)
    assert_close(param, sb.srotmg(np.float32(1.5), np.float32(2.0), np.float32(-0.75), np.float32(0.5)))
    pd1, pd2, px1, pparam = port.srotmg(1.5, 2.0, -0.75, 0.5)
    assert_close([d1[0], d2[0], x1[0]], [pd1, pd2, px1])
    assert_close(param, pparam)


# ---------------------------------------------------------------- level 2 / fused
@pytest.mark.parametrize("op", [CoralTranspose.NoTrans, CoralTranspose.Trans])
@pytest.mark.parametrize("incx,incy,pad", [(1, 1, 0), (2, 3, 32)])
def test_sgemv(gpu_lib, op, incx, incy, pad):  # coral-safe/tests/level2/sgemv.rs
    rng = np.random.default_rng(170 + incx)
    m, n = 1024, 512
    lda = m + pad
    alpha, beta = np.float32(3.14159265), np.float32(2.71828182)
    lx, ly = (n, m) if op is CoralTranspose.NoTrans else (m, n)
    x = make_strided_vec(rng, lx, incx)
    a = make_strided_mat(rng, m, n, lda)
    y = make_strided_vec(rng, ly, incy)
    y0 = y.copy()
    safe.sgemv(op, alpha, beta, MatrixRef(a, m, n, lda, 0), VectorRef(x, lx, incx, 0), VectorMut(y, ly, incy, 0))
    a2 = a.reshape(n, lda).T[:m, :]
    ref = sb.sgemv(alpha, a2, x, beta=beta, y=y0.copy(), incx=incx, incy=incy, trans=op.value)
    assert_close(y, ref)


@pytest.mark.parametrize("incx,incy,pad", [(1, 1, 0), (2, 3, 32)])
def test_sger(gpu_lib, incx, incy, pad):  # coral-safe/tests/level2/sger.rs
    rng = np.random.default_rng(180 + incx)
    m, n = 1024, 512
    lda = m + pad
    alpha = np.float32(3.14159265)
    x, y = make_strided_vec(rng, m, incx), make_strided_vec(rng, n, incy)
    a = make_strided_mat(rng, m, n, lda)
    po = a.copy()
    ref = a.copy().reshape(n, lda).T.astype(np.float64)
    ref[:m, :] += np.float64(alpha) * np.outer(_f64(x, m, incx), _f64(y, n, incy))
    safe.sger(alpha, MatrixMut(a, m, n, lda, 0), VectorRef(x, m, incx, 0), VectorRef(y, n, incy, 0))
    assert_close(a, ref.T.reshape(-1))
    port.sger(m, n, alpha, x, incx, y, incy, po, lda)
    assert np.array_equal(a, po)  # GER reproduces CORAL's rounding


@pytest.mark.parametrize("uplo", [CoralTriangular.Upper, CoralTriangular.Lower])
@pytest.mark.parametrize("diag", [CoralDiagonal.NonUnit, CoralDiagonal.Unit])
@pytest.mark.parametrize("trans", [CoralTranspose.NoTrans, CoralTranspose.Trans])
def test_strsv(gpu_lib, uplo, diag, trans):  # coral-safe/tests/level2/strsv.rs
    rng = np.random.default_rng(190 + 4 * uplo.value + 2 * diag.value + trans.value)
    n = lda = 1024
    a = make_triangular_mat(rng, uplo, diag, n, lda)
    x = make_strided_vec(rng, n, 1)
    b = x.copy()
    safe.strsv(uplo, trans, diag, MatrixRef(a, n, n, lda, 0), VectorMut(x, n, 1, 0))
    A = a.reshape(n, lda).T[:n, :n].astype(np.float64)
    ref = solve_triangular(A, b.astype(np.float64), lower=uplo is CoralTriangular.Lower,
                           trans=trans.value, unit_diagonal=diag.is_unit())
    loose = uplo is CoralTriangular.Lower and (trans is CoralTranspose.Trans) != diag.is_unit()
    assert_close(x, ref, rtol=1e-4 if loose else RTOL)  # strsv.rs:247,286 use 1e-4


@pytest.mark.parametrize("incx,incy", [(1, 1), (2, 3)])
def test_saxpyf_sdotf(gpu_lib, incx, incy):  # coral-safe/tests/fused/{saxpyf,sdotf}.rs
    rng = np.random.default_rng(200 + incx)
    n = lda = 1024
    a = make_strided_mat(rng, n, n, lda)
    A = a.reshape(n, lda).T.astype(np.float64)
    x, y = make_strided_vec(rng, n, incx), make_strided_vec(rng, n, incy)
    y0 = y.copy()
    safe.saxpyf(MatrixRef(a, n, n, lda, 0), VectorRef(x, n, incx, 0), VectorMut(y, n, incy, 0))
    ref = y0.astype(np.float64)
    ref[::incy] += A @ _f64(x, n, incx)
    assert_close(y, ref)
    y = y0.copy()
    safe.sdotf(MatrixRef(a, n, n, lda, 0), VectorRef(x, n, incx, 0), VectorMut(y, n, incy, 0))
    ref = y0.astype(np.float64)
    ref[::incy] += A.T @ _f64(x, n, incx)
    assert_close(y, ref)


# ---------------------------------------------------------------- offsets
def test_offset_views_match_the_plain_routines(gpu_lib):
    """Views with offsets reach exactly the window that starts at the offset."""
    rng = np.random.default_rng(210)
    m, n, lda, off_a, off_x, off_y = 300, 200, 317, 11, 5, 7
    a = rng.standard_normal(off_a + lda * n + 9).astype(np.float32)
    x = rng.standard_normal(off_x + 3 * m + 4).astype(np.float32)
    y = rng.standard_normal(off_y + 2 * m + 6).astype(np.float32)

    def both(f_safe, f_plain, bufs):
        got = [b.copy() for b in bufs]
        exp = [b.copy() for b in bufs]
        r1 = f_safe(*got)
        r2 = f_plain(*exp)
        for g, e in zip(got, exp):
            assert np.array_equal(g, e)
        assert (r1 == r2) or (np.isnan(r1) and np.isnan(r2))

    k = 190
    both(lambda x, y: safe.saxpy(1.5, VectorRef(x, k, 3, off_x), VectorMut(y, k, 2, off_y)),
         lambda x, y: blas.saxpy(k, 1.5, x[off_x:], 3, y[off_y:], 2), [x, y])
    both(lambda x, y: safe.srot(VectorMut(x, k, 3, off_x), VectorMut(y, k, 2, off_y), 0.6, 0.8),
         lambda x, y: blas.srot(k, x[off_x:], 3, y[off_y:], 2, 0.6, 0.8), [x, y])
    both(lambda x: safe.sscal(-0.5, VectorMut(x, k, 3, off_x)),
         lambda x: blas.sscal(k, -0.5, x[off_x:], 3), [x])
    both(lambda x: safe.snrm2(VectorRef(x, k, 3, off_x)), lambda x: blas.snrm2(k, x[off_x:], 3), [x])
    both(lambda x: safe.isamax(VectorRef(x, k, 3, off_x)), lambda x: blas.isamax(k, x[off_x:], 3), [x])
    both(lambda x, y: safe.sdot(VectorRef(x, k, 3, off_x), VectorRef(y, k, 2, off_y)),
         lambda x, y: blas.sdot(k, x[off_x:], 3, y[off_y:], 2), [x, y])
    both(lambda a, x, y: safe.sgemv(CoralTranspose.NoTrans, 0.5, 2.0, MatrixRef(a, m, n, lda, off_a),
                                    VectorRef(y, n, 2, off_y), VectorMut(x, m, 3, off_x)),
         lambda a, x, y: blas.sgemv(0, m, n, 0.5, a[off_a:], lda, y[off_y:], 2, 2.0, x[off_x:], 3), [a, x, y])
    both(lambda a, x, y: safe.sger(0.25, MatrixMut(a, m, n, lda, off_a), VectorRef(x, m, 3, off_x),
                                   VectorRef(y, n, 2, off_y)),
         lambda a, x, y: blas.sger(m, n, 0.25, x[off_x:], 3, y[off_y:], 2, a[off_a:], lda), [a, x, y])
    tri = np.tril(rng.uniform(-1e-3, 1e-3, (n, n))) + np.diag(rng.uniform(1, 2, n))
    at = np.zeros(off_a + lda * n, np.float32)
    at[off_a:].reshape(n, lda)[:, :n] = tri.T.astype(np.float32)
    both(lambda a, x: safe.strsv(CoralTriangular.Lower, CoralTranspose.Trans, CoralDiagonal.NonUnit,
                                 MatrixRef(a, n, n, lda, off_a), VectorMut(x, n, 3, off_x)),
         lambda a, x: blas.strsv(1, 1, 1, n, a[off_a:], lda, x[off_x:], 3), [at, x])
