"""GPU parity for GEMV / SYMV / HEMV / TRSV / TRMV through the C ABI.

Compared with the oracle (CORAL restatement) and the FP64 reference with the
reference tests' tolerances (dgemv_tests.rs:117-118, sgemv_tests.rs:116-117,
chemv_tests.rs:190-191, dtrsv_tests.rs:213-214), plus the backward error
|b - A x| / (|A| |x| + |b|) for TRSV.
"""
import numpy as np
import pytest

from coral_amd import blas
from oracle import port, ref64
from tests.conftest import allclose_ref

pytestmark = pytest.mark.gpu
DT = {"s": np.float32, "d": np.float64, "c": np.complex64, "z": np.complex128}
TOL = {"s": (1e-5, 1e-5), "d": (1e-13, 1e-13), "c": (1e-5, 1e-6), "z": (1e-12, 1e-12)}


def rand(rng, n, f):
    if np.issubdtype(f, np.complexfloating):
        return (rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)).astype(f)
    return rng.uniform(-1, 1, n).astype(f)


@pytest.mark.parametrize("p", "sdcz")
def test_gemv_all_variants(gpu_lib, p):
    rng = np.random.default_rng(20)
    f = DT[p]
    cplx = p in "cz"
    fn = getattr(blas, f"{p}gemv")
    pfn = getattr(port, f"{p}gemv")
    for (m, n, lda, incx, incy) in ((1, 1, 1, 1, 1), (7, 5, 9, 1, 1), (1024, 512, 1024, 1, 1),
                                    (1000, 3000, 1003, 2, 3), (4097, 33, 4100, 3, 2), (33, 4097, 33, 1, 1)):
        a = rand(rng, lda * n, f)
        for trans in (0, 1, 2):
            xl, yl = (n, m) if trans == 0 else (m, n)
            x = rand(rng, (xl - 1) * incx + 1, f)
            y = rand(rng, (yl - 1) * incy + 1, f)
            al, be = ((0.75, -0.25), (0.85, 0.1)) if cplx else (0.75, -0.25)
            out = y.copy()
            fn(trans, m, n, al, a, lda, x, incx, be, out, incy)
            ref = ref64.gemv(trans, m, n, al, a, lda, x, incx, be, y, incy)
            scale = np.abs(ref).max() + 1e-30
            rtol = 1e-5 if p in "sc" else 1e-13
            err = np.abs(out.astype(ref.dtype) - ref).max() / scale
            assert err <= rtol * max(1.0, np.sqrt(xl) / 8), (m, n, trans, err)
            po = y.copy()
            pfn(trans, m, n, al, a, lda, x, incx, be, po, incy)
            assert np.abs(out - po).max() / scale <= rtol * max(1.0, np.sqrt(xl) / 8)


def test_gemv_quick_returns_and_beta_zero(gpu_lib):
    y = np.array([1.0, np.nan, 3.0])
    out = y.copy()
    blas.dgemv(0, 3, 2, 0.0, np.zeros(6), 3, np.zeros(2), 1, 1.0, out, 1)
    assert out[0] == 1.0 and np.isnan(out[1]) and out[2] == 3.0
    # beta == 0 overwrites y (BLAS semantics; CORAL's dscal pre-scaling keeps NaN)
    out = y.copy()
    blas.dgemv(0, 3, 2, 1.0, np.ones(6), 3, np.ones(2), 1, 0.0, out, 1)
    assert np.array_equal(out, [2.0, 2.0, 2.0])
    out = y.copy()
    blas.dgemv(1, 3, 2, 0.0, np.ones(6), 3, np.ones(3), 1, 2.0, out, 1)  # alpha 0: y := beta*y (y has n = 2)
    assert out[0] == 2.0 and np.isnan(out[1]) and out[2] == 3.0


@pytest.mark.parametrize("p", "sdcz")
@pytest.mark.parametrize("uplo", [0, 1])
def test_symv_hemv(gpu_lib, p, uplo):
    herm = p in "cz"
    f = DT[p]
    fn = {"s": blas.ssymv, "d": blas.dsymv, "c": blas.chemv, "z": blas.zhemv}[p]
    rng = np.random.default_rng(21)
    for n, lda, incx, incy in ((1, 1, 1, 1), (37, 40, 1, 1), (128, 128, 2, 1), (300, 301, 1, 3), (1000, 1000, 1, 1)):
        a = rand(rng, lda * n, f)
        A = a.reshape(n, lda).T.copy()
        for j in range(n):  # NaN in the unreferenced triangle
            lo, hi = (j + 1, n) if uplo == 0 else (0, j)
            A[lo:hi, j] = np.nan
        a = A.T.reshape(-1).copy()
        x = rand(rng, (n - 1) * incx + 1, f)
        y = rand(rng, (n - 1) * incy + 1, f)
        al, be = ((1.25, -0.5), (0.5, 0.25)) if herm else (1.25, 0.5)
        out = y.copy()
        fn(uplo, n, al, a, lda, x, incx, be, out, incy)
        ref = ref64.symv(uplo, n, al, a, lda, x, incx, be, y, incy, herm=herm)
        assert np.all(np.isfinite(out))
        rtol = 2e-5 if p in "sc" else 1e-12
        assert np.abs(out - ref).max() <= rtol * (np.abs(ref).max() + 1) * max(1, np.sqrt(n) / 8), (n,)
    # upper and lower storage of the same matrix agree
    n = 65
    M = rand(rng, n * n, f).reshape(n, n)
    H = (M + M.conj().T) / 2 if herm else (M + M.T) / 2
    h = np.asfortranarray(H).reshape(-1, order="F").astype(f)
    x = rand(rng, n, f)
    o_u, o_l = np.zeros(n, f), np.zeros(n, f)
    one, zero = ((1, 0), (0, 0)) if herm else (1.0, 0.0)
    fn(0, n, one, h, n, x, 1, zero, o_u, 1)
    fn(1, n, one, h, n, x, 1, zero, o_l, 1)
    assert np.abs(o_u - o_l).max() <= (1e-5 if p in "sc" else 1e-13) * max(1, np.abs(o_u).max())


@pytest.mark.parametrize("cfg", ["blocks", None, "4,1", "2,2", "2,5"],
                         ids=["blocks", "stream", "stream-nv4", "stream-wg2", "stream-wg5"])
@pytest.mark.parametrize("p", "sdcz")
def test_symv_hemv_row_groups(gpu_lib, p, cfg, monkeypatch):
    """Orders that span many row groups of the streaming kernel (a row group is 64 x NV x
    (16 / element size) rows: 512 for float with two vectors per column, 128 for complex
    double), ragged against row groups, column runs and four-column groups, with padded
    lda, strided x and y, NaN in the unstored triangle, both triangles: by 64 x 64 blocks
    (these orders' default) and streamed (CORAL_SYMV_STREAM_MIN_N=0) in every shape
    CORAL_SYMV_CFG selects (vectors per column, workgroups per CU, which sets the run
    length); two calls give the same bits."""
    if cfg != "blocks":
        monkeypatch.setenv("CORAL_SYMV_STREAM_MIN_N", "0")
        if cfg:
            monkeypatch.setenv("CORAL_SYMV_CFG", cfg)
    herm = p in "cz"
    f = DT[p]
    fn = {"s": blas.ssymv, "d": blas.dsymv, "c": blas.chemv, "z": blas.zhemv}[p]
    rng = np.random.default_rng(22)
    for uplo in (0, 1):
        for n, lda, incx, incy in ((2049, 2051, 2, 1), (4133, 4133, 1, 3)):
            A = rand(rng, lda * n, f).reshape(n, lda).T.copy()
            unstored = np.tril(np.ones((n, n), bool), -1) if uplo == 0 else np.triu(np.ones((n, n), bool), 1)
            A[:n][unstored] = np.nan
            a = A.T.reshape(-1).copy()
            x = rand(rng, (n - 1) * incx + 1, f)
            y = rand(rng, (n - 1) * incy + 1, f)
            al, be = ((0.5, 0.75), (-0.25, 0.5)) if herm else (0.75, -0.5)
            out = y.copy()
            fn(uplo, n, al, a, lda, x, incx, be, out, incy)
            again = y.copy()
            fn(uplo, n, al, a, lda, x, incx, be, again, incy)
            assert np.array_equal(out, again)
            ref = ref64.symv(uplo, n, al, a, lda, x, incx, be, y, incy, herm=herm)
            assert np.all(np.isfinite(out))
            rtol = 2e-5 if p in "sc" else 1e-12
            err = np.abs(out - ref).max() / (np.abs(ref).max() + 1)
            assert err <= rtol * np.sqrt(n) / 8, (uplo, n, err)


def tri(n, lda, f, rng, uplo, cplx):
    a = np.full(lda * n, np.nan, f)
    for j in range(n):
        rows = range(j, n) if uplo == 1 else range(0, j + 1)
        for i in rows:
            v = rng.uniform(-1, 1) / n + (1j * rng.uniform(-1, 1) / n if cplx else 0)
            a[i + j * lda] = v if i != j else rng.uniform(1, 2) + (0.25j if cplx else 0)
    return a


def tri_vec(n, lda, f, rng, uplo, cplx):
    """Same distribution as tri() (off-diagonal U[-1,1]/n, diagonal U[1,2]), drawn
    in one vectorised pass; NaN outside the stored triangle."""
    v = rng.uniform(-1, 1, (n, n)) / n
    if cplx:
        v = v + 1j * rng.uniform(-1, 1, (n, n)) / n
    d = rng.uniform(1, 2, n) + (0.25j if cplx else 0)
    v[np.arange(n), np.arange(n)] = d
    keep = np.tril(np.ones((n, n), bool)) if uplo == 1 else np.triu(np.ones((n, n), bool))
    a = np.full((n, lda), np.nan, f)  # a[j, i] = A(i, j)
    a[:, :n] = np.where(keep.T, v.T, np.nan)
    return a.reshape(-1)


@pytest.mark.parametrize("p", "sdcz")
def test_trsv_trmv_all_variants(gpu_lib, p):
    cplx = p in "cz"
    f = DT[p]
    rng = np.random.default_rng(22)
    eps = 2.0 ** -24 if p in "sc" else 2.0 ** -53
    for n, lda, incx in ((1, 1, 1), (7, 7, 1), (64, 64, 2), (65, 70, 1), (200, 200, 3), (640, 640, 1)):
        for uplo in (0, 1):
            a = tri(n, lda, f, rng, uplo, cplx)
            for trans in (0, 1, 2):
                for diag in (0, 1):
                    a_use = a.copy()
                    if diag == 0:
                        for j in range(n):
                            a_use[j + j * lda] = np.nan  # unit diagonal is never read
                    x = rand(rng, (n - 1) * incx + 1, f)
                    xs = x.copy()
                    getattr(blas, f"{p}trsv")(uplo, trans, diag, n, a_use, lda, xs, incx)
                    M = ref64.triangular(a_use, n, lda, uplo, trans, diag)
                    xv = ref64.vec(xs, n, incx).astype(M.dtype)
                    bv = ref64.vec(x, n, incx).astype(M.dtype)
                    berr = np.linalg.norm(bv - M @ xv) / (np.linalg.norm(M) * np.linalg.norm(xv) + np.linalg.norm(bv))
                    assert berr <= 4 * n * eps, (n, uplo, trans, diag, berr)
                    ref = ref64.trsv(uplo, trans, diag, n, a_use, lda, x, incx)
                    assert allclose_ref(xs, ref, 1e-4 if p in "sc" else 1e-12, 1e-5 if p in "sc" else 1e-12)
                    # TRMV of the solution returns the right-hand side
                    getattr(blas, f"{p}trmv")(uplo, trans, diag, n, a_use, lda, xs, incx)
                    assert allclose_ref(xs, x, 1e-4 if p in "sc" else 1e-12, 1e-5 if p in "sc" else 1e-12)


@pytest.mark.parametrize("deps", ["0", "1", "2"])
@pytest.mark.parametrize("p", "sdc")
def test_trsv_precomputed_dependency_depths(gpu_lib, p, deps, monkeypatch):
    """Every depth of the precomputed block products (0: none, 1: the block just
    above, 2: the two blocks above) over all uplo/trans/diag variants, with enough
    64-row blocks (n from 449) that the precomputed path is taken, a ragged last
    block, lda > n and a strided vector."""
    monkeypatch.setenv("CORAL_TRSV_DEPS", deps)
    cplx = p in "cz"
    f = DT[p]
    rng = np.random.default_rng(31)
    eps = 2.0 ** -24 if p in "sc" else 2.0 ** -53
    for n, lda, incx in ((449, 449, 1), (1000, 1003, 2)):
        for uplo in (0, 1):
            a = tri_vec(n, lda, f, rng, uplo, cplx)
            for trans in (0, 1, 2):
                for diag in (0, 1):
                    a_use = a.copy()
                    if diag == 0:
                        a_use[np.arange(n) * (lda + 1)] = np.nan
                    x = rand(rng, (n - 1) * incx + 1, f)
                    xs = x.copy()
                    getattr(blas, f"{p}trsv")(uplo, trans, diag, n, a_use, lda, xs, incx)
                    M = ref64.triangular(a_use, n, lda, uplo, trans, diag)
                    xv = ref64.vec(xs, n, incx).astype(M.dtype)
                    bv = ref64.vec(x, n, incx).astype(M.dtype)
                    berr = np.linalg.norm(bv - M @ xv) / (np.linalg.norm(M) * np.linalg.norm(xv) + np.linalg.norm(bv))
                    assert berr <= 4 * n * eps, (deps, n, uplo, trans, diag, berr)
                    ref = ref64.trsv(uplo, trans, diag, n, a_use, lda, x, incx)
                    assert allclose_ref(xs, ref, 1e-4 if p in "sc" else 1e-12, 1e-5 if p in "sc" else 1e-12)
                    # the strided vector's gaps are left alone
                    gaps = np.ones(len(x), bool)
                    gaps[::incx] = False
                    assert np.array_equal(xs[gaps], x[gaps])


def _tri_fast(n, lda, f, rng, uplo, cplx):
    """Stored triangle of U[-1,1] values (complex: both parts), NaN everywhere else
    in the lda x n buffer, so any read outside the triangle shows up."""
    a = np.full((n, lda), np.nan, f)  # a[j, i] = A(i, j)
    v = rand(rng, n * n, f).reshape(n, n)
    keep = np.tril(np.ones((n, n), bool)) if uplo == 0 else np.triu(np.ones((n, n), bool))
    a[:, :n] = np.where(keep, v, np.nan)  # keep[j, i]: row i of column j is stored
    return a.reshape(-1)


@pytest.mark.parametrize("path", ["blocks", "stream"])
@pytest.mark.parametrize("p", "sdcz")
def test_trmv_against_fp64(gpu_lib, p, path, monkeypatch):
    """x := op(A) x for every uplo x trans x diag, ragged n around the streaming
    kernel's row groups (64 x 2 x 16 / element size rows: 512 float, 256 double or
    complex float, 128 complex double) and its four-column groups, padded lda (16-byte
    aligned columns take the unguarded loads, the others the element loads) and strided
    x: elementwise within 2*n*eps*(|op(A)| |x|) of FP64, with NaN in the unstored
    triangle, the padding and (unit) the diagonal; by 64 x 64 blocks (these orders'
    default) and streamed (CORAL_SYMV_STREAM_MIN_N=0)."""
    if path == "stream":
        monkeypatch.setenv("CORAL_SYMV_STREAM_MIN_N", "0")
    cplx = p in "cz"
    f = DT[p]
    rng = np.random.default_rng(24)
    eps = 2.0 ** -24 if p in "sc" else 2.0 ** -53
    for n, lda, incx in ((1, 1, 1), (63, 63, 1), (64, 66, 2), (129, 129, 1), (1000, 1003, 3), (1537, 1537, 1),
                         (2048, 2048, 1), (2501, 2504, 2)):
        for uplo in (0, 1):
            a = _tri_fast(n, lda, f, rng, uplo, cplx)
            for trans in (0, 1, 2):
                for diag in (0, 1):
                    a_use = a.copy()
                    if diag == 0:
                        a_use.reshape(n, lda)[np.arange(n), np.arange(n)] = np.nan  # never read
                    x = rand(rng, (n - 1) * incx + 1, f)
                    xs = x.copy()
                    getattr(blas, f"{p}trmv")(uplo, trans, diag, n, a_use, lda, xs, incx)
                    ref = ref64.trmv(uplo, trans, diag, n, a_use, lda, x, incx)
                    M = ref64.triangular(a_use, n, lda, uplo, trans, diag)
                    scale = np.abs(M) @ np.abs(ref64.vec(x, n, incx).astype(M.dtype))
                    got = ref64.vec(xs, n, incx).astype(M.dtype)
                    want = ref64.vec(ref, n, incx)
                    err = np.abs(got - want)
                    assert np.all(err <= 2 * n * eps * scale + 1e-300), (n, uplo, trans, diag, float(np.max(err / scale)))
                    # elements between the strided entries are untouched
                    mask = np.ones(xs.shape[0], bool)
                    mask[::incx] = False
                    assert np.array_equal(xs[mask], x[mask])


def test_dtrsv_n8192_backward_error(gpu_lib):
    # the DTRSV configuration: lower, non-unit, n = 8192, strictly-lower U[-1,1]/n, diag U[1,2]
    rng = np.random.default_rng(23)
    n = 8192
    A = np.tril(rng.uniform(-1, 1, (n, n)) / n, -1) + np.diag(rng.uniform(1, 2, n))
    a = np.asfortranarray(A).reshape(-1, order="F")
    b = rng.uniform(-1, 1, n)
    x = b.copy()
    blas.dtrsv(1, 0, 1, n, a, n, x, 1)
    berr = np.linalg.norm(b - A @ x) / (np.linalg.norm(A) * np.linalg.norm(x) + np.linalg.norm(b))
    assert berr <= 2.0 ** -53 * 16, berr


def test_trsv_wait_bound_is_reported(gpu_lib):
    # The host puts TRSV's handoff wait in its test mode (a bound of -1: every wait
    # gives up at once and records the error; the kernels still run to the end); the
    # host-pointer form and
    # coral_stream_check must report CORAL_ERR_WAIT_BOUND, and with the default
    # bound restored the same solve is correct and clean.
    from coral_amd._lib import CoralError, call, status_of
    from coral_amd.device import DeviceArray, Stream

    rng = np.random.default_rng(24)
    n = 16384
    A = np.tril(rng.uniform(-1, 1, (n, n)) / n, -1) + np.diag(rng.uniform(1, 2, n))
    a = np.asfortranarray(A).reshape(-1, order="F")
    b = rng.uniform(-1, 1, n)
    call("coral_testing_set_wait_limit", -1)
    try:
        x = b.copy()
        with pytest.raises(CoralError) as e:
            blas.dtrsv(1, 0, 1, n, a, n, x, 1)
        assert e.value.status == 6  # CORAL_ERR_WAIT_BOUND
        dA, dx, s = DeviceArray.from_numpy(a), DeviceArray.from_numpy(b), Stream()
        call("coral_dtrsv", 1, 0, 1, n, dA.ptr, n, dx.ptr, 1, s.handle)
        assert status_of("coral_stream_check", s.handle) == 6
        assert status_of("coral_stream_check", s.handle) == 0
    finally:
        call("coral_testing_set_wait_limit", 0)
    dx.upload(b)
    call("coral_dtrsv", 1, 0, 1, n, dA.ptr, n, dx.ptr, 1, s.handle)
    assert status_of("coral_stream_check", s.handle) == 0
    x = dx.to_numpy()
    s.destroy()
    berr = np.linalg.norm(b - A @ x) / (np.linalg.norm(A) * np.linalg.norm(x) + np.linalg.norm(b))
    assert berr <= 2.0 ** -53 * 16, berr
    x = b.copy()
    blas.dtrsv(1, 0, 1, n, a, n, x, 1)  # the host form is clean again


@pytest.mark.parametrize("p,uplo,trans,n,scale", [
    ("d", 1, 0, 8192, "4/sqrt(n)"), ("d", 0, 1, 8192, "4/sqrt(n)"), ("d", 1, 0, 1024, "1"), ("d", 0, 0, 1024, "1"),
    ("z", 1, 2, 4096, "4/sqrt(n)"), ("s", 1, 0, 4096, "4/sqrt(n)"),
])
def test_trsv_not_diagonally_dominant(gpu_lib, p, uplo, trans, n, scale):
    # Triangles whose off-diagonal rows outweigh the diagonal (U[-1,1] * 4/sqrt(n):
    # |row sum| ~ 2 sqrt(n) against a diagonal in [1, 2]), and, for n = 1024, plain
    # U[-1,1] off the diagonal: condition about 1e18 with |x| near 1e28.  The GPU
    # solves each 64-row diagonal block with its explicitly inverted block; its
    # backward error |b - op(A) x| / (|A| |x| + |b|) is held to that of CORAL's NB = 8
    # substitution (the oracle) on the same system.
    rng = np.random.default_rng(25 + n + uplo + 3 * trans)
    f = {"s": np.float32, "d": np.float64, "z": np.complex128}[p]
    eps = 2.0 ** -24 if p == "s" else 2.0 ** -53
    s = 4.0 / np.sqrt(n) if scale != "1" else 1.0
    off = rng.uniform(-1, 1, (n, n)) * s
    if p == "z":
        off = off + 1j * rng.uniform(-1, 1, (n, n)) * s
    A = (np.tril(off, -1) if uplo == 1 else np.triu(off, 1)) + np.diag(rng.uniform(1, 2, n))
    A = A.astype(f)
    a = np.asfortranarray(A).reshape(-1, order="F")
    b = rng.uniform(-1, 1, n).astype(f)
    if p == "z":
        b = (b + 1j * rng.uniform(-1, 1, n)).astype(f)
    x = b.copy()
    getattr(blas, f"{p}trsv")(uplo, trans, 1, n, a, n, x, 1)
    xo = b.copy()
    getattr(port, f"{p}trsv")(uplo, trans, 1, n, a, n, xo, 1)
    wide = np.complex128 if p == "z" else np.float64
    M = A.astype(wide)
    M = M if trans == 0 else (M.T if trans == 1 else M.conj().T)

    def berr(v):
        v = v.astype(wide)
        return np.linalg.norm(b.astype(wide) - M @ v) / (np.linalg.norm(M) * np.linalg.norm(v) + np.linalg.norm(b))

    assert np.all(np.isfinite(x))
    bg, bo = berr(x), berr(xo)
    assert bg <= 16 * max(bo, eps), (bg, bo)
