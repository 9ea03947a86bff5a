"""GPU parity for the level-3 routines, through the C ABI (libcoralhip.so).

Checked against the oracle (CORAL restatement, oracle/coral_port.c) with the
reference tests' tolerances, and against FP64 with the north-star normwise bound
|C - C_ref|_F / (k * eps * |A|_F * |B|_F) <= 1.  The split-precision forms carry
the BF16 hi+lo representation error (about 2^-17 per product), so their bound
uses eps_split = 2^-16 + k * eps32 in place of k * eps.
"""
import json
import os

import numpy as np
import pytest

from coral_amd import blas
from oracle import port, ref64
from tests.conftest import allclose_ref

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
GEMM_TOL = {"s": (1e-3, 1e-3), "d": (1e-12, 1e-12), "c": (3e-3, 2e-3), "z": (3e-11, 2e-11)}
EPS = {"s": 2.0 ** -24, "d": 2.0 ** -53, "c": 2.0 ** -24, "z": 2.0 ** -53}
GPU_FN = {"s": blas.sgemm, "d": blas.dgemm, "c": blas.cgemm, "z": blas.zgemm}
PORT_FN = {"s": port.sgemm, "d": port.dgemm, "c": port.cgemm, "z": port.zgemm}
DT = {"s": np.float32, "d": np.float64, "c": np.complex64, "z": np.complex128}


def normwise(out, ref, a_fro, b_fro, k, eps):
    return np.linalg.norm(np.asarray(out).astype(ref.dtype) - ref) / (max(k, 1) * eps * a_fro * b_fro + 1e-300)


def rand(rng, n, f):
    if np.issubdtype(f, np.complexfloating):
        return (rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)).astype(f)
    return rng.uniform(-1, 1, n).astype(f)


def test_kat(gpu_lib):
    kat = json.load(open(os.path.join(GOLD, "kat.json")))
    k = kat["sgemm_2x2"]
    for fn, f in ((blas.sgemm, np.float32), (blas.dgemm, np.float64), (blas.sgemm_bf16x3, np.float32),
                  (blas.sgemm_fp16x3, np.float32)):
        c = np.array(k["c"], f)
        fn(0, 0, 2, 2, 2, k["alpha"], np.array(k["a"], f), 2, np.array(k["b"], f), 2, k["beta"], c, 2)
        assert np.all(np.abs(c - np.array(k["expect"])) < 1e-6), fn.__name__
    k = kat["cgemm_2x2"]
    for fn, f in ((blas.cgemm, np.float32), (blas.zgemm, np.float64), (blas.cgemm_3m_bf16x3, np.float32),
                  (blas.cgemm_3m_fp16x3, np.float32)):
        c = np.array(k["c"], f)
        fn(0, 0, 2, 2, 2, k["alpha"], np.array(k["a"], f), 2, np.array(k["b"], f), 2, k["beta"], c, 2)
        assert np.all(np.abs(c - np.array(k["expect"])) < 1e-6), fn.__name__


def test_golden_reference_generators(gpu_lib):
    z = np.load(os.path.join(GOLD, "gemm_cases.npz"))
    index = json.loads(str(z["index"]))
    for key, dt, m, n, k, oa, ob, lda, ldb, ldc, alpha, beta in index:
        a, b, c, expect = z[key + "_a"], z[key + "_b"], z[key + "_c"], z[key + "_expect"]
        out = c.copy()
        GPU_FN[dt](oa, ob, m, n, k, alpha, a, lda, b, ldb, beta, out, ldc)
        assert allclose_ref(out, expect, *GEMM_TOL[dt]), key
        # ld padding between columns is untouched
        for j in range(n):
            assert np.array_equal(out[j * ldc + m:(j + 1) * ldc], c[j * ldc + m:(j + 1) * ldc]), key
        po = c.copy()
        PORT_FN[dt](oa, ob, m, n, k, alpha, a, lda, b, ldb, beta, po, ldc)
        assert allclose_ref(out, po, *GEMM_TOL[dt]), key


def test_golden_larger_reference_shapes(gpu_lib):
    # the reference tests' 128^3 and rectangular 64x192x15 / 192x64x15 runs, inputs
    # from their generators, against the committed signatures and the full FP64
    # product at the reference tolerances, and against the oracle
    from tests.test_oracle import large_reference_cases

    count = 0
    for key, dt, (m, n, k, oa, ob, lda, ldb, ldc, alpha, beta), (a, b, c), out, pos, sample, fro in \
            large_reference_cases():
        got = c.copy()
        GPU_FN[dt](oa, ob, m, n, k, alpha, a, lda, b, ldb, beta, got, ldc)
        rtol, atol = GEMM_TOL[dt]
        assert allclose_ref(got[pos], sample, rtol, atol), key
        assert allclose_ref(got, out, rtol, atol), key
        po = c.copy()
        PORT_FN[dt](oa, ob, m, n, k, alpha, a, lda, b, ldb, beta, po, ldc)
        assert allclose_ref(got, po, rtol, atol), key
        count += 1
    assert count == 360


SHAPES = [(1, 1, 1), (5, 7, 4), (64, 64, 64), (127, 129, 65), (300, 257, 129), (17, 513, 40), (512, 16, 777)]


@pytest.mark.parametrize("dt", ["s", "d", "c", "z"])
def test_normwise_bound_all_ops(gpu_lib, dt):
    rng = np.random.default_rng(10)
    f = DT[dt]
    ops = [0, 1, 2]
    worst = 0.0
    for (m, n, k) in SHAPES:
        for oa in ops:
            for ob in ops:
                lda = (m if oa == 0 else k) + 3
                ldb = (k if ob == 0 else n) + 1
                ldc = m + 2
                ar, ac = (m, k) if oa == 0 else (k, m)
                br, bc = (k, n) if ob == 0 else (n, k)
                a = rand(rng, lda * ac, f)
                b = rand(rng, ldb * bc, f)
                c = rand(rng, ldc * n, f)
                cplx = dt in "cz"
                al, be = ((1.25, -0.5), (0.5, 0.25)) if cplx else (1.25, 0.5)
                out = c.copy()
                GPU_FN[dt](oa, ob, m, n, k, al, a, lda, b, ldb, be, out, ldc)
                ref = ref64.gemm(oa, ob, m, n, k, al, a, lda, b, ldb, be, c, ldc)
                A = ref64.mat(a, ar, ac, lda)
                B = ref64.mat(b, br, bc, ldb)
                C = ref64.mat(c, m, n, ldc)
                scale = abs(complex(*al)) if cplx else abs(al)
                num = np.linalg.norm(ref64.mat(out, m, n, ldc).astype(ref.dtype) - ref64.mat(ref, m, n, ldc))
                den = max(k, 1) * EPS[dt] * (scale * np.linalg.norm(A) * np.linalg.norm(B) + np.linalg.norm(C))
                worst = max(worst, num / den)
    assert worst <= 1.0, worst


# representation error of one split operand element: hi + lo keeps 16 (BF16) or
# 22 (FP16) significand bits, so |x - hi - lo| <= 2^-17 |x| (BF16) / 2^-22 |x| (FP16);
# the dropped lo*lo term is of the same order.
SPLIT_EPS = {"bf16x3": 2.0 ** -16, "fp16x3": 2.0 ** -21}


@pytest.mark.parametrize("which", ["sgemm_bf16x3", "cgemm_3m_bf16x3", "sgemm_fp16x3", "cgemm_3m_fp16x3"])
def test_split_precision_bound(gpu_lib, which):
    rng = np.random.default_rng(11)
    cplx = which.startswith("c")
    f = np.complex64 if cplx else np.float32
    fn = getattr(blas, which)
    ops = [0, 1, 2] if cplx else [0, 1]
    worst = 0.0
    for (m, n, k) in SHAPES:
        for oa in ops:
            for ob in ops:
                lda, ldb, ldc = (m if oa == 0 else k), (k if ob == 0 else n), m
                ar, ac = (m, k) if oa == 0 else (k, m)
                br, bc = (k, n) if ob == 0 else (n, k)
                a, b, c = rand(rng, lda * ac, f), rand(rng, ldb * bc, f), rand(rng, ldc * n, f)
                al, be = ((1.0, 0.5), (0.5, -0.25)) if cplx else (1.0, 0.5)
                out = c.copy()
                fn(oa, ob, m, n, k, al, a, lda, b, ldb, be, out, ldc)
                ref = ref64.gemm(oa, ob, m, n, k, al, a, lda, b, ldb, be, c, ldc)
                A, B = ref64.mat(a, ar, ac, lda), ref64.mat(b, br, bc, ldb)
                eps_split = SPLIT_EPS[which[-6:]] + k * 2.0 ** -24
                # 3M adds the Gauss cancellation: Im = T3 - T1 - T2 (three products' errors)
                mult = 3.0 if cplx else 1.0
                num = np.linalg.norm(out.astype(ref.dtype) - ref)
                den = mult * eps_split * (abs(complex(*al)) if cplx else al) * np.linalg.norm(A) * np.linalg.norm(B)
                worst = max(worst, num / (den + np.linalg.norm(c) * 2.0 ** -24))
    assert worst <= 1.0, worst


@pytest.mark.parametrize("which,bound", [("sgemm_bf16x3", 2.0 ** -14), ("sgemm_fp16x3", 2.0 ** -17)])
def test_split_precision_wide_range_rows(gpu_lib, which, bound):
    # rows of A and columns of B scaled by 2^e, e in [-20, 20]: the power-of-two
    # row/column scaling keeps the relative error of every element at split level
    # (for FP16 halves it also keeps every scaled value inside FP16's range)
    rng = np.random.default_rng(12)
    m = n = k = 256
    a = rng.standard_normal((m, k)).astype(np.float32)
    b = rng.standard_normal((k, n)).astype(np.float32)
    ea = rng.integers(-20, 21, m)
    eb = rng.integers(-20, 21, n)
    a = (a * np.ldexp(1.0, ea)[:, None]).astype(np.float32)
    b = (b * np.ldexp(1.0, eb)[None, :]).astype(np.float32)
    af, bf = np.asfortranarray(a).reshape(-1, order="F"), np.asfortranarray(b).reshape(-1, order="F")
    out = np.zeros(m * n, np.float32)
    getattr(blas, which)(0, 0, m, n, k, 1.0, af, m, bf, k, 0.0, out, m)
    ref = a.astype(np.float64) @ b.astype(np.float64)
    scale = np.ldexp(1.0, ea)[:, None] * np.ldexp(1.0, eb)[None, :]
    rel = np.abs(out.reshape(n, m).T - ref) / (scale * np.sqrt(k))
    assert rel.max() < bound, rel.max()


@pytest.mark.parametrize("ops", [(0, 0), (1, 1)])
def test_bf16x3_scaled_and_unscaled_packing_agree_bitwise(gpu_lib, ops):
    # BF16 halves are packed unscaled unless some row of op(A) or column of op(B) has
    # its maximum outside [2^-40, 2^40] (or non-finite); then every row and column is
    # re-packed scaled by a power of two and the epilogue scales back.  Power-of-two
    # scaling commutes with every rounding while values stay normal, so scaling an
    # operand by 2^60 or 2^-60 (which forces the scaled path) changes the result by
    # exactly that factor.  NN packs A rows-contiguous and B k-contiguous; TT the
    # other way round.
    rng = np.random.default_rng(21)
    oa, ob = ops
    m, n, k = 333, 290, 203
    lda, ldb = (m if oa == 0 else k), (k if ob == 0 else n)
    ac, bc = (k if oa == 0 else m), (n if ob == 0 else k)
    a = rng.standard_normal(lda * ac).astype(np.float32)
    b = rng.standard_normal(ldb * bc).astype(np.float32)

    def run(x, y):
        out = np.zeros(m * n, np.float32)
        blas.sgemm_bf16x3(oa, ob, m, n, k, 1.0, x, lda, y, ldb, 0.0, out, m)
        return out

    base = run(a, b)
    big, tiny = np.float32(2.0 ** 60), np.float32(2.0 ** -60)
    assert np.array_equal(run(a * big, b) * tiny, base)
    assert np.array_equal(run(a, b * tiny) * big, base)
    # one row of op(A) alone out of range: that row scales exactly, the rest are unchanged
    a1 = a.copy()
    row = 17
    if oa == 0:
        a1[row::lda] *= big
    else:
        a1[row * lda:row * lda + k] *= big
    got = run(a1, b).reshape(n, m)
    ref = base.reshape(n, m)
    assert np.array_equal(got[:, row] * tiny, ref[:, row])
    keep = np.arange(m) != row
    assert np.array_equal(got[:, keep], ref[:, keep])
    # and the result is the FP64 product to split accuracy
    A = ref64.mat(a, m if oa == 0 else k, ac, lda)
    B = ref64.mat(b, k if ob == 0 else n, bc, ldb)
    ref64_c = (A.T if oa else A) @ (B.T if ob else B)
    rel = np.abs(base.reshape(n, m).T - ref64_c).max() / (np.sqrt(k) * 2.0 ** -14)
    assert rel < 1.0, rel


@pytest.mark.parametrize("ops", [(0, 0), (2, 2), (1, 2)])
def test_cgemm_3m_bf16x3_scaled_and_unscaled_packing_agree_bitwise(gpu_lib, ops):
    # the 3M form packs the planes Re, Im and Re + Im of each operand the same way:
    # unscaled unless some plane's row maximum is out of range; scaling A or B by
    # 2^+-60 forces the scaled path and changes the result by exactly that factor
    rng = np.random.default_rng(22)
    oa, ob = ops
    m, n, k = 300, 266, 197
    lda, ldb = (m if oa == 0 else k), (k if ob == 0 else n)
    ac, bc = (k if oa == 0 else m), (n if ob == 0 else k)
    a = (rng.standard_normal(lda * ac) + 1j * rng.standard_normal(lda * ac)).astype(np.complex64)
    b = (rng.standard_normal(ldb * bc) + 1j * rng.standard_normal(ldb * bc)).astype(np.complex64)

    def run(x, y):
        out = np.zeros(m * n, np.complex64)
        blas.cgemm_3m_bf16x3(oa, ob, m, n, k, (1.0, 0.0), x, lda, y, ldb, (0.0, 0.0), out, m)
        return out

    base = run(a, b)
    big, tiny = np.float32(2.0 ** 60), np.float32(2.0 ** -60)
    assert np.array_equal(run(a * big, b) * tiny, base)
    assert np.array_equal(run(a, b * tiny) * big, base)
    A = ref64.mat(a, m if oa == 0 else k, ac, lda)
    B = ref64.mat(b, k if ob == 0 else n, bc, ldb)
    opa = A if oa == 0 else (A.T if oa == 1 else A.conj().T)
    opb = B if ob == 0 else (B.T if ob == 1 else B.conj().T)
    rel = np.abs(base.reshape(n, m).T - opa @ opb).max() / (np.sqrt(k) * 2.0 ** -12)
    assert rel < 1.0, rel


def test_fp16x3_error_matches_plain_fp32(gpu_lib):
    # the FP16 split keeps 22 of FP32's 24 significand bits: its normwise error
    # against the FP64 product stays at the plain FP32-MFMA level (N(0,1) operands)
    rng = np.random.default_rng(14)
    m = n = k = 1024
    a = rng.standard_normal(m * k).astype(np.float32)
    b = rng.standard_normal(k * n).astype(np.float32)
    ref = ref64.gemm(0, 0, m, n, k, 1.0, a, m, b, k, 0.0, np.zeros(m * n, np.float32), m)
    errs = {}
    for name in ("sgemm", "sgemm_fp16x3", "sgemm_bf16x3"):
        out = np.zeros(m * n, np.float32)
        getattr(blas, name)(0, 0, m, n, k, 1.0, a, m, b, k, 0.0, out, m)
        errs[name] = np.linalg.norm(out - ref) / np.linalg.norm(ref)
    assert errs["sgemm_fp16x3"] <= 1.5 * errs["sgemm"], errs
    assert errs["sgemm_bf16x3"] > errs["sgemm_fp16x3"], errs


@pytest.mark.parametrize("dt", ["s", "d", "c", "z"])
def test_beta_zero_never_reads_c_and_quick_returns(gpu_lib, dt):
    rng = np.random.default_rng(13)
    f = DT[dt]
    cplx = dt in "cz"
    m, n, k = 33, 47, 21
    a, b = rand(rng, m * k, f), rand(rng, k * n, f)
    one, zero, half = ((1.0, 0.0), (0.0, 0.0), (0.5, 0.0)) if cplx else (1.0, 0.0, 0.5)
    c = np.full(m * n, np.nan, f)
    GPU_FN[dt](0, 0, m, n, k, one, a, m, b, k, zero, c, m)
    assert np.all(np.isfinite(c))
    # alpha == 0: C := beta*C with A and B unread (NaN operands)
    c0 = rand(rng, m * n, f)
    c1 = c0.copy()
    GPU_FN[dt](0, 0, m, n, k, zero, np.full(m * k, np.nan, f), m, np.full(k * n, np.nan, f), k, half, c1, m)
    assert np.allclose(c1, c0 * 0.5, rtol=0, atol=0)
    # k == 0 with beta == 0 zeroes C
    c2 = np.full(m * n, np.nan, f)
    GPU_FN[dt](0, 0, m, n, 0, one, a, m, b, 1, zero, c2, m)
    assert np.all(c2 == 0)


def test_real_conjtranspose_equals_transpose(gpu_lib):
    rng = np.random.default_rng(14)
    m, n, k = 40, 30, 50
    a = rng.uniform(-1, 1, k * m).astype(np.float32)
    b = rng.uniform(-1, 1, k * n).astype(np.float32)
    c1, c2 = np.zeros(m * n, np.float32), np.zeros(m * n, np.float32)
    blas.sgemm(1, 0, m, n, k, 1.0, a, k, b, k, 0.0, c1, m)
    blas.sgemm(2, 0, m, n, k, 1.0, a, k, b, k, 0.0, c2, m)
    assert np.array_equal(c1, c2)


def test_argument_errors_reported(gpu_lib):
    from coral_amd import CoralError

    a = np.zeros(16, np.float32)
    with pytest.raises(CoralError) as e:
        blas.sgemm(0, 0, 4, 4, 4, 1.0, a, 3, a, 4, 0.0, a, 4)  # lda < m
    assert e.value.status == -8
    with pytest.raises(CoralError) as e:
        blas.sgemm(7, 0, 4, 4, 4, 1.0, a, 4, a, 4, 0.0, a, 4)
    assert e.value.status == -1


# Problems with fewer 64x64 tiles than CUs take the one-launch small kernel: 16-byte
# vector loads when the operand base is 16-byte aligned and its leading dimension a
# multiple of the vector (the blocks wholly inside the operand), element loads from
# clamped indices otherwise (edge blocks, a misaligned base).  Shapes mix full and
# edge 32x32 tiles and full and partial 32- / 64-deep k steps.
SMALL_SHAPES = [(64, 96, 200), (96, 33, 64), (40, 64, 130), (32, 32, 32), (128, 64, 65), (33, 31, 7)]


@pytest.mark.parametrize("dt", ["s", "d"])
def test_small_kernel_vector_and_element_paths(gpu_lib, dt):
    from coral_amd import device
    from coral_amd._lib import call

    rng = np.random.default_rng(23)
    f = DT[dt]
    name = "coral_sgemm" if dt == "s" else "coral_dgemm"
    worst = 0.0
    for (m, n, k) in SMALL_SHAPES:
        for oa in (0, 1):
            for ob in (0, 1):
                ar, ac = (m, k) if oa == 0 else (k, m)
                br, bc = (k, n) if ob == 0 else (n, k)
                lda, ldb, ldc = -(-ar // 4) * 4 + 4, -(-br // 4) * 4, m
                a, b, c = rand(rng, lda * ac, f), rand(rng, ldb * bc, f), rand(rng, ldc * n, f)
                ref = ref64.gemm(oa, ob, m, n, k, 0.75, a, lda, b, ldb, -0.5, c, ldc)
                po = c.copy()
                PORT_FN[dt](oa, ob, m, n, k, 0.75, a, lda, b, ldb, -0.5, po, ldc)
                den = max(k, 1) * EPS[dt] * (0.75 * np.linalg.norm(ref64.mat(a, ar, ac, lda)) *
                                            np.linalg.norm(ref64.mat(b, br, bc, ldb)) +
                                            0.5 * np.linalg.norm(ref64.mat(c, m, n, ldc)))
                for shift in (0, 1):  # element offset of every operand's base: 0 aligned, 1 not
                    A, B, C = (device.DeviceArray.from_numpy(np.concatenate([np.zeros(shift, f), x]))
                               for x in (a, b, c))
                    es = np.dtype(f).itemsize
                    call(name, oa, ob, m, n, k, 0.75, A.ptr + shift * es, lda, B.ptr + shift * es, ldb, -0.5,
                         C.ptr + shift * es, ldc, None)
                    out = C.to_numpy()[shift:]
                    num = np.linalg.norm(ref64.mat(out, m, n, ldc).astype(np.float64) - ref64.mat(ref, m, n, ldc))
                    worst = max(worst, num / den)
                    assert allclose_ref(out, po, *GEMM_TOL[dt]), (m, n, k, oa, ob, shift)
    assert worst <= 1.0, worst


@pytest.mark.parametrize("which", ["sgemm", "sgemm_bf16x3", "sgemm_fp16x3"])
def test_non_finite_operands_stay_in_their_row_and_column(gpu_lib, which):
    """Inf in A(5, 7) and NaN in B(3, 9): row 5 and column 9 of C are not finite
    (plain FP32 gives +-Inf in row 5; the split forms may give NaN there, since a cross
    product multiplies the Inf by a lo half that can be zero), every other element is
    finite and within the form's normwise bound of FP64.  The row and column scaling of
    the split forms must not carry the special value into other rows or columns."""
    rng = np.random.default_rng(31)
    m = n = k = 160
    a = rng.uniform(-1, 1, m * k).astype(np.float32)
    b = rng.uniform(-1, 1, k * n).astype(np.float32)
    a[5 + 7 * m] = np.inf
    b[3 + 9 * k] = np.nan
    c = np.zeros(m * n, np.float32)
    getattr(blas, which)(0, 0, m, n, k, 1.0, a, m, b, k, 0.0, c, m)
    C = c.reshape(n, m).T
    assert not np.isfinite(C[5, :]).any() and np.isnan(C[:, 9]).all()
    rows, cols = np.arange(m) != 5, np.arange(n) != 9
    A = a.reshape(k, m).T.astype(np.float64)[rows]
    B = b.reshape(n, k).T.astype(np.float64)[:, cols]
    got = C[rows][:, cols].astype(np.float64)
    assert np.isfinite(got).all()
    eps = {"sgemm": k * 2.0 ** -24, "sgemm_bf16x3": 2.0 ** -16 + k * 2.0 ** -24,
           "sgemm_fp16x3": 2.0 ** -21 + k * 2.0 ** -24}[which]
    assert np.linalg.norm(got - A @ B) <= eps * np.linalg.norm(A) * np.linalg.norm(B)
