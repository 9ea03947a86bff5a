"""The five benchmark configurations at their full sizes, through the C ABI.

C1  SGEMM NN 512^3, alpha = 1, beta = 0.5, U[-1,1]: against the oracle (CORAL's
    algorithm restated) at the reference tests' tolerance and against FP64.
C2  DGEMM NN and TT 8192^3 N(0,1); DGEMV N 16384^2 U[-1,1].
C3  split SGEMM NN at 4096, 8192 and 16384 for N(0,1) and for the wide range
    (rows of A and columns of B scaled by 2^e, e uniform in [-20, 20]): plain FP32
    MFMA, BF16x3 and FP16x3 on the same operands.
C4  10,000 SGEMM and 10,000 DGEMM NN problems with m, n, k log-uniform integers in
    [8, 512] in one batched call (eager and hipGraph replay); 4,096 batched DTRSV
    lower non-unit with n log-uniform in [32, 256].
C5  CGEMM ConjTrans/ConjTrans 4096^3 (4M and both 3M split forms), ZGEMM NN
    4096^3, CHEMV upper 16384 on (M + M^H) / 2.

GEMM results are checked with the normwise bound
    |C - C_ref|_F <= k * eps * (|alpha| |A|_F |B|_F + |beta| |C0|_F)
(eps_split = 2^-16 + k 2^-24 for BF16x3 and 2^-21 + k 2^-24 for FP16x3 in place of
k * eps).  For the largest products the reference is FP64 on a row sample (the
Frobenius norm of the error is estimated from the sampled rows); for FP64 results
that reference is numpy's float64 product, whose own error is within the same
bound, so the check separates correct results from wrong ones by many orders of
magnitude but does not resolve the last bits.
"""
import ctypes

import numpy as np
import pytest

from coral_amd import device, synthetic
from coral_amd._lib import call
from oracle import port
from tests.conftest import allclose_ref

pytestmark = pytest.mark.gpu

E32, E64 = 2.0 ** -24, 2.0 ** -53


def host_mat(d, rows, cols, dtype):
    return d.to_numpy().view(dtype).reshape(cols, rows).T


class RowSample:
    """FP64 (complex128) reference rows of op(A) op(B), with op(X) = X, X^T or X^H,
    and the Frobenius norms of A and B, computed once for several results."""

    def __init__(self, A, B, rows, ops=(0, 0), wide=np.float64):
        self.rows, self.wide = rows, wide
        opA = A if ops[0] == 0 else (A.T if ops[0] == 1 else A.conj().T)
        opB = B if ops[1] == 0 else (B.T if ops[1] == 1 else B.conj().T)
        self.ref = opA[rows, :].astype(wide) @ opB.astype(wide)
        self.fro = float(np.linalg.norm(A.astype(wide)) * np.linalg.norm(B.astype(wide)))

    def normwise(self, C, eps):
        """Estimate of |C - op(A) op(B)|_F / (eps |A|_F |B|_F) from the sampled rows."""
        err = np.linalg.norm(C[self.rows, :].astype(self.wide) - self.ref) * np.sqrt(C.shape[0] / len(self.rows))
        return float(err / (eps * self.fro))


# ---------------------------------------------------------------------------- C1
def test_c1_sgemm_nn_512_alpha1_beta05(gpu_lib):
    n = 512
    rng = np.random.default_rng(100)
    a, b, c = (rng.uniform(-1, 1, n * n).astype(np.float32) for _ in range(3))
    A, B, C = (device.DeviceArray.from_numpy(x) for x in (a, b, c))
    call("coral_sgemm", 0, 0, n, n, n, 1.0, A.ptr, n, B.ptr, n, 0.5, C.ptr, n, None)
    out = C.to_numpy()
    po = c.copy()
    port.sgemm(0, 0, n, n, n, 1.0, a, n, b, n, 0.5, po, n)
    # the reference tests' SGEMM tolerance (coral-aarch64/tests/level3/sgemm_tests.rs:64-81)
    assert allclose_ref(out, po, 1e-3, 1e-3)
    Am, Bm, Cm = (x.reshape(n, n).T.astype(np.float64) for x in (a, b, c))
    ref = Am @ Bm + 0.5 * Cm
    bound = n * E32 * (np.linalg.norm(Am) * np.linalg.norm(Bm) + 0.5 * np.linalg.norm(Cm))
    assert np.linalg.norm(out.reshape(n, n).T - ref) <= bound
    assert np.linalg.norm(po.reshape(n, n).T - ref) <= bound  # the oracle meets it too


# ---------------------------------------------------------------------------- C2
@pytest.mark.parametrize("ops", [(0, 0), (1, 1)], ids=["NN", "TT"])
def test_c2_dgemm_8192(gpu_lib, ops):
    n = 8192
    A, B, C = (device.DeviceArray(n * n, np.float64) for _ in range(3))
    device.fill_random(A, synthetic.seed_for(2, "A"), "normal")
    device.fill_random(B, synthetic.seed_for(2, "B"), "normal")
    device.fill_random(C, synthetic.seed_for(2, "C"), "uniform")  # beta = 0: C is not read
    call("coral_dgemm", ops[0], ops[1], n, n, n, 1.0, A.ptr, n, B.ptr, n, 0.0, C.ptr, n, None)
    Ah, Bh, Ch = (host_mat(x, n, n, np.float64) for x in (A, B, C))
    rows = np.random.default_rng(2).choice(n, 48, replace=False)
    r = RowSample(Ah, Bh, rows, ops).normwise(Ch, n * E64)
    assert r <= 1.0, r
    assert np.isfinite(Ch).all()


def test_c2_dgemv_n_16384(gpu_lib):
    m = 16384
    A, x, y = device.DeviceArray(m * m, np.float64), device.DeviceArray(m, np.float64), \
        device.DeviceArray(m, np.float64)
    for d, o in ((A, "A"), (x, "x"), (y, "y")):
        device.fill_random(d, synthetic.seed_for(2, o), "uniform")
    y0 = y.to_numpy()
    Ah, xh = host_mat(A, m, m, np.float64), x.to_numpy()
    ax, absax = Ah @ xh, np.abs(Ah) @ np.abs(xh)
    del Ah
    for beta in (0.0, 0.5):
        y.upload(y0)
        call("coral_dgemv", 0, m, m, 1.0, A.ptr, m, x.ptr, 1, beta, y.ptr, 1, None)
        ref = ax + beta * y0
        scale = absax + abs(beta) * np.abs(y0)
        err = np.abs(y.to_numpy() - ref)
        assert np.all(err <= 2 * m * E64 * scale), float(np.max(err / scale))


# ---------------------------------------------------------------------------- C3
@pytest.mark.parametrize("n", [4096, 8192, 16384])
def test_c3_split_sgemm(gpu_lib, n):
    A, B = device.DeviceArray(n * n, np.float32), device.DeviceArray(n * n, np.float32)
    outs = {f: device.DeviceArray(n * n, np.float32) for f in ("plain", "bf16x3", "fp16x3")}
    fns = {"plain": "coral_sgemm", "bf16x3": "coral_sgemm_bf16x3", "fp16x3": "coral_sgemm_fp16x3"}
    eps = {"plain": n * E32, "bf16x3": 2.0 ** -16 + n * E32, "fp16x3": 2.0 ** -21 + n * E32}
    rows = np.random.default_rng(5).choice(n, 64 if n >= 16384 else 96, replace=False)
    report = {}
    for dist in ("normal", "wide"):
        device.fill_random(A, synthetic.seed_for(3, "A") + n, "normal")
        device.fill_random(B, synthetic.seed_for(3, "B") + n, "normal")
        if dist == "wide":
            rng = np.random.default_rng(n)
            ea = device.DeviceArray.from_numpy(rng.integers(-20, 21, n).astype(np.int32))
            eb = device.DeviceArray.from_numpy(rng.integers(-20, 21, n).astype(np.int32))
            call("coral_scale_pow2_f32", A.ptr, n, n, n, ea.ptr, None, None)
            call("coral_scale_pow2_f32", B.ptr, n, n, n, None, eb.ptr, None)
        for f, out in outs.items():
            call(fns[f], 0, 0, n, n, n, 1.0, A.ptr, n, B.ptr, n, 0.0, out.ptr, n, None)
        sample = RowSample(host_mat(A, n, n, np.float32), host_mat(B, n, n, np.float32), rows)
        r = {f: sample.normwise(host_mat(o, n, n, np.float32), 1.0) for f, o in outs.items()}
        report[dist] = r
        for f in outs:
            assert r[f] <= eps[f], (dist, f, r[f] / eps[f])
        # the metric's condition: FP16x3 no worse than plain FP32 MFMA against FP64;
        # BF16x3's 16-bit representation keeps it above both
        assert r["fp16x3"] <= r["plain"], (dist, r)
        assert r["bf16x3"] > r["fp16x3"], (dist, r)
        del sample
    print(f"C3 n={n} normwise |E|_F/(|A|_F|B|_F): {report}")


# ---------------------------------------------------------------------------- C4
def _log_uniform(rng, count, lo, hi, cols=3):
    return np.exp(rng.uniform(np.log(lo), np.log(hi + 1), (count, cols))).astype(np.int64).clip(lo, hi)


@pytest.mark.parametrize("dt", [np.float32, np.float64], ids=["sgemm", "dgemm"])
def test_c4_batched_10000_log_uniform(gpu_lib, dt):
    P = 10000
    rng = np.random.default_rng(4)
    dims = _log_uniform(rng, P, 8, 512)
    m, n, k = dims[:, 0], dims[:, 1], dims[:, 2]
    sa, sb, sc = m * k, k * n, m * n
    oa, ob, oc = (np.concatenate([[0], np.cumsum(s)[:-1]]) for s in (sa, sb, sc))
    A, B, C = (device.DeviceArray(int(s.sum()), dt) for s in (sa, sb, sc))
    device.fill_random(A, 41, "normal")
    device.fill_random(B, 42, "normal")
    es = np.dtype(dt).itemsize
    PA, PB, PC = (device.DeviceArray.from_numpy((base.ptr + es * o).astype(np.uint64))
                  for base, o in ((A, oa), (B, ob), (C, oc)))
    M, N, K = (device.DeviceArray.from_numpy(d.astype(np.int32)) for d in (m, n, k))
    name = "coral_sgemm_batched" if dt == np.float32 else "coral_dgemm_batched"
    stream = device.Stream()
    batched = lambda: call(name, 0, 0, P, M.ptr, N.ptr, K.ptr, 1.0, PA.ptr, M.ptr, PB.ptr, K.ptr, 0.0, PC.ptr,
                           M.ptr, stream.handle)
    batched()
    stream.synchronize()
    eager = C.to_numpy()
    # every problem against FP64 with the normwise bound
    ah, bh = A.to_numpy(), B.to_numpy()
    eps = E32 if dt == np.float32 else E64
    worst = 0.0
    for i in range(P):
        mi, ni, ki = int(m[i]), int(n[i]), int(k[i])
        Ai = ah[oa[i]:oa[i] + sa[i]].reshape(ki, mi).T.astype(np.float64)
        Bi = bh[ob[i]:ob[i] + sb[i]].reshape(ni, ki).T.astype(np.float64)
        Ci = eager[oc[i]:oc[i] + sc[i]].reshape(ni, mi).T
        r = np.linalg.norm(Ci - Ai @ Bi) / (ki * eps * np.linalg.norm(Ai) * np.linalg.norm(Bi))
        worst = max(worst, r)
        assert r <= 1.0, (i, mi, ni, ki, r)
    # a captured graph replays the same work: bit-identical to the eager call, with
    # eager work on the stream between replays
    g = ctypes.c_void_p()
    call("coral_graph_begin_capture", stream.handle)
    batched()
    call("coral_graph_end_capture", stream.handle, ctypes.byref(g))
    try:
        for _ in range(3):
            call("coral_memset", C.ptr, 0, C.nbytes)
            call("coral_graph_launch", g.value, stream.handle)
            stream.synchronize()
            assert np.array_equal(C.to_numpy(), eager)
            call("coral_sgemm" if dt == np.float32 else "coral_dgemm", 0, 0, 64, 64, 64, 1.0, A.ptr, 64, B.ptr,
                 64, 0.0, PC_scratch(dt).ptr, 64, stream.handle)
    finally:
        call("coral_graph_destroy", g.value)
        stream.destroy()
    print(f"C4 {np.dtype(dt).name}: worst normwise ratio {worst:.3e}")


_scratch = {}


def PC_scratch(dt):
    if dt not in _scratch:
        _scratch[dt] = device.DeviceArray(64 * 64, dt)
    return _scratch[dt]


def test_c4_batched_dtrsv_4096(gpu_lib):
    rng = np.random.default_rng(44)
    ns = np.exp(rng.uniform(np.log(32), np.log(257), 4096)).astype(np.int64).clip(32, 256)
    mats, xs = [], []
    for nn in ns:
        Ah = np.diag(rng.uniform(1, 2, nn)) + np.tril(rng.uniform(-1, 1, (nn, nn)) / nn, -1)
        mats.append(np.asfortranarray(Ah).reshape(-1, order="F"))
        xs.append(rng.uniform(-1, 1, nn))
    oa = np.concatenate([[0], np.cumsum([a.size for a in mats])[:-1]])
    ox = np.concatenate([[0], np.cumsum([x.size for x in xs])[:-1]])
    DA = device.DeviceArray.from_numpy(np.concatenate(mats))
    DX = device.DeviceArray.from_numpy(np.concatenate(xs))
    PA = device.DeviceArray.from_numpy((DA.ptr + 8 * oa).astype(np.uint64))
    PX = device.DeviceArray.from_numpy((DX.ptr + 8 * ox).astype(np.uint64))
    NS = device.DeviceArray.from_numpy(ns.astype(np.int32))
    call("coral_dtrsv_batched", 1, 0, 1, len(ns), NS.ptr, PA.ptr, NS.ptr, PX.ptr, None)
    out = DX.to_numpy()
    for i, nn in enumerate(ns):
        A = mats[i].reshape(nn, nn).T
        x = out[ox[i]:ox[i] + nn]
        b = xs[i]
        berr = np.linalg.norm(b - A @ x) / (np.linalg.norm(A) * np.linalg.norm(x) + np.linalg.norm(b))
        assert berr <= 4 * nn * E64, (i, nn, berr)


# ---------------------------------------------------------------------------- C5
@pytest.mark.parametrize("fn,eps", [("coral_cgemm", None), ("coral_cgemm_3m_bf16x3", 2.0 ** -16),
                                    ("coral_cgemm_3m_fp16x3", 2.0 ** -21)], ids=["4M", "3M-bf16x3", "3M-fp16x3"])
def test_c5_cgemm_cc_4096(gpu_lib, fn, eps):
    n = 4096
    A, B, C = (device.DeviceArray(n * n, np.complex64) for _ in range(3))
    device.fill_random(A, synthetic.seed_for(5, "A"), "uniform")
    device.fill_random(B, synthetic.seed_for(5, "B"), "uniform")
    one = (ctypes.c_float * 2)(1.0, 0.0)
    zero = (ctypes.c_float * 2)(0.0, 0.0)
    call(fn, 2, 2, n, n, n, one, A.ptr, n, B.ptr, n, zero, C.ptr, n, None)
    Ah, Bh, Ch = (host_mat(x, n, n, np.complex64) for x in (A, B, C))
    rows = np.random.default_rng(6).choice(n, 64, replace=False)
    # 3M forms combine three real products (Re, Im and Re + Im planes), so the
    # imaginary part carries up to about three times one product's error
    bound = n * E32 if eps is None else 3 * (eps + n * E32)
    r = RowSample(Ah, Bh, rows, (2, 2), np.complex128).normwise(Ch, 1.0)
    assert r <= bound, (r, r / bound)


def test_c5_zgemm_nn_4096(gpu_lib):
    n = 4096
    A, B, C = (device.DeviceArray(n * n, np.complex128) for _ in range(3))
    device.fill_random(A, synthetic.seed_for(5, "A"), "uniform")
    device.fill_random(B, synthetic.seed_for(5, "B"), "uniform")
    one = (ctypes.c_double * 2)(1.0, 0.0)
    zero = (ctypes.c_double * 2)(0.0, 0.0)
    call("coral_zgemm", 0, 0, n, n, n, one, A.ptr, n, B.ptr, n, zero, C.ptr, n, None)
    Ah, Bh, Ch = (host_mat(x, n, n, np.complex128) for x in (A, B, C))
    rows = np.random.default_rng(7).choice(n, 64, replace=False)
    r = RowSample(Ah, Bh, rows, wide=np.complex128).normwise(Ch, n * E64)
    assert r <= 1.0, r


def test_c5_chemv_upper_16384(gpu_lib):
    n = 16384
    Md = device.DeviceArray(n * n, np.complex64)
    device.fill_random(Md, synthetic.seed_for(5, "C"), "uniform")
    M = host_mat(Md, n, n, np.complex64)
    H = ((M + M.conj().T) * np.float32(0.5)).astype(np.complex64)
    del M
    ref_full = H.astype(np.complex128)
    stored = np.triu(H)
    stored[np.tril_indices(n, -1)] = np.complex64(complex(np.nan, np.nan))  # never read
    Md.upload(np.asfortranarray(stored).reshape(-1, order="F").view(np.float32))
    del stored, H
    rng = np.random.default_rng(8)
    x = (rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)).astype(np.complex64)
    y0 = (rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)).astype(np.complex64)
    X, Y = device.DeviceArray.from_numpy(x), device.DeviceArray.from_numpy(y0)
    alpha = (ctypes.c_float * 2)(0.75, -0.25)
    beta = (ctypes.c_float * 2)(0.5, 0.125)
    call("coral_chemv", 0, n, alpha, Md.ptr, n, X.ptr, 1, beta, Y.ptr, 1, None)
    al, be = 0.75 - 0.25j, 0.5 + 0.125j
    ref = al * (ref_full @ x.astype(np.complex128)) + be * y0
    scale = abs(al) * (np.abs(ref_full) @ np.abs(x)) + abs(be) * np.abs(y0)
    err = np.abs(Y.to_numpy() - ref)
    assert np.all(err <= 4 * n * E32 * scale), float(np.max(err / scale))
