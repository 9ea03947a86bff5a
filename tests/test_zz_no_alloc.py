"""The allocation-free DGEMV entry point (coral_dgemv_noalloc) against coral_dgemv.

Both run the same kernels; the allocation-free form keeps its partial sums in the
library's fixed per-GPU scratch array instead of a stream workspace.  Two calls in
a row (the second smaller, with strides) must give results bit-identical to
coral_dgemv's for NoTranspose and Transpose; y above the fixed capacity (2^19
elements) is rejected with -2 (m) / -3 (n) and y is left untouched.
"""
import numpy as np
import pytest

from coral_amd import device
from coral_amd._lib import call, status_of
from coral_amd.device import DeviceArray

pytestmark = pytest.mark.gpu

CAP = 1 << 19


def run_pair(trans, m, n, incx, incy, alpha, beta, seed, stream):
    rng = np.random.default_rng(seed)
    lda = m + 3
    a = rng.uniform(-1, 1, lda * n)
    xl, yl = (n, m) if trans == 0 else (m, n)
    x = rng.uniform(-1, 1, (xl - 1) * incx + 1)
    y = rng.uniform(-1, 1, (yl - 1) * incy + 1)
    A, X = DeviceArray.from_numpy(a), DeviceArray.from_numpy(x)
    Y1, Y2 = DeviceArray.from_numpy(y), DeviceArray.from_numpy(y)
    h = stream.handle if stream else None
    call("coral_dgemv", trans, m, n, alpha, A.ptr, lda, X.ptr, incx, beta, Y1.ptr, incy, h)
    call("coral_dgemv_noalloc", trans, m, n, alpha, A.ptr, lda, X.ptr, incx, beta, Y2.ptr, incy, h)
    device.synchronize()
    ref, got = Y1.to_numpy(), Y2.to_numpy()
    assert np.array_equal(ref, got), (trans, m, n)
    # and both are the product: elementwise within 2 k eps (|A||x|) + |beta y|
    M = a.reshape(n, lda).T[:m, :]
    op = M if trans == 0 else M.T
    xv, yv = x[::incx], y[::incy]
    want = alpha * (op @ xv) + beta * yv
    scale = abs(alpha) * (np.abs(op) @ np.abs(xv)) + abs(beta) * np.abs(yv)
    assert np.all(np.abs(got[::incy] - want) <= 2 * max(m, n) * 2.0 ** -53 * scale + 1e-300)
    mask = np.ones(got.size, bool)
    mask[::incy] = False
    assert np.array_equal(got[mask], y[mask])


@pytest.mark.parametrize("trans", [0, 1], ids=["N", "T"])
def test_noalloc_matches_dgemv_twice_second_smaller(gpu_lib, trans):
    stream = device.Stream()
    try:
        run_pair(trans, 4096, 3072, 1, 1, 1.0, 0.0, 1 + trans, stream)
        run_pair(trans, 1000, 777, 2, 3, -0.75, 0.5, 3 + trans, stream)
    finally:
        stream.destroy()
    run_pair(trans, 333, 129, 1, 2, 1.5, 1.0, 5 + trans, None)  # the library's default stream


def test_noalloc_at_capacity(gpu_lib):
    run_pair(0, CAP, 16, 1, 1, 1.0, 0.25, 7, None)
    run_pair(1, 16, CAP, 1, 1, 1.0, 0.25, 8, None)


def test_noalloc_rejects_y_beyond_capacity(gpu_lib):
    for trans, m, n, code in ((0, CAP + 1, 1, -2), (1, 1, CAP + 1, -3)):
        lda = m
        A = DeviceArray(lda * n, np.float64)
        X = DeviceArray(max(m, n), np.float64)
        y0 = np.arange(max(m, n), dtype=np.float64)
        Y = DeviceArray.from_numpy(y0)
        assert status_of("coral_dgemv_noalloc", trans, m, n, 1.0, A.ptr, lda, X.ptr, 1, 0.0, Y.ptr, 1, None) == code
        device.synchronize()
        assert np.array_equal(Y.to_numpy(), y0)
