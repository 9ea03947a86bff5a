"""GPU parity for the batched uneven GEMM and batched DTRSV (device-pointer C ABI),
including hipGraph capture and repeated replay of the persistent launch.

Each problem is compared with FP64 under the GEMM normwise bound, and with the
result of the single-problem routine (coral_sgemm / coral_dgemm / coral_dtrsv).
"""
import ctypes

import numpy as np
import pytest

from coral_amd import device
from coral_amd._lib import call
from coral_amd.device import DeviceArray
from oracle import ref64

pytestmark = pytest.mark.gpu


class Batch:
    """Problems in one device arena plus the pointer / dim arrays the ABI takes."""

    def __init__(self, dims, dtype, op_a, op_b, rng, lds_pad=0):
        self.dtype = np.dtype(dtype)
        self.dims = dims
        self.op_a, self.op_b = op_a, op_b
        self.host_a, self.host_b, self.host_c, self.ld = [], [], [], []
        offs_a, offs_b, offs_c = [], [], []
        tot_a = tot_b = tot_c = 0
        for (m, n, k) in dims:
            lda = (m if op_a == 0 else k) + lds_pad
            ldb = (k if op_b == 0 else n) + lds_pad
            ldc = m + lds_pad
            ca = k if op_a == 0 else m
            cb = n if op_b == 0 else k
            a = rng.standard_normal(max(lda * ca, 1)).astype(dtype)
            b = rng.standard_normal(max(ldb * cb, 1)).astype(dtype)
            c = rng.standard_normal(max(ldc * n, 1)).astype(dtype)
            self.host_a.append(a)
            self.host_b.append(b)
            self.host_c.append(c)
            self.ld.append((max(lda, 1), max(ldb, 1), max(ldc, 1)))
            offs_a.append(tot_a)
            offs_b.append(tot_b)
            offs_c.append(tot_c)
            tot_a += a.size
            tot_b += b.size
            tot_c += c.size
        self.A = DeviceArray.from_numpy(np.concatenate(self.host_a))
        self.B = DeviceArray.from_numpy(np.concatenate(self.host_b))
        self.C = DeviceArray.from_numpy(np.concatenate(self.host_c))
        self.offs_c = offs_c
        es = self.dtype.itemsize
        pa = np.array([self.A.ptr + o * es for o in offs_a], np.uint64)
        pb = np.array([self.B.ptr + o * es for o in offs_b], np.uint64)
        pc = np.array([self.C.ptr + o * es for o in offs_c], np.uint64)
        self.PA, self.PB, self.PC = (DeviceArray.from_numpy(x) for x in (pa, pb, pc))
        d = np.array(dims, np.int32).reshape(-1, 3)
        lds = np.array(self.ld, np.int32).reshape(-1, 3)
        self.M, self.N, self.K = (DeviceArray.from_numpy(np.ascontiguousarray(d[:, i])) for i in range(3))
        self.LDA, self.LDB, self.LDC = (DeviceArray.from_numpy(np.ascontiguousarray(lds[:, i])) for i in range(3))

    def run(self, alpha, beta, stream=None):
        fn = "coral_sgemm_batched" if self.dtype == np.float32 else "coral_dgemm_batched"
        call(fn, self.op_a, self.op_b, len(self.dims), self.M.ptr, self.N.ptr, self.K.ptr, alpha, self.PA.ptr,
             self.LDA.ptr, self.PB.ptr, self.LDB.ptr, beta, self.PC.ptr, self.LDC.ptr,
             stream.handle if stream else None)

    def results(self):
        allc = self.C.to_numpy()
        return [allc[o:o + c.size] for o, c in zip(self.offs_c, self.host_c)]


def check(batch, alpha, beta, eps):
    out = batch.results()
    worst = 0.0
    for (m, n, k), a, b, c, (lda, ldb, ldc), o in zip(batch.dims, batch.host_a, batch.host_b, batch.host_c,
                                                     batch.ld, out):
        if m == 0 or n == 0:
            assert np.array_equal(o, c)
            continue
        ref = ref64.gemm(batch.op_a, batch.op_b, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc)
        ar, ac = (m, k) if batch.op_a == 0 else (k, m)
        br, bc = (k, n) if batch.op_b == 0 else (n, k)
        A = ref64.mat(a, ar, ac, lda) if k else np.zeros((1, 1))
        B = ref64.mat(b, br, bc, ldb) if k else np.zeros((1, 1))
        C = ref64.mat(c, m, n, ldc)
        num = np.linalg.norm(ref64.mat(o, m, n, ldc).astype(np.float64) - ref64.mat(ref, m, n, ldc))
        # gamma_{k+2}: k roundings in the dot products, one in the alpha scaling and one
        # in the beta update (which the compiler may fuse either way: fma(alpha, acc,
        # beta c) or fma(beta, c, alpha acc)); at k = 1 the k-only bound is exceeded by
        # a correctly rounded fma.
        den = (k + 2) * eps * (abs(alpha) * np.linalg.norm(A) * np.linalg.norm(B) + abs(beta) * np.linalg.norm(C))
        worst = max(worst, num / (den + 1e-300))
    return worst


def log_uniform_dims(rng, count, lo, hi):
    return [tuple(int(v) for v in np.exp(rng.uniform(np.log(lo), np.log(hi + 1), 3)).astype(int).clip(lo, hi))
            for _ in range(count)]


@pytest.mark.parametrize("dtype,eps", [(np.float32, 2.0 ** -24), (np.float64, 2.0 ** -53)])
@pytest.mark.parametrize("ops", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_batched_uneven(gpu_lib, dtype, eps, ops):
    rng = np.random.default_rng(40 + ops[0] * 2 + ops[1])
    dims = log_uniform_dims(rng, 300, 1, 300) + [(0, 5, 5), (5, 0, 5), (7, 9, 0), (1, 1, 1), (64, 64, 64)]
    batch = Batch(dims, dtype, ops[0], ops[1], rng, lds_pad=1)
    batch.run(1.25, 0.5)
    device.synchronize()
    assert check(batch, 1.25, 0.5, eps) <= 1.0


def test_batched_queue_shares_longer_than_one_chunk(gpu_lib):
    """20,000 small problems: each of the 8 work queues holds 2,500, more than the setup
    kernel's 128 threads x 16 counts per queue in registers, so its per-queue prefix sums
    re-read the counts chunk by chunk; every problem is checked against FP64."""
    rng = np.random.default_rng(55)
    dims = [tuple(int(v) for v in rng.integers(1, 24, 3)) for _ in range(20000)]
    batch = Batch(dims, np.float32, 0, 0, rng)
    batch.run(1.0, 0.0)
    device.synchronize()
    assert check(batch, 1.0, 0.0, 2.0 ** -24) <= 1.0


def test_batched_alpha_zero_and_beta_zero(gpu_lib):
    rng = np.random.default_rng(50)
    dims = log_uniform_dims(rng, 50, 1, 100)
    batch = Batch(dims, np.float64, 0, 0, rng)
    # beta == 0 must not read C: plant NaN in C
    nan_c = np.full(batch.C.count, np.nan)
    batch.C.upload(nan_c)
    batch.run(1.0, 0.0)
    out = batch.results()
    for (m, n, k), o, (lda, ldb, ldc) in zip(dims, out, batch.ld):
        assert np.all(np.isfinite(ref64.mat(o, m, n, ldc)))
    # alpha == 0: C := beta*C, A/B unread (NaN operands)
    batch.A.upload(np.full(batch.A.count, np.nan))
    batch.C.upload(np.concatenate(batch.host_c))
    batch.run(0.0, 2.0)
    for c, o in zip(batch.host_c, batch.results()):
        assert np.array_equal(o, 2.0 * c)


def test_batched_graph_capture_replays(gpu_lib):
    rng = np.random.default_rng(51)
    dims = log_uniform_dims(rng, 400, 8, 200)
    batch = Batch(dims, np.float32, 0, 0, rng)
    stream = device.Stream()
    batch.run(1.0, 0.0, stream)  # sizes the workspace outside the capture
    stream.synchronize()
    eager = batch.results()
    batch.C.upload(np.zeros(batch.C.count, np.float32))
    g = ctypes.c_void_p()
    call("coral_graph_begin_capture", stream.handle)
    batch.run(1.0, 0.0, stream)
    call("coral_graph_end_capture", stream.handle, ctypes.byref(g))
    for replay in range(4):
        batch.C.upload(np.full(batch.C.count, np.nan, np.float32))  # eager work between replays
        call("coral_graph_launch", g.value, stream.handle)
        stream.synchronize()
        for e, o in zip(eager, batch.results()):
            assert np.array_equal(e, o), replay
    call("coral_graph_destroy", g.value)
    stream.destroy()


def test_batched_matches_single_problem_routine(gpu_lib):
    rng = np.random.default_rng(52)
    dims = log_uniform_dims(rng, 40, 8, 512)
    batch = Batch(dims, np.float64, 0, 0, rng)
    batch.run(1.0, 0.0)
    out = batch.results()
    for (m, n, k), a, b, c, (lda, ldb, ldc), o in zip(dims, batch.host_a, batch.host_b, batch.host_c, batch.ld, out):
        da, db, dc = DeviceArray.from_numpy(a), DeviceArray.from_numpy(b), DeviceArray.from_numpy(c)
        call("coral_dgemm", 0, 0, m, n, k, 1.0, da.ptr, lda, db.ptr, ldb, 0.0, dc.ptr, ldc, None)
        single = dc.to_numpy()
        assert np.allclose(single, o, rtol=1e-12, atol=1e-12)


def test_dtrsv_batched(gpu_lib):
    rng = np.random.default_rng(53)
    for uplo, trans, diag in ((1, 0, 1), (0, 0, 1), (1, 1, 0), (0, 2, 1)):
        ns = [int(v) for v in np.exp(rng.uniform(np.log(1), np.log(257), 120)).astype(int).clip(1, 256)]
        mats, xs = [], []
        for n in ns:
            A = np.diag(rng.uniform(1, 2, n)) + (np.tril(rng.uniform(-1, 1, (n, n)) / n, -1) if uplo == 1
                                                 else np.triu(rng.uniform(-1, 1, (n, n)) / n, 1))
            mats.append(np.asfortranarray(A).reshape(-1, order="F"))
            xs.append(rng.uniform(-1, 1, n))
        offs_a = np.cumsum([0] + [m.size for m in mats])[:-1]
        offs_x = np.cumsum([0] + [x.size for x in xs])[:-1]
        DA = DeviceArray.from_numpy(np.concatenate(mats))
        DX = DeviceArray.from_numpy(np.concatenate(xs))
        PA = DeviceArray.from_numpy(np.array([DA.ptr + 8 * int(o) for o in offs_a], np.uint64))
        PX = DeviceArray.from_numpy(np.array([DX.ptr + 8 * int(o) for o in offs_x], np.uint64))
        Ns = DeviceArray.from_numpy(np.array(ns, np.int32))
        call("coral_dtrsv_batched", uplo, trans, diag, len(ns), Ns.ptr, PA.ptr, Ns.ptr, PX.ptr, None)
        res = DX.to_numpy()
        for n, a, x, o in zip(ns, mats, xs, offs_x):
            ref = ref64.trsv(uplo, trans, diag, n, a, n, x, 1)
            got = res[o:o + n]
            assert np.allclose(got, ref, rtol=1e-12, atol=1e-12), (n, uplo, trans, diag)


def test_batched_graph_keeps_its_workspace(gpu_lib):
    # A captured graph reads the stream's batched workspace (prefix sums, kinds, queue
    # head).  Eager calls that grow the stream's workspace afterwards, and destroying
    # the stream, must leave the graph's copy alive: the replay on another stream
    # still gives the eager result.
    rng = np.random.default_rng(54)
    small = Batch(log_uniform_dims(rng, 200, 8, 128), np.float32, 0, 0, rng)
    big = Batch(log_uniform_dims(rng, 3000, 8, 64), np.float32, 0, 0, rng)
    s1 = device.Stream()
    small.run(1.0, 0.0, s1)
    s1.synchronize()
    eager = small.results()
    g = ctypes.c_void_p()
    call("coral_graph_begin_capture", s1.handle)
    small.run(1.0, 0.0, s1)
    call("coral_graph_end_capture", s1.handle, ctypes.byref(g))
    big.run(1.0, 0.0, s1)  # grows the batched workspace of s1
    n = 1536  # and a split-form GEMM grows the packing and row-maximum workspaces
    A, B, C = (DeviceArray(n * n, np.float32) for _ in range(3))
    device.fill_random(A, 1, "normal", s1)
    device.fill_random(B, 2, "normal", s1)
    call("coral_sgemm_bf16x3", 0, 0, n, n, n, 1.0, A.ptr, n, B.ptr, n, 0.0, C.ptr, n, s1.handle)
    s1.synchronize()
    s1.destroy()
    s2 = device.Stream()
    try:
        for replay in range(3):
            small.C.upload(np.full(small.C.count, np.nan, np.float32))
            call("coral_graph_launch", g.value, s2.handle)
            s2.synchronize()
            for e, o in zip(eager, small.results()):
                assert np.array_equal(e, o), replay
    finally:
        call("coral_graph_destroy", g.value)
        s2.destroy()


def test_batched_invalid_problem_is_skipped_and_reported(gpu_lib):
    rng = np.random.default_rng(55)
    dims = log_uniform_dims(rng, 30, 4, 80)
    batch = Batch(dims, np.float64, 0, 0, rng)
    bad = 7
    lds = np.array(batch.ld, np.int32)
    lda = lds[:, 0].copy()
    lda[bad] = dims[bad][0] - 1 if dims[bad][0] > 1 else 0  # below max(1, m)
    batch.LDA.upload(lda)
    s = device.Stream()
    batch.run(1.0, 0.5, s)
    assert load_status("coral_stream_check", s.handle) == 7  # CORAL_ERR_PROBLEM_ARG
    assert load_status("coral_stream_check", s.handle) == 0  # the check clears it
    out = batch.results()
    assert np.array_equal(out[bad], batch.host_c[bad])
    # lda above 2^24 (the tile kernel's 32-bit offsets) is refused the same way; the
    # matrix is never read
    big = 11
    lda2 = lds[:, 0].copy()
    lda2[big] = (1 << 24) + 1
    batch.LDA.upload(lda2)
    batch.C.upload(np.concatenate(batch.host_c))
    batch.run(1.0, 0.5, s)
    assert load_status("coral_stream_check", s.handle) == 7
    assert np.array_equal(batch.results()[big], batch.host_c[big])
    batch.LDA.upload(lds[:, 0].copy())
    ok = [i for i in range(len(dims)) if i != bad]
    for i in ok:
        m, n, k = dims[i]
        ref = ref64.gemm(0, 0, m, n, k, 1.0, batch.host_a[i], batch.ld[i][0], batch.host_b[i], batch.ld[i][1], 0.5,
                         batch.host_c[i], batch.ld[i][2])
        assert np.allclose(out[i], ref, rtol=1e-11, atol=1e-11), i
    s.destroy()


def load_status(name, *args):
    from coral_amd._lib import status_of

    return status_of(name, *args)


def test_dtrsv_batched_large_n_and_bad_lda(gpu_lib):
    # n above the LDS-staged size is solved in place in x; a problem with lda < n is
    # skipped (x unchanged) and reported through coral_stream_check
    rng = np.random.default_rng(56)
    ns = [100, 5000, 64, 40]
    ldas = [100, 5000, 64, 39]
    mats, xs = [], []
    for n, lda in zip(ns, ldas):
        A = np.diag(rng.uniform(1, 2, n)) + np.tril(rng.uniform(-1, 1, (n, n)) / n, -1)
        mats.append(np.asfortranarray(A).reshape(-1, order="F"))  # stored with ld = n
        xs.append(rng.uniform(-1, 1, n))
    offs_a = np.cumsum([0] + [m.size for m in mats])[:-1]
    offs_x = np.cumsum([0] + [x.size for x in xs])[:-1]
    DA = DeviceArray.from_numpy(np.concatenate(mats))
    DX = DeviceArray.from_numpy(np.concatenate(xs))
    PA = DeviceArray.from_numpy(np.array([DA.ptr + 8 * int(o) for o in offs_a], np.uint64))
    PX = DeviceArray.from_numpy(np.array([DX.ptr + 8 * int(o) for o in offs_x], np.uint64))
    Ns = DeviceArray.from_numpy(np.array(ns, np.int32))
    LDs = DeviceArray.from_numpy(np.array(ldas, np.int32))
    s = device.Stream()
    call("coral_dtrsv_batched", 1, 0, 1, len(ns), Ns.ptr, PA.ptr, LDs.ptr, PX.ptr, s.handle)
    assert load_status("coral_stream_check", s.handle) == 7
    res = DX.to_numpy()
    for i, (n, a, x, o) in enumerate(zip(ns, mats, xs, offs_x)):
        got = res[o:o + n]
        if ldas[i] < n:
            assert np.array_equal(got, x)
            continue
        A = a.reshape(n, n).T
        berr = np.linalg.norm(x - A @ got) / (np.linalg.norm(A) * np.linalg.norm(got) + np.linalg.norm(x))
        assert berr <= 4 * n * 2.0 ** -53, (n, berr)
    s.destroy()
