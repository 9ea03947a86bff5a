"""Every tile of every GEMM menu, and the autotuner's persistence, on the GPU.

The default tile choice only exercises a few menu entries at the usual sizes; here
each entry is forced in turn (coral_autotune_set for the shape's bucket) on an odd
shape with padded leading dimensions and checked against the FP64 product with the
normwise bound of test_gpu_level3.py.  Then coral_autotune_run / save / load / get
round-trip a table through a file.
"""
import ctypes
import os

import numpy as np
import pytest

from coral_amd import blas
from coral_amd._lib import load
from oracle import ref64

pytestmark = pytest.mark.gpu

# routine -> (host function, dtype, eps of the normwise bound per unit k, extra eps)
ROUTINES = {
    "sgemm": (blas.sgemm, np.float32, 2.0 ** -24, 0.0),
    "dgemm": (blas.dgemm, np.float64, 2.0 ** -53, 0.0),
    "sgemm_bf16x3": (blas.sgemm_bf16x3, np.float32, 2.0 ** -24, 2.0 ** -16),
    "sgemm_fp16x3": (blas.sgemm_fp16x3, np.float32, 2.0 ** -24, 2.0 ** -21),
    "cgemm": (blas.cgemm, np.complex64, 2.0 ** -24, 0.0),
    "zgemm": (blas.zgemm, np.complex128, 2.0 ** -53, 0.0),
    "cgemm_3m_bf16x3": (blas.cgemm_3m_bf16x3, np.complex64, 2.0 ** -24, 3 * 2.0 ** -16),
    "cgemm_3m_fp16x3": (blas.cgemm_3m_fp16x3, np.complex64, 2.0 ** -24, 3 * 2.0 ** -21),
}


def rand(rng, n, f):
    if np.issubdtype(f, np.complexfloating):
        return (rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)).astype(f)
    return rng.uniform(-1, 1, n).astype(f)


def menu_size(lib, routine, m, n, k, op_a, op_b):
    count = 0
    while lib.coral_autotune_set(routine.encode(), op_a, op_b, m, n, k, count) == 0:
        count += 1
    return count


@pytest.mark.parametrize("routine", list(ROUTINES))
def test_every_menu_tile(gpu_lib, routine):
    lib = load()
    fn, f, eps, eps_split = ROUTINES[routine]
    cplx = np.issubdtype(f, np.complexfloating)
    rng = np.random.default_rng(60)
    m, n, k = 300, 520, 200
    cases = [(0, 0), (1, 1)] + ([(2, 0)] if cplx else [])
    try:
        for op_a, op_b in cases:
            tiles = menu_size(lib, routine, m, n, k, op_a, op_b)
            assert tiles >= 3, (routine, tiles)
            lda, ldb, ldc = (m if op_a == 0 else k) + 3, (k if op_b == 0 else n) + 5, m + 2
            ar, ac = (m, k) if op_a == 0 else (k, m)
            br, bc = (k, n) if op_b == 0 else (n, k)
            a, b, c = rand(rng, lda * ac, f), rand(rng, ldb * bc, f), rand(rng, ldc * n, f)
            al, be = ((0.75, -0.25), (0.5, 0.125)) if cplx else (0.75, -0.5)
            ref = ref64.gemm(op_a, op_b, m, n, k, al, a, lda, b, ldb, be, c, ldc)
            A, B = ref64.mat(a, ar, ac, lda), ref64.mat(b, br, bc, ldb)
            scale = abs(complex(*al)) if cplx else abs(al)
            den = (eps_split + k * eps) * scale * np.linalg.norm(A) * np.linalg.norm(B) + eps * np.linalg.norm(c)
            for t in range(tiles):
                assert lib.coral_autotune_set(routine.encode(), op_a, op_b, m, n, k, t) == 0
                out = c.copy()
                fn(op_a, op_b, m, n, k, al, a, lda, b, ldb, be, out, ldc)
                got = ref64.mat(out, m, n, ldc).astype(ref.dtype)
                want = ref64.mat(ref, m, n, ldc)
                assert np.linalg.norm(got - want) <= den, (routine, op_a, op_b, t)
                # the padding between columns of C is untouched
                pad = out.reshape(n, ldc)[:, m:] if ldc > m else None
                if pad is not None:
                    assert np.array_equal(pad, c.reshape(n, ldc)[:, m:]), (routine, t)
    finally:
        lib.coral_autotune_clear()
        shipped = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "coral_amd",
                               "autotune_table.txt")
        if os.path.exists(shipped):  # the table the library loaded at start
            assert lib.coral_autotune_load(shipped.encode()) == 0


def test_autotune_run_save_load_roundtrip(gpu_lib, tmp_path):
    lib = load()
    try:
        lib.coral_autotune_clear()
        chosen = ctypes.c_int32(-1)
        assert lib.coral_autotune_run(b"sgemm_bf16x3", 0, 0, 1024, 1024, 1024, ctypes.byref(chosen)) == 0
        assert chosen.value >= 0
        path = str(tmp_path / "tune.txt")
        assert lib.coral_autotune_save(path.encode()) == 0
        assert os.path.getsize(path) > 0
        lib.coral_autotune_clear()
        got = ctypes.c_int32(7)
        assert lib.coral_autotune_get(b"sgemm_bf16x3", 0, 0, 1024, 1024, 1024, ctypes.byref(got)) == 0
        assert got.value == -1  # no entry after clear
        assert lib.coral_autotune_load(path.encode()) == 0
        assert lib.coral_autotune_get(b"sgemm_bf16x3", 0, 0, 1024, 1024, 1024, ctypes.byref(got)) == 0
        assert got.value == chosen.value
        # the tuned tile gives a correct product
        rng = np.random.default_rng(61)
        m = n = k = 1024
        a, b = rng.standard_normal(m * k).astype(np.float32), rng.standard_normal(k * n).astype(np.float32)
        out = np.zeros(m * n, np.float32)
        blas.sgemm_bf16x3(0, 0, m, n, k, 1.0, a, m, b, k, 0.0, out, m)
        ref = ref64.gemm(0, 0, m, n, k, 1.0, a, m, b, k, 0.0, np.zeros(m * n, np.float32), m)
        assert np.linalg.norm(out - ref) <= (2.0 ** -16 + k * 2.0 ** -24) * np.linalg.norm(a) * np.linalg.norm(b)
    finally:
        lib.coral_autotune_clear()
        shipped = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "coral_amd",
                               "autotune_table.txt")
        if os.path.exists(shipped):  # the table the library loaded at start
            assert lib.coral_autotune_load(shipped.encode()) == 0
