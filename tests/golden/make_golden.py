"""Writes the golden fixtures under tests/golden/ (run: python tests/golden/make_golden.py).

Fixtures are data only: inputs and expected outputs.

* ``kat.json`` -- the reference's known-answer examples, transcribed from its
  rustdoc: SGEMM/DGEMM 2x2 (coral-aarch64/src/level3/sgemm.rs:36-81, dgemm.rs:37-80)
  and CGEMM/ZGEMM 2x2 complex (cgemm.rs:38-94, zgemm.rs:38-94).
* ``gemm_cases.npz`` -- the reference GEMM tests' deterministic affine generators
  (coral-aarch64/tests/level3/sgemm_tests.rs:112-114, cgemm_tests.rs:136-153) on
  the reference test shapes (small 5x7x4, not-block-multiple 13x14x11, padded ld
  17x19x13 with +3/+2/+5/+4/+7, sgemm_tests.rs:187-235) for every (alpha, beta)
  of sgemm_tests.rs:151-156 and op pair, with the expected C from numpy float64 /
  complex128 matmul on the rounded inputs (numpy's OpenBLAS is the same library
  family as the reference tests' CBLAS oracle, coral-aarch64/Cargo.toml:25-31).
* ``gemm_reference_shapes.npz`` -- the larger reference shapes (128^3 and the
  rectangular 64x192x15 / 192x64x15, sgemm_tests.rs:221-235, cgemm_tests.rs:260-274)
  for every run of the reference's run_all_ops, as signatures: the expected C at
  192 fixed sampled entries and its Frobenius norm (inputs come from the same
  generators, so tests rebuild them).
* ``iamax_cases.npz`` -- the planted-maximum vectors of
  coral-aarch64/tests/level1/iamax_tests.rs:79-140 with their expected indices.
"""
from __future__ import annotations

import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

# ---------------------------------------------------------------- KATs
KAT = {
    "sgemm_2x2": {
        "m": 2, "n": 2, "k": 2, "alpha": 2.0, "beta": 1.0,
        "a": [1.0, 2.0, 3.0, 4.0], "b": [5.0, 6.0, 7.0, 8.0], "c": [1.0, 0.0, 0.0, 1.0],
        "expect": [47.0, 68.0, 62.0, 93.0],
    },
    "cgemm_2x2": {
        "m": 2, "n": 2, "k": 2, "alpha": [2.0, 1.0], "beta": [1.0, -1.0],
        "a": [1.0, 2.0, 5.0, 6.0, 3.0, 4.0, 7.0, 8.0],
        "b": [9.0, 1.0, 4.0, 5.0, 2.0, 3.0, 6.0, 7.0],
        "c": [1.0, 1.0, 3.0, 0.0, 0.0, 2.0, 4.0, 1.0],
        "expect": [-50.0, 99.0, -69.0, 276.0, -78.0, 92.0, -163.0, 223.0],
    },
}

# ---------------------------------------------------------------- GEMM cases
SHAPES = [  # (name, m, n, k, lda_n, lda_t, ldb_n, ldb_t, ldc)
    ("small", 5, 7, 4, 5, 4, 4, 7, 5),
    ("not_block_multiple", 13, 14, 11, 13, 11, 11, 14, 13),
    ("padded_lds", 17, 19, 13, 20, 15, 18, 23, 24),
]
REAL_CASES = [(1.0, 0.0), (0.5, 1.0), (0.75, -0.25), (0.0, 0.7)]
# (alpha, beta) lists of sgemm_tests.rs:151-156 and cgemm_tests.rs:196-201
CPLX_CASES = [((1.0, 0.0), (0.0, 0.0)), ((0.5, 0.25), (1.0, 0.0)), ((0.75, -0.25), (-0.5, 0.3)),
              ((0.0, 0.0), (0.7, -0.2))]
# the larger reference shapes (sgemm_tests.rs:221-235, cgemm_tests.rs:260-274), kept
# as sampled signatures: their inputs come from the same generators
LARGE_SHAPES = [  # (name, m, n, k, lda_n, lda_t, ldb_n, ldb_t, ldc)
    ("medium_square", 128, 128, 128, 128, 128, 128, 128, 128),
    ("rectangular_wide", 64, 192, 15, 64, 15, 15, 192, 64),
    ("rectangular_tall", 192, 64, 15, 192, 15, 15, 64, 192),
]
SAMPLES = 192


def colmajor(rows, cols, ld, f, dtype):
    """Column-major buffer with element (i, j) = f(i, j), evaluated elementwise in the
    generator's own precision (f receives integer-valued arrays of that type)."""
    a = np.zeros(ld * cols, dtype)
    real = np.float32 if np.dtype(dtype) in (np.float32, np.complex64) else np.float64
    i = np.arange(rows, dtype=real)[:, None]
    j = np.arange(cols, dtype=real)[None, :]
    vals = f(i, j)
    if isinstance(vals, tuple):
        vals = vals[0] + 1j * vals[1]
    for c in range(cols):
        a[c * ld: c * ld + rows] = np.broadcast_to(vals, (rows, cols))[:, c]
    return a


def op(mat, o):
    return mat if o == 0 else (mat.T if o == 1 else mat.conj().T)


def view(buf, rows, cols, ld):
    return buf.reshape(-1, ld).T[:rows, :cols] if False else np.stack(
        [buf[j * ld: j * ld + rows] for j in range(cols)], axis=1)


def real_case(dt, m, n, k, oa, ob, lda, ldb, ldc, alpha, beta):
    f32 = np.float32 if dt == "s" else np.float64
    ar, ac = (m, k) if oa == 0 else (k, m)
    br, bc = (k, n) if ob == 0 else (n, k)
    a = colmajor(ar, ac, lda, lambda i, j: f32(0.1) + f32(i) * f32(0.25) + f32(j) * f32(0.125), f32)
    b = colmajor(br, bc, ldb, lambda i, j: f32(-0.2) + f32(i) * f32(0.05) - f32(j) * f32(0.075), f32)
    c = colmajor(m, n, ldc, lambda i, j: f32(0.3) - f32(i) * f32(0.01) + f32(j) * f32(0.02), f32)
    A = view(a, ar, ac, lda).astype(np.float64)
    B = view(b, br, bc, ldb).astype(np.float64)
    C = view(c, m, n, ldc).astype(np.float64)
    ref = alpha * (op(A, oa) @ op(B, ob)) + beta * C
    out = c.astype(np.float64).copy()
    for j in range(n):
        out[j * ldc: j * ldc + m] = ref[:, j]
    return a, b, c, out


def cplx_case(dt, m, n, k, oa, ob, lda, ldb, ldc, alpha, beta):
    f = np.float32 if dt == "c" else np.float64
    cd = np.complex64 if dt == "c" else np.complex128
    ar, ac = (m, k) if oa == 0 else (k, m)
    br, bc = (k, n) if ob == 0 else (n, k)
    a = colmajor(ar, ac, lda, lambda i, j: (f(0.1) + f(i) * f(0.25) + f(j) * f(0.125),
                                            f(-0.05) + f(i) * f(0.20) - f(j) * f(0.075)), cd)
    b = colmajor(br, bc, ldb, lambda i, j: (f(-0.2) + f(i) * f(0.05) - f(j) * f(0.075),
                                            f(0.15) - f(i) * f(0.03) + f(j) * f(0.02)), cd)
    c = colmajor(m, n, ldc, lambda i, j: (f(0.3) - f(i) * f(0.01) + f(j) * f(0.02),
                                          f(-0.1) + f(i) * f(0.015) - f(j) * f(0.025)), cd)
    A = view(a, ar, ac, lda).astype(np.complex128)
    B = view(b, br, bc, ldb).astype(np.complex128)
    C = view(c, m, n, ldc).astype(np.complex128)
    al, be = complex(*alpha), complex(*beta)
    ref = al * (op(A, oa) @ op(B, ob)) + be * C
    out = c.astype(np.complex128).copy()
    for j in range(n):
        out[j * ldc: j * ldc + m] = ref[:, j]
    return a, b, c, out


def gemm_cases():
    data = {}
    index = []
    for dt in "sdcz":
        cplx = dt in "cz"
        ops = [0, 1, 2] if cplx else [0, 1, 2]
        cases = CPLX_CASES if cplx else REAL_CASES
        for (name, m, n, k, lda_n, lda_t, ldb_n, ldb_t, ldc) in SHAPES:
            for ci, (alpha, beta) in enumerate(cases):
                for oa in ops:
                    for ob in ops:
                        if not cplx and (oa == 2 or ob == 2) and ci > 1:
                            continue  # the reference runs ConjTranspose for two cases only
                        lda = lda_n if oa == 0 else lda_t
                        ldb = ldb_n if ob == 0 else ldb_t
                        fn = cplx_case if cplx else real_case
                        a, b, c, out = fn(dt, m, n, k, oa, ob, lda, ldb, ldc, alpha, beta)
                        key = f"{dt}_{name}_{ci}_{oa}{ob}"
                        data[key + "_a"] = a
                        data[key + "_b"] = b
                        data[key + "_c"] = c
                        data[key + "_expect"] = out
                        index.append([key, dt, m, n, k, oa, ob, lda, ldb, ldc,
                                      list(alpha) if cplx else alpha, list(beta) if cplx else beta])
    data["index"] = np.array(json.dumps(index))
    np.savez_compressed(os.path.join(HERE, "gemm_cases.npz"), **data)
    return len(index)


def reference_runs(dt):
    """(case index, op_a, op_b) of one run_all_ops call: real routines run the four
    N/T pairs per (alpha, beta) plus ConjTranspose x NoTranspose with the first case
    and NoTranspose x ConjTranspose with the second (sgemm_tests.rs:158-185); complex
    routines run all nine pairs per case (cgemm_tests.rs:203-224)."""
    runs = []
    for ci in range(4):
        if dt in "cz":
            runs += [(ci, oa, ob) for oa in (0, 1, 2) for ob in (0, 1, 2)]
        else:
            runs += [(ci, oa, ob) for oa in (0, 1) for ob in (0, 1)] + [(0, 2, 0), (1, 0, 2)]
    return runs


def large_case(dt, name, m, n, k, lda_n, lda_t, ldb_n, ldb_t, ldc, ci, oa, ob):
    """Inputs of one larger reference run and its expected C (float64 / complex128)."""
    lda = lda_n if oa == 0 else lda_t
    ldb = ldb_n if ob == 0 else ldb_t
    if dt in "cz":
        alpha, beta = CPLX_CASES[ci]
        a, b, c, out = cplx_case(dt, m, n, k, oa, ob, lda, ldb, ldc, alpha, beta)
    else:
        alpha, beta = REAL_CASES[ci]
        a, b, c, out = real_case(dt, m, n, k, oa, ob, lda, ldb, ldc, alpha, beta)
    return lda, ldb, alpha, beta, a, b, c, out


def sample_positions(m, n, ldc):
    """Fixed buffer offsets i + j ldc of the sampled entries of an m x n result."""
    rng = np.random.default_rng(m * 100000 + n * 100 + ldc)
    cells = rng.choice(m * n, min(SAMPLES, m * n), replace=False)
    return np.sort((cells % m) + (cells // m) * ldc)


def large_gemm_cases():
    data, index = {}, []
    for dt in "sdcz":
        for (name, m, n, k, lda_n, lda_t, ldb_n, ldb_t, ldc) in LARGE_SHAPES:
            pos = sample_positions(m, n, ldc)
            data[f"{name}_pos"] = pos
            for ci, oa, ob in reference_runs(dt):
                lda, ldb, alpha, beta, a, b, c, out = large_case(dt, name, m, n, k, lda_n, lda_t, ldb_n, ldb_t,
                                                                 ldc, ci, oa, ob)
                key = f"{dt}_{name}_{ci}_{oa}{ob}"
                data[key + "_sample"] = out[pos]
                region = np.stack([out[j * ldc: j * ldc + m] for j in range(n)], axis=1)
                data[key + "_fro"] = np.float64(np.linalg.norm(region))
                index.append([key, dt, name, ci, oa, ob, m, n, k, lda, ldb, ldc])
    data["index"] = np.array(json.dumps(index))
    np.savez_compressed(os.path.join(HERE, "gemm_reference_shapes.npz"), **data)
    return len(index)


def iamax_cases():
    n1 = 1024
    k = np.arange(n1, dtype=np.float32)
    x1 = (np.abs(np.sin(k) * np.float32(3.0)) * np.where(np.arange(n1) % 2 == 0, 1.0, -1.0)).astype(np.float32)
    x1[777] = np.float32(-1234.5)
    n2, inc2 = 777, 3
    x2 = np.zeros((n2 - 1) * inc2 + 1, np.float32)
    kk = np.arange(n2)
    vals = ((np.float32(0.2) - np.float32(0.001) * kk.astype(np.float32)) *
            np.where(kk % 2 == 0, 1.0, -1.0)).astype(np.float32)
    vals[321] = np.float32(9999.0)
    x2[::inc2] = vals
    np.savez_compressed(os.path.join(HERE, "iamax_cases.npz"), x1=x1, expect1=np.int64(777),
                        x2=x2, inc2=np.int64(inc2), n2=np.int64(n2), expect2=np.int64(321))


if __name__ == "__main__":
    with open(os.path.join(HERE, "kat.json"), "w") as f:
        json.dump(KAT, f, indent=1)
    print("gemm cases:", gemm_cases())
    print("larger reference gemm shapes:", large_gemm_cases())
    iamax_cases()
    print("wrote", sorted(p for p in os.listdir(HERE) if not p.endswith(".py")))
