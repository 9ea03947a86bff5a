"""The C-ABI boundary without a GPU: the library loads, exports every symbol
include/coral_hip.h declares, and rejects invalid arguments with CORAL's
argument index before touching memory or the device."""
import ctypes
import re
import subprocess

import numpy as np
import pytest

import coral_amd
from coral_amd._lib import HEADER_PATH, LIB_PATH, PROTOTYPES


def test_header_declares_the_call_surface():
    names = set(PROTOTYPES)
    for p in "sdcz":
        for r in ("gemm", "gemv", "trsv", "trmv", "axpy"):
            assert f"coral_{p}{r}" in names and f"coralh_{p}{r}" in names
        assert f"coral_i{p}amax" in names
    for r in ("ssymv", "dsymv", "chemv", "zhemv", "sdot", "ddot", "cdotu", "cdotc", "zdotu", "zdotc",
              "snrm2", "dnrm2", "scnrm2", "dznrm2", "sgemm_bf16x3", "cgemm_3m_bf16x3", "sgemm_fp16x3",
              "cgemm_3m_fp16x3", "sgemm_batched",
              "dgemm_batched", "dtrsv_batched", "autotune_run", "autotune_save", "autotune_load"):
        assert f"coral_{r}" in names, r
    for r in ("sger", "dger", "cgeru", "cgerc", "zgeru", "zgerc", "ssyr", "dsyr", "cher", "zher",
              "ssyr2", "dsyr2", "cher2", "zher2"):
        assert f"coral_{r}" in names and f"coralh_{r}" in names, r
    for r in ("sscal", "dscal", "cscal", "zscal", "csscal", "zdscal", "sasum", "dasum", "scasum", "dzasum",
              "scopy", "dcopy", "ccopy", "zcopy", "sswap", "dswap", "cswap", "zswap", "srot", "drot", "csrot",
              "zdrot", "srotm", "drotm", "srotg", "drotg", "srotmg", "drotmg"):
        assert f"coral_{r}" in names and f"coralh_{r}" in names, r


def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB_PATH], check=True, capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (\w+)$", out, flags=re.M))
    missing = sorted(set(PROTOTYPES) - exported)
    assert not missing, missing
    lib = coral_amd.load()
    for name in PROTOTYPES:
        assert getattr(lib, name) is not None


def _status(name, *args):
    return coral_amd.status_of(name, *args)


def test_gemm_argument_validation_codes():
    buf = np.zeros(64, np.float32)
    p = buf.ctypes.data
    # CORAL argument order: op_a=1 op_b=2 m=3 n=4 k=5 alpha=6 a=7 lda=8 b=9 ldb=10 beta=11 c=12 ldc=13
    assert _status("coral_sgemm", 3, 0, 4, 4, 4, 1.0, p, 4, p, 4, 0.0, p, 4, None) == -1
    assert _status("coral_sgemm", 0, -1, 4, 4, 4, 1.0, p, 4, p, 4, 0.0, p, 4, None) == -2
    assert _status("coral_sgemm", 0, 0, -1, 4, 4, 1.0, p, 4, p, 4, 0.0, p, 4, None) == -3
    assert _status("coral_sgemm", 0, 0, 4, 4, 4, 1.0, p, 3, p, 4, 0.0, p, 4, None) == -8
    assert _status("coral_sgemm", 0, 1, 4, 5, 4, 1.0, p, 4, p, 4, 0.0, p, 4, None) == -10  # ldb < n for B^T
    assert _status("coral_sgemm", 0, 0, 4, 4, 4, 1.0, p, 4, p, 4, 0.0, p, 3, None) == -13
    assert _status("coral_dgemm", 0, 0, 4, 4, 4, 1.0, None, 4, p, 4, 0.0, p, 4, None) == -7


def test_level2_level1_validation_codes():
    buf = np.zeros(64)
    p = buf.ctypes.data
    assert _status("coral_dgemv", 5, 4, 4, 1.0, p, 4, p, 1, 0.0, p, 1, None) == -1
    assert _status("coral_dgemv", 0, 4, 4, 1.0, p, 3, p, 1, 0.0, p, 1, None) == -6
    assert _status("coral_dgemv", 0, 4, 4, 1.0, p, 4, p, 0, 0.0, p, 1, None) == -8
    assert _status("coral_dgemv", 0, 4, 4, 1.0, p, 4, p, 1, 0.0, p, -2, None) == -11
    assert _status("coral_dtrsv", 2, 0, 1, 4, p, 4, p, 1, None) == -1
    assert _status("coral_dtrsv", 1, 3, 1, 4, p, 4, p, 1, None) == -2
    assert _status("coral_dtrsv", 1, 0, 2, 4, p, 4, p, 1, None) == -3
    assert _status("coral_dtrsv", 1, 0, 1, 4, p, 2, p, 1, None) == -6
    assert _status("coral_dsymv", 1, 4, 1.0, p, 3, p, 1, 0.0, p, 1, None) == -5
    assert _status("coral_ddot", -1, p, 1, p, 1, p, None) == -1
    assert _status("coral_ddot", 4, p, 0, p, 1, p, None) == -3
    assert _status("coral_dnrm2", 4, p, 1, None, None) == -4


def test_rank_update_validation_codes():
    buf = np.zeros(64)
    p = buf.ctypes.data
    alv, zv = np.array([1.0, 0.5]), np.zeros(2)  # kept alive while their pointers are used
    al, zero = alv.ctypes.data, zv.ctypes.data
    # ger: m=1 n=2 alpha=3 x=4 incx=5 y=6 incy=7 a=8 lda=9
    assert _status("coral_dger", -1, 4, 1.0, p, 1, p, 1, p, 4, None) == -1
    assert _status("coral_dger", 4, -1, 1.0, p, 1, p, 1, p, 4, None) == -2
    assert _status("coral_dger", 4, 4, 1.0, p, 0, p, 1, p, 4, None) == -5
    assert _status("coral_dger", 4, 4, 1.0, p, 1, p, 0, p, 4, None) == -7
    assert _status("coral_dger", 4, 4, 1.0, p, 1, p, 1, p, 3, None) == -9
    assert _status("coral_dger", 4, 4, 1.0, None, 1, p, 1, p, 4, None) == -4
    assert _status("coral_zgerc", 4, 4, None, p, 1, p, 1, p, 4, None) == -3
    assert _status("coral_zgeru", 4, 4, al, p, 1, None, 1, p, 4, None) == -6
    # syr / her: uplo=1 n=2 alpha=3 x=4 incx=5 a=6 lda=7
    assert _status("coral_dsyr", 2, 4, 1.0, p, 1, p, 4, None) == -1
    assert _status("coral_dsyr", 0, -1, 1.0, p, 1, p, 4, None) == -2
    assert _status("coral_zher", 0, 4, 1.0, p, 0, p, 4, None) == -5
    assert _status("coral_zher", 1, 4, 1.0, p, 1, p, 3, None) == -7
    assert _status("coral_dsyr", 1, 4, 1.0, p, 1, None, 4, None) == -6
    # syr2 / her2: uplo=1 n=2 alpha=3 x=4 incx=5 y=6 incy=7 a=8 lda=9
    assert _status("coral_dsyr2", 0, 4, 1.0, p, 1, p, 0, p, 4, None) == -7
    assert _status("coral_zher2", 0, 4, None, p, 1, p, 1, p, 4, None) == -3
    assert _status("coral_zher2", 0, 4, al, p, 1, p, 1, p, 2, None) == -9
    # quick returns (sger.rs:80, ssyr.rs:73, cher2.rs:77) need no device
    assert _status("coral_dger", 0, 4, 1.0, p, 1, p, 1, p, 1, None) == 0
    assert _status("coral_dger", 4, 4, 0.0, p, 1, p, 1, p, 4, None) == 0
    assert _status("coral_dsyr", 0, 0, 1.0, p, 1, p, 1, None) == 0
    assert _status("coral_zher2", 1, 4, zero, p, 1, p, 1, p, 4, None) == 0


def test_level1_rest_validation_codes():
    buf = np.zeros(64)
    p = buf.ctypes.data
    pvf = np.array([-1.0, 1.0, 0.0, 0.0, 1.0], np.float32)
    # copy/swap/rot/rotm: n=1 x=2 incx=3 y=4 incy=5
    assert _status("coral_dcopy", -1, p, 1, p, 1, None) == -1
    assert _status("coral_dcopy", 4, p, 0, p, 1, None) == -3
    assert _status("coral_zswap", 4, p, 1, p, 0, None) == -5
    assert _status("coral_drot", 4, None, 1, p, 1, 1.0, 0.0, None) == -2
    assert _status("coral_srotm", 4, p, 1, None, 1, pvf.ctypes.data, None) == -4
    assert _status("coral_drotm", 4, p, 1, p, 1, None, None) == -6
    # asum: n=1 x=2 incx=3 result=4; scal: n=1 alpha=2 x=3 incx=4
    assert _status("coral_dasum", 4, p, 0, p, None) == -3
    assert _status("coral_dzasum", 4, p, 1, None, None) == -4
    assert _status("coral_zdscal", 4, 2.0, p, 0, None) == -4
    assert _status("coral_csscal", 4, 2.0, None, 1, None) == -3
    # rotg: a=1 b=2 c=3 s=4; rotmg: d1=1 d2=2 x1=3 y1=4 param=5
    assert _status("coral_drotg", p, p, None, p, None) == -3
    assert _status("coral_srotmg", p, p, p, 1.0, None, None) == -5
    # n == 0 and flag == -2 return before any device call
    assert _status("coral_dcopy", 0, p, 1, p, 1, None) == 0
    pid = np.array([-2.0, 0, 0, 0, 0])
    assert _status("coral_drotm", 4, p, 1, p, 1, pid.ctypes.data, None) == 0


def test_quick_returns_need_no_device():
    # m == 0 / n == 0 return before any device call, as CORAL's quick returns do
    buf = np.zeros(4)
    p = buf.ctypes.data
    assert _status("coral_dgemm", 0, 0, 0, 4, 4, 1.0, p, 1, p, 4, 0.0, p, 1, None) == 0
    assert _status("coral_dgemv", 0, 0, 4, 1.0, p, 1, p, 1, 0.0, p, 1, None) == 0
    assert _status("coral_dtrsv", 1, 0, 1, 0, p, 1, p, 1, None) == 0


def test_no_device_is_reported_not_faked():
    lib = coral_amd.load()
    if lib.coral_init() == 0:
        pytest.skip("a GPU is present")
    buf = np.zeros(64)
    p = buf.ctypes.data
    st = _status("coral_dgemm", 0, 0, 4, 4, 4, 1.0, p, 4, p, 4, 0.0, p, 4, None)
    assert st == 2  # CORAL_ERR_NO_DEVICE: no CPU fallback
    assert b"device" in lib.coral_last_error()


def test_host_api_rejects_before_copy():
    buf = np.zeros(16, np.float32)
    st = _status("coralh_sgemm", 0, 0, 4, 4, 4, 1.0, buf.ctypes.data, 2, buf.ctypes.data, 4, 0.0,
                 buf.ctypes.data, 4)
    assert st in (-8, 2)  # -8 when validation precedes device init


def test_synthetic_generator_matches_reference_implementation():
    # pure-Python splitmix64 restatement pins the vectorised numpy generator
    from coral_amd import synthetic

    M = (1 << 64) - 1
    G = 0x9E3779B97F4A7C15

    def mix(z):
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        return z ^ (z >> 31)

    seed = synthetic.seed_for(2, "A")
    s = mix((seed + G) & M)
    h = [mix((s + (i + 1) * G) & M) for i in range(8)]
    assert list(synthetic.raw(seed, 8)) == h
    u = synthetic.uniform(seed, 8, np.float64)
    assert np.array_equal(u, np.array([2.0 * ((x >> 11) * 2.0 ** -53) - 1.0 for x in h]))
    uf = synthetic.uniform(seed, 8, np.float32)
    assert np.array_equal(uf, np.array([2.0 * ((x >> 40) * 2.0 ** -24) - 1.0 for x in h], np.float32))
    assert synthetic.uniform(seed, 100000).min() >= -1.0 and synthetic.uniform(seed, 100000).max() < 1.0
    z = synthetic.normal(seed, 200000)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1) < 0.01
