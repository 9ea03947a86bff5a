"""GPU parity for the level-2 rank updates (?ger, ?geru, ?gerc, ?syr, ?syr2, ?her,
?her2) through the C ABI.

The kernels reproduce CORAL's per-element arithmetic, so every case is compared
with the oracle (the CPU restatement of coral-aarch64/src/level2/sger.rs,
cgerc.rs, ssyr.rs, ssyr2.rs, cher.rs, cher2.rs and the axpys they call) bit for
bit, over the whole buffer: padding rows and the unstored triangle must come back
untouched.  Cases cover unit and non-unit strides (which change CORAL's rounding),
padded and unaligned leading dimensions, both triangles, ragged sizes around the
kernel's 256*EV-row and 32-column blocks, zero coefficients with -0.0 and
non-finite entries, and the closed-form rounding cases of the oracle's tests.
"""
import numpy as np
import pytest

from coral_amd import blas, device
from coral_amd._lib import call
from coral_amd.device import DeviceArray
from oracle import port

pytestmark = pytest.mark.gpu

DT = {"s": np.float32, "d": np.float64, "c": np.complex64, "z": np.complex128}
REAL = {"s": np.float32, "d": np.float64, "c": np.float32, "z": np.float64}
GER = ["sger", "dger", "cgeru", "cgerc", "zgeru", "zgerc"]
SYR = ["ssyr", "dsyr", "cher", "zher"]
SYR2 = ["ssyr2", "dsyr2", "cher2", "zher2"]


def rand(rng, n, dt):
    if np.issubdtype(dt, np.complexfloating):
        return (rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)).astype(dt)
    return rng.uniform(-1, 1, n).astype(dt)


def bits(a):
    return np.ascontiguousarray(a).view(np.uint8)


def alpha_for(name, rng):
    if name[0] in "cz" and name not in SYR:
        return (float(rng.uniform(-1.5, 1.5)), float(rng.uniform(-1.5, 1.5)))
    return float(rng.uniform(-1.5, 1.5))


def check_ger(name, m, n, lda, incx, incy, rng, off=0, alpha=None, x=None, y=None, a=None):
    dt = DT[name[0]]
    al = alpha_for(name, rng) if alpha is None else alpha
    if x is None:
        x = rand(rng, (m - 1) * incx + 1 if m else 1, dt)
    if y is None:
        y = rand(rng, (n - 1) * incy + 1 if n else 1, dt)
    if a is None:
        a = rand(rng, lda * n + off if n else 1 + off, dt)
    g, o = a.copy(), a.copy()
    getattr(blas, name)(m, n, al, x, incx, y, incy, g[off:], lda)
    getattr(port, name)(m, n, al, x, incx, y, incy, o[off:], lda)
    assert np.array_equal(bits(g), bits(o)), (name, m, n, lda, incx, incy, off)
    return g


def check_tri(name, uplo, n, lda, incx, incy, rng, off=0, alpha=None, x=None, y=None, a=None):
    dt = DT[name[0]]
    al = alpha_for(name, rng) if alpha is None else alpha
    if x is None:
        x = rand(rng, (n - 1) * incx + 1 if n else 1, dt)
    if y is None:
        y = rand(rng, (n - 1) * incy + 1 if n else 1, dt)
    if a is None:
        a = rand(rng, lda * n + off if n else 1 + off, dt)
    g, o = a.copy(), a.copy()
    if name in SYR:
        getattr(blas, name)(uplo, n, al, x, incx, g[off:], lda)
        getattr(port, name)(uplo, n, al, x, incx, o[off:], lda)
    else:
        getattr(blas, name)(uplo, n, al, x, incx, y, incy, g[off:], lda)
        getattr(port, name)(uplo, n, al, x, incx, y, incy, o[off:], lda)
    assert np.array_equal(bits(g), bits(o)), (name, uplo, n, lda, incx, incy, off)
    return g


@pytest.mark.parametrize("name", GER)
def test_ger_bit_exact(gpu_lib, name):
    rng = np.random.default_rng(70)
    for m, n, lda, incx, incy, off in ((1, 1, 1, 1, 1, 0), (7, 5, 7, 1, 1, 0), (300, 257, 303, 1, 1, 0),
                                       (300, 257, 300, 2, 3, 0), (1029, 515, 1031, 1, 1, 0),
                                       (513, 64, 513, 1, 2, 0), (64, 700, 64, 3, 1, 0), (1030, 33, 1030, 1, 1, 1),
                                       (2049, 40, 2050, 1, 1, 0)):
        check_ger(name, m, n, lda, incx, incy, rng, off)


@pytest.mark.parametrize("name", SYR + SYR2)
@pytest.mark.parametrize("uplo", [0, 1])
def test_syr_her_bit_exact(gpu_lib, name, uplo):
    rng = np.random.default_rng(71 + uplo)
    strides = [(1, 1), (2, 3), (1, 3), (2, 1)] if name in SYR2 else [(1, 1), (2, 1)]
    for n, lda, off in ((1, 1, 0), (7, 7, 0), (64, 67, 0), (257, 257, 0), (1000, 1003, 0), (1029, 1029, 1)):
        for incx, incy in strides:
            check_tri(name, uplo, n, lda, incx, incy, rng, off)


@pytest.mark.parametrize("name", GER + SYR + SYR2)
def test_zero_coefficients_signed_zeros_and_non_finite(gpu_lib, name):
    """Columns whose coefficient is zero are skipped as CORAL skips them: -0.0 in A
    survives and an infinite x element never reaches them."""
    rng = np.random.default_rng(72)
    p = name[0]
    dt = DT[p]
    n = 41
    for incx, incy in ((1, 1), (2, 3)):
        x = rand(rng, (n - 1) * incx + 1, dt)
        y = rand(rng, (n - 1) * incy + 1, dt)
        x[5 * incx] = 0
        y[7 * incy] = 0
        y[5 * incy] = 0
        x[11 * incx] = np.inf
        a = rand(rng, n * n, dt)
        for j in (5, 7):
            a[j * n:(j + 1) * n] = complex(-0.0, -0.0) if p in "cz" else -0.0
        if name in GER:
            check_ger(name, n, n, n, incx, incy, rng, x=x, y=y, a=a)
        else:
            for uplo in (0, 1):
                check_tri(name, uplo, n, n, incx, incy, rng, x=x, y=y, a=a)


F_TINY = {np.float32: (1 + 2.0 ** -13, 2.0 ** -26), np.float64: (1 + 2.0 ** -28, 2.0 ** -56)}
Q = {"s": 8, "d": 4, "c": 4, "z": 2}


@pytest.mark.parametrize("name", GER + SYR)
def test_rounding_closed_form(gpu_lib, name):
    """x = y = u, alpha = 1, A = -fl(u*u): an element gets 2^-26 (f32) / 2^-56 (f64)
    where CORAL fuses and 0 where it rounds twice."""
    p = name[0]
    dt, real = DT[p], REAL[p]
    u, tiny = F_TINY[real]
    u = real(u)
    uu = real(u * u)
    n = 21
    one = (1.0, 0.0) if (p in "cz" and name not in SYR) else 1.0
    for incx in (1, 2):
        x = np.zeros((n - 1) * incx + 1, dt)
        x[::incx] = u
        y = np.full(n, u, dt)
        a = np.full(n * n, -uu, dt)
        if name in GER:
            getattr(blas, name)(n, n, one, x, incx, y, 1, a, n)
            mode = "unit" if incx == 1 else ("fused" if p in "sd" else "plain")
        else:
            getattr(blas, name)(0, n, one, x, incx, a, n)
            mode = "unit" if incx == 1 else ("fused" if p in "sd" else "plain")
        A = np.real(a.reshape(n, n).T)
        for j in range(n):
            lo, hi = (0, n) if name in GER else (0, j + 1)
            for i in range(lo, hi):
                t, length = i - lo, hi - lo
                fused = mode == "fused" or (mode == "unit" and t < length // Q[p] * Q[p])
                assert A[i, j] == (tiny if fused else 0.0), (name, incx, i, j)


def test_device_pointer_form_matches_host_form(gpu_lib):
    """coral_dger / coral_zher2 on device buffers and a stream equal the host form."""
    rng = np.random.default_rng(73)
    m, n, lda = 1000, 300, 1003
    x, y, a = rand(rng, m, np.float64), rand(rng, n, np.float64), rand(rng, lda * n, np.float64)
    want = a.copy()
    blas.dger(m, n, 0.75, x, 1, y, 1, want, lda)
    dx, dy, da = DeviceArray.from_numpy(x), DeviceArray.from_numpy(y), DeviceArray.from_numpy(a)
    st = device.Stream()
    call("coral_dger", m, n, 0.75, dx.ptr, 1, dy.ptr, 1, da.ptr, lda, st.handle)
    st.synchronize()
    assert np.array_equal(bits(da.to_numpy()), bits(want))
    nn = 517
    xc, yc = rand(rng, nn, np.complex128), rand(rng, nn, np.complex128)
    ac = rand(rng, nn * nn, np.complex128)
    want = ac.copy()
    blas.zher2(1, nn, (0.5, -0.25), xc, 1, yc, 1, want, nn)
    dxc, dyc, dac = DeviceArray.from_numpy(xc), DeviceArray.from_numpy(yc), DeviceArray.from_numpy(ac)
    alpha = np.array([0.5, -0.25])
    call("coral_zher2", 1, nn, alpha.ctypes.data, dxc.ptr, 1, dyc.ptr, 1, dac.ptr, nn, st.handle)
    st.synchronize()
    assert np.array_equal(bits(dac.to_numpy()), bits(want))
    st.destroy()


def test_large_dger_bit_exact(gpu_lib):
    rng = np.random.default_rng(74)
    m, n = 4099, 3001
    check_ger("dger", m, n, m + 1, 1, 1, rng)
    check_tri("zher2", 1, 2051, 2051, 1, 1, rng)
