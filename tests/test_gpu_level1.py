"""GPU parity for dot / axpy / nrm2 / i?amax / scal through the C ABI.

i?amax is bit-exact (0-based first index, NaN skipped, |re|+|im|) against brute
force and the oracle; the float reductions are checked against correctly rounded
FP64 references with the reference tests' tolerances (dot_tests.rs:176-179).
"""
import os

import numpy as np
import pytest

from coral_amd import blas
from oracle import port, ref64

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _abs_sum(x, y):
    """sum |x_i| |y_i|: the scale of a dot product's rounding error."""
    return float(np.sum(np.abs(np.asarray(x).astype(np.complex128)) * np.abs(np.asarray(y).astype(np.complex128))))


def test_dot(gpu_lib):
    # |computed - exact| <= c * u * sqrt(n) * sum |x_i y_i| (random-walk accumulation bound)
    rng = np.random.default_rng(30)
    for n, inc in ((0, 1), (1, 1), (31, 1), (1000, 2), (100003, 1), (1 << 20, 1)):
        x = rng.uniform(-1, 1, (n - 1) * inc + 1 if n else 1)
        y = rng.uniform(-1, 1, (n - 1) * inc + 1 if n else 1)
        xv, yv = ref64.vec(x, n, inc), ref64.vec(y, n, inc)
        s = max(_abs_sum(xv, yv), 1e-300)
        r = max(1.0, np.sqrt(n))
        e = ref64.dot_exact(xv, yv)
        assert abs(blas.ddot(n, x, inc, y, inc) - e) <= 4 * 2.0 ** -53 * r * s + 1e-300
        xf, yf = x.astype(np.float32), y.astype(np.float32)
        ef = ref64.dot_exact(ref64.vec(xf, n, inc), ref64.vec(yf, n, inc))
        assert abs(blas.sdot(n, xf, inc, yf, inc) - ef) <= 4 * 2.0 ** -24 * r * s + 1e-30
        xc = (x + 1j * y).astype(np.complex128)
        yc = (y - 0.5j * x).astype(np.complex128)
        sc = max(_abs_sum(ref64.vec(xc, n, inc), ref64.vec(yc, n, inc)), 1e-300)
        for name, conj in (("zdotu", False), ("zdotc", True)):
            rr = getattr(blas, name)(n, xc, inc, yc, inc)
            e = ref64.dot_exact(ref64.vec(xc, n, inc), ref64.vec(yc, n, inc), conj=conj)
            assert abs(complex(*rr) - e) <= 8 * 2.0 ** -53 * r * sc + 1e-300, name
        for name, conj in (("cdotu", False), ("cdotc", True)):
            rr = getattr(blas, name)(n, xc.astype(np.complex64), inc, yc.astype(np.complex64), inc)
            e = ref64.dot_exact(ref64.vec(xc.astype(np.complex64), n, inc), ref64.vec(yc.astype(np.complex64), n, inc),
                                conj=conj)
            assert abs(complex(*rr) - e) <= 8 * 2.0 ** -24 * r * sc + 1e-30, name


@pytest.mark.parametrize("p", "sdcz")
def test_axpy(gpu_lib, p):
    """Bit-identical to the oracle: fused over the NEON body of a unit-stride call,
    twice rounded in its tail and in strided calls (saxpy.rs:63-126)."""
    rng = np.random.default_rng(31)
    dt = {"s": np.float32, "d": np.float64, "c": np.complex64, "z": np.complex128}[p]
    cplx = p in "cz"
    alpha = (0.5, -2.0) if cplx else 0.5
    for n, incx, incy, off in ((0, 1, 1, 0), (1, 1, 1, 0), (13, 1, 1, 0), (1000, 1, 1, 0), (1003, 1, 1, 0),
                               (1003, 1, 1, 1), (777, 3, 2, 0), (777, 1, 2, 0)):
        lx, ly = (n - 1) * incx + 1 + off if n else 1, (n - 1) * incy + 1 + off if n else 1
        x = rng.uniform(-1, 1, lx) + (1j * rng.uniform(-1, 1, lx) if cplx else 0)
        y = rng.uniform(-1, 1, ly) + (1j * rng.uniform(-1, 1, ly) if cplx else 0)
        x, y = x.astype(dt), y.astype(dt)
        out = y.copy()
        getattr(blas, f"{p}axpy")(n, alpha, x[off:], incx, out[off:], incy)  # off=1: unaligned, unit stride
        po = y.copy()
        getattr(port, f"{p}axpy")(n, alpha, x[off:], incx, po[off:], incy)
        assert np.array_equal(out.view(np.uint8), po.view(np.uint8)), (p, n, incx, incy, off)
    out = y.copy()
    getattr(blas, f"{p}axpy")(n, (0.0, 0.0) if cplx else 0.0, np.full_like(x, np.nan), incx, out, incy)
    assert np.array_equal(out, y)  # alpha == 0 quick return


def test_nrm2(gpu_lib):
    rng = np.random.default_rng(32)
    for scale in (1e-30, 1e-10, 1.0, 1e10, 1e30):
        x = (rng.uniform(-1, 1, 10001) * scale).astype(np.float32)
        e = ref64.nrm2(x)
        assert abs(blas.snrm2(10001, x, 1) - e) <= 1e-6 * e
    for scale in (1e-300, 1e-160, 1.0, 1e160, 1e300):
        x = rng.uniform(-1, 1, 10001) * scale
        e = ref64.nrm2(x)
        assert abs(blas.dnrm2(10001, x, 1) - e) <= 1e-14 * e
        xc = (x + 1j * x[::-1]).astype(np.complex128)
        e = ref64.nrm2(xc)
        assert abs(blas.dznrm2(10001, xc, 1) - e) <= 1e-14 * e
    x = rng.uniform(-1, 1, 3001)
    xs = ref64.vec(x, 1000, 3)
    assert abs(blas.dnrm2(1000, x, 3) - ref64.nrm2(xs)) <= 1e-14 * ref64.nrm2(xs)
    assert blas.dnrm2(0, x, 1) == 0.0
    assert np.isnan(blas.dnrm2(3, np.array([1.0, np.nan, 2.0]), 1))
    assert np.isinf(blas.dnrm2(3, np.array([1.0, np.inf, 2.0]), 1))


def test_iamax_bit_exact(gpu_lib):
    z = np.load(os.path.join(GOLD, "iamax_cases.npz"))
    assert blas.isamax(1024, z["x1"], 1) == int(z["expect1"])
    assert blas.isamax(int(z["n2"]), z["x2"], int(z["inc2"])) == int(z["expect2"])
    assert blas.isamax(0, np.ones(5, np.float32), 1) == 0
    x = np.array([1.0, -3.0, np.nan, 3.0, 2.0])
    assert blas.idamax(5, x, 1) == 1
    assert blas.idamax(3, np.full(3, np.nan), 1) == 0
    xc = np.array([1 + 2j, -2 - 1j, 3 + 0j, 0 - 3j], np.complex64)
    assert blas.icamax(4, xc, 1) == 0
    rng = np.random.default_rng(33)
    for n in (1, 2, 63, 64, 65, 1000, 100000, 1 << 20):
        v = rng.integers(-50, 51, n).astype(np.float64)  # many ties
        v[rng.integers(0, n, max(1, n // 100))] = np.nan
        assert blas.idamax(n, v, 1) == port.idamax(n, v, 1) == ref64.iamax(v, n, 1), n
        vf = v.astype(np.float32)
        assert blas.isamax(n, vf, 1) == port.isamax(n, vf, 1), n
        vc = (v + 1j * v[::-1]).astype(np.complex128)
        assert blas.izamax(n, vc, 1) == port.izamax(n, vc, 1), n
        assert blas.icamax(n, vc.astype(np.complex64), 1) == port.icamax(n, vc.astype(np.complex64), 1), n
        if n <= 100000:  # brute force (first index of the largest |x| or |re| + |im|) for every variant
            assert blas.isamax(n, vf, 1) == ref64.iamax(vf, n, 1), n
            assert blas.izamax(n, vc, 1) == ref64.iamax(vc, n, 1, cplx=True), n
            assert blas.icamax(n, vc.astype(np.complex64), 1) == ref64.iamax(vc.astype(np.complex64), n, 1,
                                                                             cplx=True), n
    v = rng.uniform(-1, 1, 3000)
    assert blas.idamax(1000, v, 3) == port.idamax(1000, v, 3)


def test_scal(gpu_lib):
    from coral_amd.device import DeviceArray
    from coral_amd._lib import call

    x = np.array([1.0, np.nan, -2.0, 4.0])
    d = DeviceArray.from_numpy(x)
    call("coral_dscal", 4, 0.0, d.ptr, 1, None)  # multiplies: NaN stays NaN (dscal.rs:30-66)
    r = d.to_numpy()
    assert r[0] == 0.0 and np.isnan(r[1]) and r[2] == 0.0


# ---------------------------------------------------------------- rest of level 1
DTP = {"s": np.float32, "d": np.float64, "c": np.complex64, "z": np.complex128}
RP = {"s": np.float32, "d": np.float64, "c": np.float32, "z": np.float64}
# (n, incx, incy, offset in elements): unit aligned with ragged tails, unaligned, strided
XY_CASES = [(0, 1, 1, 0), (1, 1, 1, 0), (13, 1, 1, 0), (1024, 1, 1, 0), (1023, 1, 1, 0), (1023, 1, 1, 1),
            (777, 3, 2, 0), (1, 5, 7, 0), (257, 2, 3, 0), (100003, 1, 1, 0)]


def _vec(rng, n, inc, off, dt):
    m = ((n - 1) * inc + 1 if n else 1) + off
    v = rng.uniform(-1, 1, m)
    if np.issubdtype(dt, np.complexfloating):
        v = v + 1j * rng.uniform(-1, 1, m)
    return v.astype(dt)


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint8)


@pytest.mark.parametrize("p", "sdcz")
def test_scal_all_forms_bit_exact(gpu_lib, p):
    """sscal/dscal (one rounding), cscal/zscal (CORAL's fused body and rounded tail),
    csscal/zdscal (real alpha) against the oracle, bit for bit."""
    rng = np.random.default_rng(40)
    forms = [(f"{p}scal", 0.7)] if p in "sd" else [(f"{p}scal", (0.75, -0.2)), (f"{p}scal", (0.0, 1.0)),
                                                    ({"c": "csscal", "z": "zdscal"}[p], -1.25)]
    for name, alpha in forms:
        for n, inc, _, off in XY_CASES:
            x = _vec(rng, n, inc, off, DTP[p])
            g, o = x.copy(), x.copy()
            getattr(blas, name)(n, alpha, g[off:], inc)
            getattr(port, name)(n, alpha, o[off:], inc)
            assert np.array_equal(_bits(g), _bits(o)), (name, n, inc, off)


@pytest.mark.parametrize("p", "sdcz")
def test_asum(gpu_lib, p):
    """Within n*eps of the oracle's CORAL-order sum (the GPU sums in a tree); exact on
    dyadic inputs, where every order gives the exact sum."""
    rng = np.random.default_rng(41)
    name = {"s": "sasum", "d": "dasum", "c": "scasum", "z": "dzasum"}[p]
    eps = float(np.finfo(RP[p]).eps)
    for n, inc, _, off in XY_CASES:
        x = _vec(rng, n, inc, off, DTP[p])
        got = float(getattr(blas, name)(n, x[off:], inc))
        want = float(getattr(port, name)(n, x[off:], inc))
        assert abs(got - want) <= 2 * max(n, 1) * eps * abs(want), (n, inc)
    v = rng.integers(-64, 65, 20000) / 8.0
    x = v.astype(DTP[p]) if p in "sd" else (v[0::2] + 1j * v[1::2]).astype(DTP[p])
    assert float(getattr(blas, name)(len(x), x, 1)) == float(np.abs(v).sum())


@pytest.mark.parametrize("p", "sdcz")
def test_copy_swap_bit_exact(gpu_lib, p):
    rng = np.random.default_rng(42)
    for n, incx, incy, off in XY_CASES:
        x, y = _vec(rng, n, incx, off, DTP[p]), _vec(rng, n, incy, off, DTP[p])
        gy, oy = y.copy(), y.copy()
        getattr(blas, f"{p}copy")(n, x[off:], incx, gy[off:], incy)
        getattr(port, f"{p}copy")(n, x[off:], incx, oy[off:], incy)
        assert np.array_equal(_bits(gy), _bits(oy)), ("copy", n, incx, incy, off)
        gx, gy, ox, oy = x.copy(), y.copy(), x.copy(), y.copy()
        getattr(blas, f"{p}swap")(n, gx[off:], incx, gy[off:], incy)
        getattr(port, f"{p}swap")(n, ox[off:], incx, oy[off:], incy)
        assert np.array_equal(_bits(gx), _bits(ox)) and np.array_equal(_bits(gy), _bits(oy)), ("swap", n)


@pytest.mark.parametrize("p", "sdcz")
def test_rot_bit_exact(gpu_lib, p):
    rng = np.random.default_rng(43)
    name = {"s": "srot", "d": "drot", "c": "csrot", "z": "zdrot"}[p]
    for n, incx, incy, off in XY_CASES:
        c, s = float(rng.uniform(-1, 1)), float(rng.uniform(-1, 1))
        x, y = _vec(rng, n, incx, off, DTP[p]), _vec(rng, n, incy, off, DTP[p])
        gx, gy, ox, oy = x.copy(), y.copy(), x.copy(), y.copy()
        getattr(blas, name)(n, gx[off:], incx, gy[off:], incy, c, s)
        getattr(port, name)(n, ox[off:], incx, oy[off:], incy, c, s)
        assert np.array_equal(_bits(gx), _bits(ox)) and np.array_equal(_bits(gy), _bits(oy)), (name, n, incx)


@pytest.mark.parametrize("p", "sd")
def test_rotm_bit_exact(gpu_lib, p):
    rng = np.random.default_rng(44)
    params = [[-2.0, 1.0, 2.0, 3.0, 4.0], [-1.0, 1.1, -0.4, 0.7, 0.9], [0.0, 0.0, -0.6, 0.8, 0.0],
              [1.0, 1.2, 0.0, 0.0, -0.75], [-1.0, -0.3, 0.5, -0.9, 0.4]]
    for param in params:
        for n, incx, incy, off in XY_CASES:
            x, y = _vec(rng, n, incx, off, DTP[p]), _vec(rng, n, incy, off, DTP[p])
            gx, gy, ox, oy = x.copy(), y.copy(), x.copy(), y.copy()
            getattr(blas, f"{p}rotm")(n, gx[off:], incx, gy[off:], incy, param)
            getattr(port, f"{p}rotm")(n, ox[off:], incx, oy[off:], incy, param)
            assert np.array_equal(_bits(gx), _bits(ox)) and np.array_equal(_bits(gy), _bits(oy)), (param, n)


@pytest.mark.parametrize("p", "sd")
def test_rotg_rotmg_bit_exact(gpu_lib, p):
    """The single-thread device constructions give the oracle's bits (IEEE division
    and square root on both sides), including the zero, rescaling and infinite cases."""
    rng = np.random.default_rng(45)
    pairs = [(3.5, -2.25), (0.0, 5.0), (-4.0, 0.0), (-1.25, -7.5), (12.0, -0.75), (0.0, 0.0), (2.5, 2.5),
             (1e-30, 2e-30), (7.0, -1e-3)] + [tuple(rng.uniform(-10, 10, 2)) for _ in range(40)]
    for a, b in pairs:
        got = getattr(blas, f"{p}rotg")(a, b)
        want = getattr(port, f"{p}rotg")(a, b)
        assert _bits(np.array(got)).tobytes() == _bits(np.array(want)).tobytes(), (a, b, got, want)
    quads = [(1.5, 2.0, 0.75, -1.25), (1e-3, 3e2, -4.5, 7.25), (5.0, 0.75, -2.0, 3.0), (1e4, 1e-5, 8.0, -6.0),
             (-1.0, 2.0, 3.0, 4.0), (4.0, 2.0, 3.0, 0.0), (1e-9, 1.0, 1.0, 1.0), (1e9, 1.0, 1.0, 1e-3),
             (1.0, 1e-12, 1.0, 1.0), (1.0, -2.0, 0.1, 5.0), (np.inf, 1.0, 1.0, 1.0)] + \
        [tuple(rng.uniform(0.01, 10, 4) * [1, 1, rng.choice([-1, 1]), rng.choice([-1, 1])]) for _ in range(40)]
    for q in quads:
        got = getattr(blas, f"{p}rotmg")(*q)
        want = getattr(port, f"{p}rotmg")(*q)
        gb = _bits(np.concatenate([np.array(got[:3]), got[3]]))
        wb = _bits(np.concatenate([np.array(want[:3]), want[3]]))
        assert gb.tobytes() == wb.tobytes(), (q, got, want)
