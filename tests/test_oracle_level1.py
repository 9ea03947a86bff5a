"""Pins the oracle's restatement of the rest of CORAL's level 1 (asum, copy, swap,
rot, rotm, rotg, rotmg, scal; oracle/coral_port.c) before it serves as the checker
of the GPU kernels.  CPU only.

* The reference tests' own cases (coral-aarch64/tests/level1/{asum,copy,swap,rot,
  scal}_tests.rs: sizes, strides, alphas, rotation parameters, rotg/rotmg inputs)
  against OpenBLAS through scipy.linalg.blas, the cblas those tests call, at their
  tolerances (copy and swap exactly).
* Closed-form rounding cases: with u = 1+2^-13 (f32) or 1+2^-28 (f64), fl(u*u)
  misses u*u by exactly tiny = 2^-26 (2^-56), so an update built as
  -fl(u*u) + u*u gives tiny when CORAL fuses it and 0 when it rounds the product
  first.  This pins which elements the NEON body of a unit-stride call covers
  (floor(len/Q)*Q) for rot, rotm and complex scal, and that strided calls and
  scalar tails round every product.
* Mathematical properties of rotg / rotmg: the rotation annihilates b, c^2 + s^2 = 1,
  z encodes (c, s) by the BLAS convention, and rotmg's H annihilates y1 with the
  weighted norm d1*x1^2 + d2*y1^2 preserved.
"""
from fractions import Fraction

import numpy as np
import pytest
import scipy.linalg.blas as sblas

from oracle import port
from tests.conftest import allclose_ref

DT = {"s": np.float32, "d": np.float64, "c": np.complex64, "z": np.complex128}
REAL = {"s": np.float32, "d": np.float64, "c": np.float32, "z": np.float64}
TOL = {"s": (1e-6, 1e-6), "d": (1e-12, 1e-12), "c": (1e-6, 1e-6), "z": (1e-12, 1e-12)}

# (n, incx, incy) of the reference tests (asum/copy/swap/rot_tests.rs)
REAL_CASES = [(1024, 1, 1), (777, 3, 2), (0, 1, 1), (1, 5, 7), (1536, 1, 1), (1023, 4, 3), (1, 6, 5)]
CPLX_CASES = [(800, 1, 1), (513, 2, 3), (0, 1, 1), (1, 7, 5), (640, 1, 1), (511, 3, 2), (1, 9, 4)]


def strided(rng, n, inc, dt):
    m = (n - 1) * inc + 1 if n else 1
    v = rng.uniform(-1, 1, m)
    if np.issubdtype(dt, np.complexfloating):
        v = v + 1j * rng.uniform(-1, 1, m)
    return v.astype(dt)


def cases(p):
    return CPLX_CASES if p in "cz" else REAL_CASES


@pytest.mark.parametrize("p", "sdcz")
def test_asum_reference_cases_vs_blas(p):
    rng = np.random.default_rng(1)
    name = {"s": "sasum", "d": "dasum", "c": "scasum", "z": "dzasum"}[p]
    for n, incx, _ in cases(p):
        x = strided(rng, n, incx, DT[p])
        got = getattr(port, name)(n, x, incx)
        want = getattr(sblas, name)(x[::incx][:n]) if n else 0.0
        assert allclose_ref(got, want, *TOL[p]), (name, n, incx)
    # dyadic values: every partial sum is exact, so the result is the exact sum
    v = rng.integers(-64, 65, 998) / 8.0
    x = v.astype(DT[p]) if p in "sd" else (v[0::2] + 1j * v[1::2]).astype(DT[p])
    exact = sum(Fraction(float(abs(t))) for t in v)
    assert Fraction(float(getattr(port, name)(len(x), x, 1))) == exact


@pytest.mark.parametrize("p", "sdcz")
def test_copy_swap_exact(p):
    rng = np.random.default_rng(2)
    for n, incx, incy in cases(p):
        x, y = strided(rng, n, incx, DT[p]), strided(rng, n, incy, DT[p])
        y2 = y.copy()
        getattr(port, f"{p}copy")(n, x, incx, y2, incy)
        want = y.copy()
        if n:
            want[::incy][:n] = x[::incx][:n]
        assert np.array_equal(y2, want)
        x3, y3 = x.copy(), y.copy()
        getattr(port, f"{p}swap")(n, x3, incx, y3, incy)
        wx, wy = x.copy(), y.copy()
        if n:
            wx[::incx][:n], wy[::incy][:n] = y[::incy][:n], x[::incx][:n]
        assert np.array_equal(x3, wx) and np.array_equal(y3, wy)


ROT = {"s": "srot", "d": "drot", "c": "csrot", "z": "zdrot"}


@pytest.mark.parametrize("p", "sdcz")
def test_rot_reference_cases_vs_blas(p):
    rng = np.random.default_rng(3)
    c, s = 0.8, 0.6
    for n, incx, incy in cases(p):
        x, y = strided(rng, n, incx, DT[p]), strided(rng, n, incy, DT[p])
        gx, gy = x.copy(), y.copy()
        getattr(port, ROT[p])(n, gx, incx, gy, incy, c, s)
        if n:
            wx, wy = getattr(sblas, ROT[p])(x[::incx][:n].copy(), y[::incy][:n].copy(), c, s)
            assert allclose_ref(gx[::incx][:n], wx, *TOL[p]) and allclose_ref(gy[::incy][:n], wy, *TOL[p])
        # elements between the strided ones are untouched
        mask = np.ones(len(x), bool)
        mask[::incx] = False
        assert np.array_equal(gx[mask], x[mask])


F_TINY = {np.float32: (1 + 2.0 ** -13, 2.0 ** -26), np.float64: (1 + 2.0 ** -28, 2.0 ** -56)}
Q_ROT = {"s": 4, "d": 2, "c": 4, "z": 2}  # real lanes per NEON step (complex: of the 2n floats)


def _u(p):
    real = REAL[p]
    u, tiny = F_TINY[real]
    u = real(u)
    uu = real(u * u)
    assert Fraction(float(u)) ** 2 - Fraction(float(uu)) == Fraction(tiny)
    return u, uu, tiny


def _fused_mask(n, unit, q, w=1):
    """Elements covered by the NEON body: the first floor(w*n/q)*q real lanes."""
    nf = (w * n) // q * q // w if unit else 0
    return np.arange(n) < nf


@pytest.mark.parametrize("p", "sdcz")
@pytest.mark.parametrize("n", [1, 3, 5, 13])
def test_rot_rounding(p, n):
    u, uu, tiny = _u(p)
    dt, w = DT[p], (2 if p in "cz" else 1)
    for inc in (1, 2):
        fused = _fused_mask(n, inc == 1, Q_ROT[p], w)
        # x' = c*x + s*y with c = 1, s = u, x = -fl(u*u), y = u
        x = np.full((n - 1) * inc + 1, -uu, dt)
        y = np.full((n - 1) * inc + 1, u, dt)
        getattr(port, ROT[p])(n, x, inc, y, inc, 1.0, u)
        assert np.array_equal(np.real(x[::inc]), np.where(fused, tiny, 0.0).astype(REAL[p])), (p, inc)
        # y' = c*y - s*x with c = 1, s = u, x = u, y = fl(u*u): real vectors fuse it
        x = np.full((n - 1) * inc + 1, u, dt)
        y = np.full((n - 1) * inc + 1, uu, dt)
        getattr(port, ROT[p])(n, x, inc, y, inc, 1.0, u)
        yf = fused if p in "sd" else np.zeros(n, bool)
        assert np.array_equal(np.real(y[::inc]), np.where(yf, -tiny, 0.0).astype(REAL[p])), (p, inc)


ROTM_PARAMS = [[-2.0, 0.0, 0.0, 0.0, 0.0], [-1.0, 1.1, -0.4, 0.7, 0.9], [0.0, 0.0, -0.6, 0.8, 0.0],
               [1.0, 1.2, 0.0, 0.0, -0.75], [-1.0, -0.3, 0.5, -0.9, 0.4]]


@pytest.mark.parametrize("p", "sd")
def test_rotm_reference_cases_vs_blas(p):
    rng = np.random.default_rng(4)
    for (n, incx, incy), param in zip([(256, 1, 1), (320, 1, 1), (300, 1, 1), (300, 1, 1), (257, 2, 3)],
                                      ROTM_PARAMS):
        x, y = strided(rng, n, incx, DT[p]), strided(rng, n, incy, DT[p])
        gx, gy = x.copy(), y.copy()
        getattr(port, f"{p}rotm")(n, gx, incx, gy, incy, param)
        wx, wy = getattr(sblas, f"{p}rotm")(x[::incx][:n].copy(), y[::incy][:n].copy(),
                                             np.array(param, REAL[p]))
        assert allclose_ref(gx[::incx][:n], wx, *TOL[p]) and allclose_ref(gy[::incy][:n], wy, *TOL[p]), param


@pytest.mark.parametrize("p", "sd")
@pytest.mark.parametrize("n", [1, 3, 5, 13])
def test_rotm_rounding(p, n):
    u, uu, tiny = _u(p)
    dt = DT[p]
    for inc in (1, 2):
        fused = _fused_mask(n, inc == 1, Q_ROT[p])
        want = np.where(fused, tiny, 0.0).astype(dt)
        m = (n - 1) * inc + 1
        # flag < 0: x' = h11*x + h12*y with h11 = 1, h12 = u; y' = h21*x + h22*y with h21 = u, h22 = 1
        x, y = np.full(m, -uu, dt), np.full(m, u, dt)
        getattr(port, f"{p}rotm")(n, x, inc, y, inc, [-1.0, 1.0, 0.0, u, 1.0])
        assert np.array_equal(x[::inc], want), ("flag<0 x", inc)
        x, y = np.full(m, u, dt), np.full(m, -uu, dt)
        getattr(port, f"{p}rotm")(n, x, inc, y, inc, [-1.0, 1.0, u, 0.0, 1.0])
        assert np.array_equal(y[::inc], want), ("flag<0 y", inc)
        # flag == 0: x' = x + h12*y, y' = h21*x + y
        x, y = np.full(m, -uu, dt), np.full(m, u, dt)
        getattr(port, f"{p}rotm")(n, x, inc, y, inc, [0.0, 0.0, 0.0, u, 0.0])
        assert np.array_equal(x[::inc], want), ("flag0 x", inc)
        x, y = np.full(m, u, dt), np.full(m, -uu, dt)
        getattr(port, f"{p}rotm")(n, x, inc, y, inc, [0.0, 0.0, u, 0.0, 0.0])
        assert np.array_equal(y[::inc], want), ("flag0 y", inc)
        # flag > 0: x' = h11*x + y; y' = h22*y - x is never fused
        x, y = np.full(m, u, dt), np.full(m, -uu, dt)
        getattr(port, f"{p}rotm")(n, x, inc, y, inc, [1.0, u, 0.0, 0.0, 1.0])
        assert np.array_equal(x[::inc], want), ("flag>0 x", inc)


ROTG_INPUTS = {"s": [(3.5, -2.25), (0.0, 5.0), (-4.0, 0.0), (-1.25, -7.5)],
               "d": [(12.0, -0.75), (0.0, 3.0), (-5.0, 0.0), (-2.0, 9.0)]}


@pytest.mark.parametrize("p", "sd")
def test_rotg_vs_blas_and_properties(p):
    tol = 1e-6 if p == "s" else 1e-12
    extra = [(1.0, 1.0), (-3.0, 4.0), (1e-30, 2e-30), (7.0, -1e-3), (0.0, 0.0), (2.5, 2.5)]
    for a, b in ROTG_INPUTS[p] + extra:
        r, z, c, s = getattr(port, f"{p}rotg")(a, b)
        if a == 0 and b == 0:
            assert (r, z, c, s) == (0, 0, 1, 0)
            continue
        wc, ws = getattr(sblas, f"{p}rotg")(a, b)
        assert allclose_ref([c, s], [wc, ws], tol, tol), (a, b)
        assert abs(float(c) ** 2 + float(s) ** 2 - 1) < 10 * tol
        assert abs(float(c) * a + float(s) * b - float(r)) <= 10 * tol * max(abs(a), abs(b))
        assert abs(-float(s) * a + float(c) * b) <= 10 * tol * max(abs(a), abs(b))
        # z encodes (c, s): |z| < 1 -> s = z; |z| > 1 -> c = 1/z; z == 1 -> c = 0
        z = float(z)
        if abs(z) < 1:
            assert abs(float(s) - z) <= tol
        elif z == 1 and float(c) == 0:
            pass
        else:
            assert abs(float(c) - 1 / z) <= 10 * tol


ROTMG_INPUTS = {"s": [(1.5, 2.0, 0.75, -1.25), (1e-3, 3e2, -4.5, 7.25)],
                "d": [(5.0, 0.75, -2.0, 3.0), (1e4, 1e-5, 8.0, -6.0)]}


def _h(param):
    flag, h11, h21, h12, h22 = (float(v) for v in param)
    if flag < 0:
        return np.array([[h11, h12], [h21, h22]])
    if flag == 0:
        return np.array([[1.0, h12], [h21, 1.0]])
    return np.array([[h11, 1.0], [-1.0, h22]])


@pytest.mark.parametrize("p", "sd")
def test_rotmg_vs_blas_and_properties(p):
    tol = 1e-5 if p == "s" else 1e-12
    extra = [(1.0, 1.0, 1.0, 1.0), (2.0, 3.0, 0.5, 4.0), (1e-9, 1.0, 1.0, 1.0), (1e9, 1.0, 1.0, 1e-3),
             (1.0, 1e-12, 1.0, 1.0), (4.0, 2.0, 3.0, 0.0), (-1.0, 2.0, 3.0, 4.0), (1.0, -2.0, 0.1, 5.0)]
    for d1, d2, x1, y1 in ROTMG_INPUTS[p] + extra:
        nd1, nd2, nx1, param = getattr(port, f"{p}rotmg")(d1, d2, x1, y1)
        want = getattr(sblas, f"{p}rotmg")(d1, d2, x1, y1)
        assert allclose_ref(param, want, tol, tol), (d1, d2, x1, y1, param, want)
        flag = float(param[0])
        if flag == -2 or (flag == -1 and float(nd1) == 0 and float(nd2) == 0):
            continue
        H = _h(param)
        v = H @ np.array([x1, y1])
        assert abs(v[1]) <= 10 * tol * max(abs(x1), abs(y1), 1) * max(1, np.abs(H).max())
        assert abs(v[0] - float(nx1)) <= 10 * tol * max(1, abs(v[0]))
        lhs = float(nd1) * v[0] ** 2
        rhs = d1 * x1 ** 2 + d2 * y1 ** 2
        assert abs(lhs - rhs) <= 100 * tol * abs(rhs), (d1, d2, x1, y1)


def test_rotmg_infinite_d1_terminates():
    nd1, nd2, nx1, param = port.drotmg(np.inf, 1.0, 1.0, 1.0)
    assert np.isfinite(param[0])


SCAL_CASES = [(1024, 3.1415926, 1), (777, -0.75, 3), (4096, 0.0, 1), (257, 1.0, 1), (0, 1.23, 1),
              (1536, -2.5, 1), (1023, 1.0 / 3.0, 4)]
CSCAL_CASES = [(800, (0.75, -0.20), 1), (513, (-1.1, 0.4), 3), (256, (0.0, 0.0), 1), (300, (2.0, 0.0), 1),
               (300, (0.0, 1.0), 1), (640, (1.25, -0.75), 1), (511, (0.5, 0.25), 2)]


@pytest.mark.parametrize("p", "sdcz")
def test_scal_reference_cases(p):
    rng = np.random.default_rng(5)
    if p in "sd":
        for n, alpha, inc in SCAL_CASES:
            x = strided(rng, n, inc, DT[p])
            g = x.copy()
            getattr(port, f"{p}scal")(n, alpha, g, inc)
            want = x.copy()
            if n:
                want[::inc][:n] = x[::inc][:n] * REAL[p](alpha)  # one IEEE rounding
            assert np.array_equal(g, want)
        return
    for n, alpha, inc in CSCAL_CASES:
        x = strided(rng, n, inc, DT[p])
        g = x.copy()
        getattr(port, f"{p}scal")(n, alpha, g, inc)
        if n:
            want = getattr(sblas, f"{p}scal")(complex(*alpha), x[::inc][:n].copy())
            assert allclose_ref(g[::inc][:n], want, *TOL[p]), (n, alpha, inc)
        # real alpha on complex x: each part one rounding
        g = x.copy()
        getattr(port, {"c": "csscal", "z": "zdscal"}[p])(n, alpha[0], g, inc)
        want = x.copy()
        if n:
            want[::inc][:n] = (x[::inc][:n].real * REAL[p](alpha[0])) + 1j * (x[::inc][:n].imag * REAL[p](alpha[0]))
        assert np.array_equal(g, want)


@pytest.mark.parametrize("p", "cz")
@pytest.mark.parametrize("n", [1, 3, 5, 13])
def test_cscal_rounding(p, n):
    """re' = ar*xr - ai*xi with ar = 1, xr = -fl(u*u), ai = u, xi = -u: tiny where fused."""
    u, uu, tiny = _u(p)
    q = 4 if p == "c" else 2
    for inc in (1, 2):
        fused = _fused_mask(n, inc == 1, q)
        x = np.full((n - 1) * inc + 1, complex(-uu, -u), DT[p])
        getattr(port, f"{p}scal")(n, (1.0, float(u)), x, inc)
        assert np.array_equal(x[::inc].real, np.where(fused, tiny, 0.0).astype(REAL[p])), (p, inc)
