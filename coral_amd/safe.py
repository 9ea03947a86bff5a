"""Checked front end: CORAL's ``coral-safe`` view API over the gfx950 library.

``coral-safe`` wraps single-precision BLAS routines in checked buffer views:
``VectorRef``/``VectorMut`` (data, n, stride, offset) and ``MatrixRef``/``MatrixMut``
(data, n_rows, n_cols, lda, offset), validated when the view is built, with a
``BufferError`` describing what is wrong (coral-safe/src/types.rs:37-367,
coral-safe/src/errors.rs:3-25).  Routines then take views instead of raw
(slice, length, stride) triples, so a routine can never read outside its buffer.

This module keeps those names, argument orders and error values.  A view holds a
1-D numpy array; every routine validates its views' lengths against each other
(coral-safe checks these with ``debug_assert!`` only, so a release build would
run on with inconsistent lengths; here they raise ``ValueError`` always) and then
calls the library's synchronous host-pointer entry point on the window that
starts at the view's offset.  The arithmetic is the library's: the element-wise
routines (axpy, scal, copy, swap, rot, rotm) reproduce coral-aarch64's rounding
bit for bit; the reductions, GEMV, GER and TRSV agree with coral-safe to the
tolerances of coral-safe's own tests (RTOL 1e-5, ATOL 1e-6, tests/common.rs:8-9).
"""
from __future__ import annotations

from enum import Enum

import numpy as np

from . import blas

_USIZE_MAX = (1 << 64) - 1


# ---------------------------------------------------------------- errors
class BufferError(ValueError):
    """A view that does not fit its buffer (coral-safe/src/errors.rs:3-25).

    Four kinds, compared by kind and fields as the Rust enum derives ``PartialEq``:
    ``BufferError.ZeroStride()``, ``BufferError.OutOfBounds(required, len)``,
    ``BufferError.BadOffset(offset, len)``, ``BufferError.InvalidLda(lda, n_rows)``.
    ``str()`` gives the reference's ``Display`` text.
    """

    kind = ""
    fields: tuple = ()

    def __init__(self, **values):
        self.values = values
        super().__init__(self._message())

    def _message(self) -> str:  # pragma: no cover - overridden
        return self.kind

    def __eq__(self, other):
        return type(other) is type(self) and other.values == self.values

    def __hash__(self):
        return hash((self.kind, tuple(sorted(self.values.items()))))

    def __repr__(self):
        inner = ", ".join(f"{k}: {v}" for k, v in self.values.items())
        return f"BufferError::{self.kind}" + (f" {{ {inner} }}" if inner else "")

    def __getattr__(self, name):
        values = self.__dict__.get("values", {})
        if name in values:
            return values[name]
        raise AttributeError(name)


class _ZeroStride(BufferError):
    kind = "ZeroStride"

    def __init__(self):
        super().__init__()

    def _message(self):
        return "stride must be nonzero"


class _OutOfBounds(BufferError):
    kind = "OutOfBounds"

    def __init__(self, required: int, len: int):  # noqa: A002 - the reference's field name
        super().__init__(required=required, len=len)

    def _message(self):
        return f"buffer too small: required {self.values['required']}, have {self.values['len']}"


class _BadOffset(BufferError):
    kind = "BadOffset"

    def __init__(self, offset: int, len: int):  # noqa: A002
        super().__init__(offset=offset, len=len)

    def _message(self):
        return f"offset {self.values['offset']} out of range for buffer of len {self.values['len']}"


class _InvalidLda(BufferError):
    kind = "InvalidLda"

    def __init__(self, lda: int, n_rows: int):
        super().__init__(lda=lda, n_rows=n_rows)

    def _message(self):
        return f"leading dimension {self.values['lda']} must exceed num rows {self.values['n_rows']}"


BufferError.ZeroStride = _ZeroStride
BufferError.OutOfBounds = _OutOfBounds
BufferError.BadOffset = _BadOffset
BufferError.InvalidLda = _InvalidLda


# ---------------------------------------------------------------- op selectors
class CoralTranspose(Enum):  # coral-safe/src/types.rs:3-9
    NoTrans = 0
    Trans = 1


class CoralTriangular(Enum):  # coral-safe/src/types.rs:11-17
    Upper = 0
    Lower = 1


class CoralDiagonal(Enum):  # coral-safe/src/types.rs:19-34
    Unit = 0
    NonUnit = 1

    def is_unit(self) -> bool:
        return self is CoralDiagonal.Unit


# ---------------------------------------------------------------- views
def _usize(v, name):
    if isinstance(v, bool) or not isinstance(v, (int, np.integer)):
        raise TypeError(f"{name} must be an unsigned integer, got {type(v).__name__}")
    v = int(v)
    if v < 0 or v > _USIZE_MAX:
        raise ValueError(f"{name} must lie in [0, 2^64), got {v}")
    return v


def _sat_add(a, b):
    return min(a + b, _USIZE_MAX)


def _sat_mul(a, b):
    return min(a * b, _USIZE_MAX)


def _array(data, mutable):
    a = data if isinstance(data, np.ndarray) else np.asarray(data)
    if a.ndim != 1:
        raise ValueError("a view's buffer must be a 1-D array")
    if mutable and not a.flags["WRITEABLE"]:
        raise ValueError("a mutable view needs a writeable buffer")
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("a view's buffer must be contiguous (strides belong to the view)")
    return a


class _Vector:
    _mutable = False

    def __init__(self, data, n, stride, offset):
        """coral-safe/src/types.rs:72-103 (VectorRef::new), :128-160 (VectorMut::new)."""
        data = _array(data, self._mutable)
        n, stride, offset = _usize(n, "n"), _usize(stride, "stride"), _usize(offset, "offset")
        if n != 0:
            if stride == 0:
                raise BufferError.ZeroStride()
            required = _sat_add(_sat_add(_sat_mul(n - 1, stride), offset), 1)
            if required > data.shape[0]:
                raise BufferError.OutOfBounds(required=required, len=data.shape[0])
        self._data, self._n, self._stride, self._offset = data, n, stride, offset

    @classmethod
    def new(cls, data, n, stride, offset):
        return cls(data, n, stride, offset)

    def n(self) -> int:
        return self._n

    def stride(self) -> int:
        return self._stride

    def offset(self) -> int:
        return self._offset

    def as_slice(self):
        return self._data

    def contiguous_slice(self):
        """The logical window when the stride is 1, else ``None`` (types.rs:114-117)."""
        return self._data[self._offset:self._offset + self._n] if self._stride == 1 else None

    def compare_n(self, n: int) -> bool:
        return self._n == n

    def _window(self):
        """Buffer from the first logical element on: what the library's pointer sees."""
        return self._data[self._offset:] if self._n else self._data[:0]

    def __repr__(self):
        return f"{type(self).__name__}(n={self._n}, stride={self._stride}, offset={self._offset}, len={self._data.shape[0]})"


class VectorRef(_Vector):
    """Immutable strided vector view (coral-safe/src/types.rs:37-44)."""


class VectorMut(_Vector):
    """Mutable strided vector view (coral-safe/src/types.rs:47-54)."""

    _mutable = True

    def as_slice_mut(self):
        return self._data

    def contiguous_slice_mut(self):
        return self.contiguous_slice()


class _Matrix:
    _mutable = False

    def __init__(self, data, n_rows, n_cols, lda, offset):
        """coral-safe/src/types.rs:183-222 (MatrixRef::new), :262-301 (MatrixMut::new)."""
        data = _array(data, self._mutable)
        n_rows, n_cols = _usize(n_rows, "n_rows"), _usize(n_cols, "n_cols")
        lda, offset = _usize(lda, "lda"), _usize(offset, "offset")
        if n_rows != 0 and n_cols != 0:
            if lda == 0:
                raise BufferError.ZeroStride()
            if lda < n_rows:
                raise BufferError.InvalidLda(lda=lda, n_rows=n_rows)
            required = _sat_add(_sat_add(_sat_mul(n_cols - 1, lda), n_rows), offset)
            if required > data.shape[0]:
                raise BufferError.OutOfBounds(required=required, len=data.shape[0])
        self._data, self._rows, self._cols, self._lda, self._offset = data, n_rows, n_cols, lda, offset

    @classmethod
    def new(cls, data, n_rows, n_cols, lda, offset):
        return cls(data, n_rows, n_cols, lda, offset)

    def n_rows(self) -> int:
        return self._rows

    def n_cols(self) -> int:
        return self._cols

    def lda(self) -> int:
        return self._lda

    def offset(self) -> int:
        return self._offset

    def as_slice(self):
        return self._data

    def contiguous_slice(self):
        """The packed window when lda == n_rows, else ``None`` (types.rs:236-244)."""
        if self._lda != self._rows:
            return None
        if self._cols == 0:
            return self._data[self._offset:self._offset]
        return self._data[self._offset:self._offset + (self._cols - 1) * self._lda + self._rows]

    def compare_m_n(self) -> bool:
        return self._rows == self._cols

    def _window(self):
        return self._data[self._offset:] if self._rows and self._cols else self._data[:0]

    def __repr__(self):
        return (f"{type(self).__name__}(n_rows={self._rows}, n_cols={self._cols}, lda={self._lda}, "
                f"offset={self._offset}, len={self._data.shape[0]})")


class MatrixRef(_Matrix):
    """Immutable column-major matrix view (coral-safe/src/types.rs:57-64)."""


class MatrixMut(_Matrix):
    """Mutable column-major matrix view (coral-safe/src/types.rs:67-74)."""

    _mutable = True

    def as_slice_mut(self):
        return self._data

    def contiguous_slice_mut(self):
        return self.contiguous_slice()


# ---------------------------------------------------------------- checks
def _f32(*views):
    for v in views:
        if not isinstance(v, (_Vector, _Matrix)):
            raise TypeError(f"expected a coral-safe view, got {type(v).__name__}")
        if v.as_slice().dtype != np.float32:
            raise TypeError(f"coral-safe routines are single precision; got a {v.as_slice().dtype} view")


def _need(cond, msg):
    """coral-safe's debug_assert! checks, enforced in every build."""
    if not cond:
        raise ValueError(msg)


def _kind(v, cls, what):
    if not isinstance(v, cls):
        raise TypeError(f"{what} must be a {cls.__name__}, got {type(v).__name__}")


def _n_eq(x, y):  # debug_assert_n_eq!, coral-safe/src/types.rs:357-366
    _need(x.compare_n(y.n()), "number of logical elements must be equal")


def _inc(v):
    """The library's stride argument: any positive value when the view is empty."""
    return v.stride() if v.n() else 1


# ---------------------------------------------------------------- level 1
def saxpy(alpha, x: VectorRef, y: VectorMut) -> None:
    """y := alpha*x + y (coral-safe/src/level1/saxpy.rs:27-81)."""
    _kind(x, _Vector, "x"), _kind(y, VectorMut, "y"), _f32(x, y), _n_eq(x, y)
    blas.saxpy(x.n(), float(alpha), x._window(), _inc(x), y._window(), _inc(y))


def sscal(alpha, x: VectorMut) -> None:
    """x := alpha*x (coral-safe/src/level1/sscal.rs:25-65)."""
    _kind(x, VectorMut, "x"), _f32(x)
    blas.sscal(x.n(), float(alpha), x._window(), _inc(x))


def sasum(x: VectorRef) -> float:
    """sum |x_i| (coral-safe/src/level1/sasum.rs:26-67); 0 for n == 0."""
    _kind(x, _Vector, "x"), _f32(x)
    return blas.sasum(x.n(), x._window(), _inc(x))


def snrm2(x: VectorRef) -> float:
    """Euclidean norm (coral-safe/src/level1/snrm2.rs:25-62); 0 for n == 0."""
    _kind(x, _Vector, "x"), _f32(x)
    return blas.snrm2(x.n(), x._window(), _inc(x))


def scopy(x: VectorRef, y: VectorMut) -> None:
    """y := x (coral-safe/src/level1/scopy.rs:26-60)."""
    _kind(x, _Vector, "x"), _kind(y, VectorMut, "y"), _f32(x, y), _n_eq(x, y)
    blas.scopy(x.n(), x._window(), _inc(x), y._window(), _inc(y))


def sswap(x: VectorMut, y: VectorMut) -> None:
    """x <-> y (coral-safe/src/level1/sswap.rs:24-59)."""
    _kind(x, VectorMut, "x"), _kind(y, VectorMut, "y"), _f32(x, y), _n_eq(x, y)
    blas.sswap(x.n(), x._window(), _inc(x), y._window(), _inc(y))


def sdot(x: VectorRef, y: VectorRef) -> float:
    """x . y (coral-safe/src/level1/sdot.rs:28-82); 0 for n == 0."""
    _kind(x, _Vector, "x"), _kind(y, _Vector, "y"), _f32(x, y), _n_eq(x, y)
    return blas.sdot(x.n(), x._window(), _inc(x), y._window(), _inc(y))


def isamax(x: VectorRef) -> int:
    """0-based index of the first max |x_i| (coral-safe/src/level1/isamax.rs:27-95); 0 for n == 0."""
    _kind(x, _Vector, "x"), _f32(x)
    return blas.isamax(x.n(), x._window(), _inc(x))


def srot(x: VectorMut, y: VectorMut, c, s) -> None:
    """Plane rotation (x, y) := (c x + s y, c y - s x) (coral-safe/src/level1/srot.rs:32-104)."""
    _kind(x, VectorMut, "x"), _kind(y, VectorMut, "y"), _f32(x, y), _n_eq(x, y)
    blas.srot(x.n(), x._window(), _inc(x), y._window(), _inc(y), float(c), float(s))


def srotm(x: VectorMut, y: VectorMut, param) -> None:
    """Modified rotation with CORAL's five-element ``param`` (coral-safe/src/level1/srotm.rs:94-244)."""
    _kind(x, VectorMut, "x"), _kind(y, VectorMut, "y"), _f32(x, y), _n_eq(x, y)
    p = np.asarray(param, np.float32)
    _need(p.shape == (5,), "param must hold 5 values")
    blas.srotm(x.n(), x._window(), _inc(x), y._window(), _inc(y), p)


def _cell(v, name):
    """A ``&mut f32`` argument: a writeable one-element float32 array."""
    if not (isinstance(v, np.ndarray) and v.dtype == np.float32 and v.size == 1 and v.flags["WRITEABLE"]):
        raise TypeError(f"{name} must be a writeable one-element float32 array (the &mut f32 of the reference)")
    return v.reshape(-1)


def srotg(a, b, c, s) -> None:
    """Givens rotation setup, in place on four one-element arrays (coral-safe/src/level1/srotg.rs:34-81).

    On return a = r, b = z, and c, s hold the rotation.
    """
    ca, cb, cc, cs = (_cell(v, n) for v, n in ((a, "a"), (b, "b"), (c, "c"), (s, "s")))
    r, z, cv, sv = blas.srotg(float(ca[0]), float(cb[0]))
    ca[0], cb[0], cc[0], cs[0] = r, z, cv, sv


def srotmg(sd1, sd2, sx1, sy1, param) -> None:
    """Modified Givens setup, in place (coral-safe/src/level1/srotmg.rs:32-185).

    ``sd1``, ``sd2``, ``sx1`` are one-element float32 arrays, ``param`` a writeable
    float32 array of 5 (the reference's ``&mut [f32; 5]``).
    """
    c1, c2, c3 = _cell(sd1, "sd1"), _cell(sd2, "sd2"), _cell(sx1, "sx1")
    if not (isinstance(param, np.ndarray) and param.dtype == np.float32 and param.shape == (5,)
            and param.flags["WRITEABLE"]):
        raise TypeError("param must be a writeable float32 array of 5 (the &mut [f32; 5] of the reference)")
    d1, d2, x1, p = blas.srotmg(float(c1[0]), float(c2[0]), float(c3[0]), float(sy1))
    c1[0], c2[0], c3[0] = d1, d2, x1
    param[:] = p


# ---------------------------------------------------------------- fused (level 1.5)
def saxpyf(a: MatrixRef, x: VectorRef, y: VectorMut) -> None:
    """y := y + A x over all columns at once (coral-safe/src/fused/saxpyf.rs:1-10, :150-202)."""
    _kind(a, _Matrix, "a"), _kind(x, _Vector, "x"), _kind(y, VectorMut, "y"), _f32(a, x, y)
    _need(x.n() == a.n_cols(), "logical length of x must equal n_cols")
    _need(y.n() == a.n_rows(), "logical length of y must equal n_rows")
    blas.sgemv(0, a.n_rows(), a.n_cols(), 1.0, a._window(), max(a.lda(), 1), x._window(), _inc(x), 1.0,
               y._window(), _inc(y))


def sdotf(a: MatrixRef, x: VectorRef, y: VectorMut) -> None:
    """y := y + A^T x, every column's dot product in one sweep (coral-safe/src/fused/sdotf.rs:1-10, :150-203)."""
    _kind(a, _Matrix, "a"), _kind(x, _Vector, "x"), _kind(y, VectorMut, "y"), _f32(a, x, y)
    _need(x.n() == a.n_rows(), "logical length of x must equal n_rows")
    _need(y.n() == a.n_cols(), "logical length of y must equal n_cols")
    blas.sgemv(1, a.n_rows(), a.n_cols(), 1.0, a._window(), max(a.lda(), 1), x._window(), _inc(x), 1.0,
               y._window(), _inc(y))


# ---------------------------------------------------------------- level 2
def sgemv(op: CoralTranspose, alpha, beta, a: MatrixRef, x: VectorRef, y: VectorMut) -> None:
    """y := alpha op(A) x + beta y (coral-safe/src/level2/sgemv.rs:29-41, sgemv_n.rs:12-30, sgemv_t.rs)."""
    _kind(a, _Matrix, "a"), _kind(x, _Vector, "x"), _kind(y, VectorMut, "y"), _f32(a, x, y)
    if not isinstance(op, CoralTranspose):
        raise TypeError("op must be a CoralTranspose")
    rows, cols = a.n_rows(), a.n_cols()
    if op is CoralTranspose.NoTrans:
        _need(y.n() == rows, "logical length of y must equal n_rows")
        _need(x.n() == cols, "logical length of x must equal n_cols")
    else:
        _need(y.n() == cols, "logical length of y must equal n_cols")
        _need(x.n() == rows, "logical length of x must equal n_rows")
    blas.sgemv(op.value, rows, cols, float(alpha), a._window(), max(a.lda(), 1), x._window(), _inc(x),
               float(beta), y._window(), _inc(y))


def sger(alpha, a: MatrixMut, x: VectorRef, y: VectorRef) -> None:
    """A := alpha x y^T + A (coral-safe/src/level2/sger.rs:218-273)."""
    _kind(a, MatrixMut, "a"), _kind(x, _Vector, "x"), _kind(y, _Vector, "y"), _f32(a, x, y)
    _need(x.n() == a.n_rows(), "logical length of x must equal n_rows")
    _need(y.n() == a.n_cols(), "logical length of y must equal n_cols")
    blas.sger(a.n_rows(), a.n_cols(), float(alpha), x._window(), _inc(x), y._window(), _inc(y), a._window(),
              max(a.lda(), 1))


def strsv(uplo: CoralTriangular, trans: CoralTranspose, diag: CoralDiagonal, a: MatrixRef, x: VectorMut) -> None:
    """Solve op(A) x = b in place, A triangular n x n (coral-safe/src/level2/strsv.rs:27-38)."""
    _kind(a, _Matrix, "a"), _kind(x, VectorMut, "x"), _f32(a, x)
    if not (isinstance(uplo, CoralTriangular) and isinstance(trans, CoralTranspose)
            and isinstance(diag, CoralDiagonal)):
        raise TypeError("uplo, trans, diag must be CoralTriangular, CoralTranspose, CoralDiagonal")
    _need(a.compare_m_n(), "matrix must be square")
    _need(x.n() == a.n_rows(), "logical length of x must equal n")
    blas.strsv(uplo.value, trans.value, diag.value, a.n_rows(), a._window(), max(a.lda(), 1), x._window(),
               _inc(x))
