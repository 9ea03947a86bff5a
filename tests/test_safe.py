"""Checked front end (coral_amd.safe), host side only: no GPU call is made here.

The first three tests are coral-safe's own view tests (coral-safe/tests/types.rs:6-132)
with the same buffers and expected errors.  The rest pin the view arithmetic
(required lengths, saturation, empty views, contiguous windows) and show that
every routine checks its views' lengths and hands the library the window that
starts at the view's offset, with the view's stride, before anything reaches
the GPU (a recorder stands in for the host-pointer bindings).
"""
import numpy as np
import pytest

from coral_amd import safe
from coral_amd.safe import BufferError, MatrixMut, MatrixRef, VectorMut, VectorRef


def test_ensure_nonzero_stride_vec():  # coral-safe/tests/types.rs:6-31
    n = 5
    data_ref = np.ones(n, np.float32)
    data_mut = np.ones(n, np.float32)
    for cls, data in ((VectorRef, data_ref), (VectorMut, data_mut)):
        with pytest.raises(BufferError) as e:
            cls.new(data, n, 0, 0)
        assert e.value == BufferError.ZeroStride()


def test_ensure_within_bounds_vec():  # coral-safe/tests/types.rs:33-79
    n, stride, offset = 3, 2, 1
    for cls in (VectorRef, VectorMut):
        with pytest.raises(BufferError) as e:
            cls.new(np.ones(5, np.float32), n, stride, offset)
        assert e.value == BufferError.OutOfBounds(required=6, len=5)
        assert cls.new(np.ones(6, np.float32), n, stride, offset).n() == 3


def test_ensure_valid_lda_mat():  # coral-safe/tests/types.rs:82-132
    n_rows, n_cols, lda_valid, lda_invalid = 5, 6, 7, 4
    for cls in (MatrixRef, MatrixMut):
        with pytest.raises(BufferError) as e:
            cls.new(np.ones(n_cols * lda_invalid, np.float32), n_rows, n_cols, lda_invalid, 0)
        assert e.value == BufferError.InvalidLda(lda=lda_invalid, n_rows=n_rows)
        assert cls.new(np.ones(n_cols * lda_valid, np.float32), n_rows, n_cols, lda_valid, 0).lda() == 7


def test_error_values_and_display():  # coral-safe/src/errors.rs:3-25
    assert str(BufferError.ZeroStride()) == "stride must be nonzero"
    assert str(BufferError.OutOfBounds(required=6, len=5)) == "buffer too small: required 6, have 5"
    assert str(BufferError.BadOffset(offset=9, len=4)) == "offset 9 out of range for buffer of len 4"
    assert str(BufferError.InvalidLda(lda=4, n_rows=5)) == "leading dimension 4 must exceed num rows 5"
    assert BufferError.OutOfBounds(required=6, len=5) != BufferError.OutOfBounds(required=7, len=5)
    assert BufferError.OutOfBounds(required=6, len=5) != BufferError.BadOffset(offset=6, len=5)
    assert BufferError.InvalidLda(lda=4, n_rows=5).lda == 4
    assert isinstance(BufferError.ZeroStride(), ValueError)


def test_vector_required_length():
    # required = (n-1)*stride + offset + 1, checked against the whole buffer (types.rs:89-101)
    buf = np.zeros(20, np.float32)
    for n, stride, offset in ((1, 1, 19), (4, 5, 4), (20, 1, 0), (7, 3, 1)):
        req = (n - 1) * stride + offset + 1
        v = VectorRef(buf[:req], n, stride, offset)
        assert (v.n(), v.stride(), v.offset()) == (n, stride, offset)
        with pytest.raises(BufferError) as e:
            VectorRef(buf[:req - 1], n, stride, offset)
        assert e.value == BufferError.OutOfBounds(required=req, len=req - 1)
    # saturating arithmetic: a huge stride reports usize::MAX, as the reference does
    with pytest.raises(BufferError) as e:
        VectorRef(buf, 3, 1 << 63, 0)
    assert e.value == BufferError.OutOfBounds(required=(1 << 64) - 1, len=20)


def test_empty_views_are_valid():
    # n == 0 skips every check, including the zero stride (types.rs:78-86); a matrix with
    # no rows or no columns skips them too (types.rs:190-199)
    e = np.zeros(0, np.float32)
    assert VectorRef(e, 0, 0, 7).n() == 0
    assert VectorMut(np.zeros(0, np.float32), 0, 0, 0).stride() == 0
    assert MatrixRef(e, 0, 5, 0, 3).n_cols() == 5
    assert MatrixMut(np.zeros(0, np.float32), 4, 0, 0, 0).n_rows() == 4


def test_matrix_required_length_and_lda():
    buf = np.zeros(100, np.float32)
    # required = (n_cols-1)*lda + n_rows + offset (types.rs:207-219)
    req = (6 - 1) * 9 + 5 + 3
    assert MatrixRef(buf[:req], 5, 6, 9, 3).offset() == 3
    with pytest.raises(BufferError) as e:
        MatrixRef(buf[:req - 1], 5, 6, 9, 3)
    assert e.value == BufferError.OutOfBounds(required=req, len=req - 1)
    with pytest.raises(BufferError) as e:
        MatrixRef(buf, 5, 6, 0, 0)
    assert e.value == BufferError.ZeroStride()
    assert MatrixRef(buf, 5, 6, 5, 0).compare_m_n() is False
    assert MatrixRef(buf, 6, 6, 6, 0).compare_m_n() is True


def test_contiguous_windows():
    buf = np.arange(40, dtype=np.float32)
    assert np.array_equal(VectorRef(buf, 5, 1, 3).contiguous_slice(), buf[3:8])
    assert VectorRef(buf, 5, 2, 3).contiguous_slice() is None
    m = MatrixRef(buf, 4, 3, 4, 2)
    assert np.array_equal(m.contiguous_slice(), buf[2:14])
    assert MatrixRef(buf, 4, 3, 5, 2).contiguous_slice() is None
    w = VectorMut(buf, 5, 1, 0)
    w.contiguous_slice_mut()[0] = -1
    assert buf[0] == -1  # the window is a view of the caller's buffer
    assert VectorRef(buf, 5, 1, 0).compare_n(5) and not VectorRef(buf, 5, 1, 0).compare_n(4)


def test_view_argument_checks():
    ro = np.ones(8, np.float32)
    ro.flags.writeable = False
    VectorRef(ro, 8, 1, 0)
    with pytest.raises(ValueError):
        VectorMut(ro, 8, 1, 0)
    with pytest.raises(ValueError):
        VectorRef(np.ones((2, 4), np.float32), 8, 1, 0)
    with pytest.raises(ValueError):
        VectorRef(np.ones(8, np.float32), -1, 1, 0)
    with pytest.raises(TypeError):
        VectorRef(np.ones(8, np.float32), 2.0, 1, 0)


class _Recorder:
    """Stands in for coral_amd.blas: records each call instead of reaching the library."""

    def __init__(self):
        self.calls = []

    def __getattr__(self, name):
        def f(*args):
            self.calls.append((name, args))
            return {"sasum": 0.0, "snrm2": 0.0, "sdot": 0.0, "isamax": 0,
                    "srotg": (1.0, 2.0, 3.0, 4.0),
                    "srotmg": (1.0, 2.0, 3.0, np.arange(5, dtype=np.float32))}.get(name)

        return f


@pytest.fixture
def rec(monkeypatch):
    r = _Recorder()
    monkeypatch.setattr(safe, "blas", r)
    return r


def _same_window(w, buf, offset):
    """w is the caller's buffer from `offset` on (a view, not a copy)."""
    return np.shares_memory(w, buf) and w.__array_interface__["data"][0] == buf[offset:].__array_interface__["data"][0]


def test_routines_pass_offset_windows(rec):
    x = np.zeros(40, np.float32)
    y = np.zeros(60, np.float32)
    a = np.zeros(200, np.float32)
    xv = VectorRef(x, 10, 3, 2)
    yv = VectorMut(y, 10, 5, 4)
    safe.saxpy(2.5, xv, yv)
    name, args = rec.calls[-1]
    assert name == "saxpy" and args[0] == 10 and args[1] == 2.5 and args[3] == 3 and args[5] == 5
    assert _same_window(args[2], x, 2) and _same_window(args[4], y, 4)

    am = MatrixMut(a, 10, 12, 13, 7)
    safe.sger(1.5, am, VectorRef(x, 10, 3, 2), VectorRef(y, 12, 4, 1))
    name, args = rec.calls[-1]
    assert name == "sger" and args[:3] == (10, 12, 1.5) and args[4] == 3 and args[6] == 4 and args[8] == 13
    assert _same_window(args[3], x, 2) and _same_window(args[5], y, 1) and _same_window(args[7], a, 7)

    ar = MatrixRef(a, 10, 12, 13, 7)
    safe.sgemv(safe.CoralTranspose.Trans, 2.0, 0.5, ar, VectorRef(x, 10, 3, 2), VectorMut(y, 12, 4, 1))
    name, args = rec.calls[-1]
    assert name == "sgemv" and args[:3] == (1, 10, 12) and args[5] == 13 and args[7] == 3 and args[8] == 0.5
    safe.saxpyf(ar, VectorRef(y, 12, 4, 1), VectorMut(x, 10, 3, 2))
    assert rec.calls[-1][1][:4] == (0, 10, 12, 1.0) and rec.calls[-1][1][8] == 1.0
    safe.sdotf(ar, VectorRef(x, 10, 3, 2), VectorMut(y, 12, 4, 1))
    assert rec.calls[-1][1][:4] == (1, 10, 12, 1.0) and rec.calls[-1][1][8] == 1.0

    sq = MatrixRef(a, 12, 12, 13, 5)
    safe.strsv(safe.CoralTriangular.Lower, safe.CoralTranspose.NoTrans, safe.CoralDiagonal.Unit, sq,
               VectorMut(y, 12, 4, 1))
    name, args = rec.calls[-1]
    assert name == "strsv" and args[:4] == (1, 0, 0, 12) and args[5] == 13 and args[7] == 4
    assert _same_window(args[4], a, 5) and _same_window(args[6], y, 1)


def test_scalar_routines_write_back(rec):
    a, b, c, s = (np.array([v], np.float32) for v in (0.5, -1.25, 0.0, 0.0))
    safe.srotg(a, b, c, s)
    assert (a[0], b[0], c[0], s[0]) == (1.0, 2.0, 3.0, 4.0)
    d1, d2, x1 = (np.array([v], np.float32) for v in (1.5, 2.0, -0.75))
    param = np.zeros(5, np.float32)
    safe.srotmg(d1, d2, x1, 0.5, param)
    assert (d1[0], d2[0], x1[0]) == (1.0, 2.0, 3.0) and np.array_equal(param, np.arange(5))
    with pytest.raises(TypeError):
        safe.srotg(0.5, b, c, s)  # a plain float cannot be written back
    with pytest.raises(TypeError):
        safe.srotmg(d1, d2, x1, 0.5, [0.0] * 5)


def test_length_mismatch_is_an_error_before_the_gpu(rec):
    x = np.zeros(64, np.float32)
    y = np.zeros(64, np.float32)
    a = np.zeros(64 * 64, np.float32)
    cases = [
        lambda: safe.saxpy(1.0, VectorRef(x, 4, 1, 0), VectorMut(y, 5, 1, 0)),
        lambda: safe.scopy(VectorRef(x, 4, 1, 0), VectorMut(y, 3, 1, 0)),
        lambda: safe.sswap(VectorMut(x, 4, 1, 0), VectorMut(y, 3, 1, 0)),
        lambda: safe.sdot(VectorRef(x, 4, 1, 0), VectorRef(y, 3, 1, 0)),
        lambda: safe.srot(VectorMut(x, 4, 1, 0), VectorMut(y, 3, 1, 0), 1.0, 0.0),
        lambda: safe.srotm(VectorMut(x, 4, 1, 0), VectorMut(y, 3, 1, 0), np.zeros(5, np.float32)),
        lambda: safe.srotm(VectorMut(x, 4, 1, 0), VectorMut(y, 4, 1, 0), np.zeros(4, np.float32)),
        lambda: safe.sgemv(safe.CoralTranspose.NoTrans, 1.0, 0.0, MatrixRef(a, 8, 6, 8, 0),
                           VectorRef(x, 8, 1, 0), VectorMut(y, 8, 1, 0)),
        lambda: safe.sgemv(safe.CoralTranspose.Trans, 1.0, 0.0, MatrixRef(a, 8, 6, 8, 0),
                           VectorRef(x, 6, 1, 0), VectorMut(y, 6, 1, 0)),
        lambda: safe.sger(1.0, MatrixMut(a, 8, 6, 8, 0), VectorRef(x, 6, 1, 0), VectorRef(y, 8, 1, 0)),
        lambda: safe.strsv(safe.CoralTriangular.Upper, safe.CoralTranspose.NoTrans, safe.CoralDiagonal.NonUnit,
                           MatrixRef(a, 8, 6, 8, 0), VectorMut(x, 8, 1, 0)),
        lambda: safe.strsv(safe.CoralTriangular.Upper, safe.CoralTranspose.NoTrans, safe.CoralDiagonal.NonUnit,
                           MatrixRef(a, 8, 8, 8, 0), VectorMut(x, 7, 1, 0)),
        lambda: safe.saxpyf(MatrixRef(a, 8, 6, 8, 0), VectorRef(x, 8, 1, 0), VectorMut(y, 8, 1, 0)),
        lambda: safe.sdotf(MatrixRef(a, 8, 6, 8, 0), VectorRef(x, 6, 1, 0), VectorMut(y, 6, 1, 0)),
    ]
    for i, case in enumerate(cases):
        with pytest.raises(ValueError):
            case()
        assert rec.calls == [], f"case {i} reached the library"


def test_type_checks(rec):
    x64 = np.zeros(8, np.float64)
    with pytest.raises(TypeError):
        safe.sasum(VectorRef(x64, 8, 1, 0))  # single precision only
    with pytest.raises(TypeError):
        safe.saxpy(1.0, VectorRef(np.zeros(8, np.float32), 8, 1, 0), VectorRef(np.zeros(8, np.float32), 8, 1, 0))
    with pytest.raises(TypeError):
        safe.sscal(2.0, np.zeros(8, np.float32))
    with pytest.raises(TypeError):
        safe.sgemv(0, 1.0, 0.0, MatrixRef(np.zeros(4, np.float32), 2, 2, 2, 0), VectorRef(np.zeros(2, np.float32), 2, 1, 0),
                   VectorMut(np.zeros(2, np.float32), 2, 1, 0))
    assert rec.calls == []


def test_empty_views_pass_a_positive_stride(rec):
    e = np.zeros(0, np.float32)
    safe.saxpy(1.0, VectorRef(e, 0, 0, 0), VectorMut(np.zeros(0, np.float32), 0, 0, 0))
    name, args = rec.calls[-1]
    assert args[0] == 0 and args[3] == 1 and args[5] == 1
