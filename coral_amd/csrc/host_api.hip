// Host-pointer ("drop-in") forms of every routine: CORAL's synchronous
// call-and-return semantics over host memory.  Each wrapper stages the operand
// extents CORAL's length contracts describe (coral-aarch64/src/level2/
// assert_length_helpers.rs:4-33: (cols-1)*ld + rows elements for a matrix,
// (n-1)*inc + 1 for a vector) into device buffers, runs the device routine on the
// library's staging stream and copies outputs back before returning.
#include "coral_common.hpp"

#include <mutex>

namespace coral {
namespace host {

static std::mutex g_host_mu;  // one staging area; host calls are serialised
static hipStream_t g_stream = nullptr;
static void *g_buf[6] = {};
static size_t g_cap[6] = {};

static int stage_stream() {
  if (!g_stream) CORAL_HIP(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
  return CORAL_OK;
}

static int buf(int slot, size_t bytes, void **p) {
  if (bytes == 0) bytes = 16;
  if (g_cap[slot] < bytes) {
    if (g_buf[slot]) CORAL_HIP(hipFree(g_buf[slot]));
    g_buf[slot] = nullptr;
    g_cap[slot] = 0;
    CORAL_HIP(hipMalloc(&g_buf[slot], bytes));
    g_cap[slot] = bytes;
  }
  *p = g_buf[slot];
  return CORAL_OK;
}

static size_t mat_bytes(int64_t rows, int64_t cols, int64_t ld, size_t es) {
  if (rows <= 0 || cols <= 0) return 0;
  return (size_t)((cols - 1) * ld + rows) * es;
}
static size_t vec_bytes(int64_t n, int64_t inc, size_t es) {
  if (n <= 0 || inc <= 0) return 0;
  return (size_t)((n - 1) * inc + 1) * es;
}

// copy host -> staging slot (nullptr when there is nothing to copy)
static int put(int slot, const void *h, size_t bytes, void **d) {
  *d = nullptr;
  if (!h || bytes == 0) return CORAL_OK;
  CORAL_TRY(buf(slot, bytes, d));
  CORAL_HIP(hipMemcpyAsync(*d, h, bytes, hipMemcpyHostToDevice, g_stream));
  return CORAL_OK;
}
static int get(void *h, const void *d, size_t bytes) {
  if (!h || !d || bytes == 0) return CORAL_OK;
  CORAL_HIP(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, g_stream));
  return CORAL_OK;
}
static int finish() {
  CORAL_HIP(hipStreamSynchronize(g_stream));
  return CORAL_OK;
}

#define HOST_PROLOGUE()                       \
  std::lock_guard<std::mutex> lk(g_host_mu);  \
  CORAL_TRY(ensure_init());                   \
  CORAL_TRY(stage_stream());                  \
  coral_stream_t st = (coral_stream_t)g_stream

template <typename F>
static int gemm_host(size_t es, int32_t op_a, int32_t op_b, int64_t m, int64_t n, int64_t k, const void *a,
                     int64_t lda, const void *b, int64_t ldb, void *c, int64_t ldc, F call) {
  HOST_PROLOGUE();
  const bool sane = m >= 0 && n >= 0 && k >= 0 && lda >= 1 && ldb >= 1 && ldc >= 1;
  if (!sane) return call(a, b, c, st);  // argument error, reported without touching memory
  const bool an = op_a == CORAL_NO_TRANSPOSE, bn = op_b == CORAL_NO_TRANSPOSE;
  size_t ab = mat_bytes(an ? m : k, an ? k : m, lda, es);
  size_t bb = mat_bytes(bn ? k : n, bn ? n : k, ldb, es);
  size_t cb = mat_bytes(m, n, ldc, es);
  void *da, *db, *dc;
  CORAL_TRY(put(0, a, ab, &da));
  CORAL_TRY(put(1, b, bb, &db));
  CORAL_TRY(put(2, c, cb, &dc));
  int s = call(da ? da : a, db ? db : b, dc ? dc : c, st);
  if (s != CORAL_OK) { finish(); return s; }
  CORAL_TRY(get(c, dc, cb));
  return finish();
}

// level-2 with a general/triangular matrix, input x and in/out y.
template <typename F>
static int mv_host(size_t es, int64_t rows, int64_t cols, int64_t lda, const void *a, int64_t xlen, const void *x,
                   int64_t incx, int64_t ylen, void *y, int64_t incy, bool sane, F call) {
  HOST_PROLOGUE();
  if (!sane || incx <= 0 || incy <= 0 || lda < 1 || rows < 0 || cols < 0) return call(a, x, y, st);
  void *da, *dx, *dy;
  size_t yb = vec_bytes(ylen, incy, es);
  CORAL_TRY(put(0, a, mat_bytes(rows, cols, lda, es), &da));
  CORAL_TRY(put(1, x, vec_bytes(xlen, incx, es), &dx));
  CORAL_TRY(put(2, y, yb, &dy));
  int s = call(da ? da : a, dx ? dx : x, dy ? dy : y, st);
  if (s != CORAL_OK) { finish(); return s; }
  CORAL_TRY(get(y, dy, yb));
  return finish();
}

// triangular solve / multiply: matrix in, x in/out
template <typename F>
static int tr_host(size_t es, int64_t n, int64_t lda, const void *a, void *x, int64_t incx, F call) {
  HOST_PROLOGUE();
  if (n < 0 || lda < 1 || incx <= 0) return call(a, x, st);
  void *da, *dx;
  size_t xb = vec_bytes(n, incx, es);
  CORAL_TRY(put(0, a, mat_bytes(n, n, lda, es), &da));
  CORAL_TRY(put(1, x, xb, &dx));
  int s = call(da ? da : a, dx ? dx : x, st);
  if (s != CORAL_OK) { finish(); return s; }
  CORAL_TRY(get(x, dx, xb));
  CORAL_TRY(finish());
  // a TRSV whose handoff wait gave up has written an invalid x: report it
  return coral_stream_check(st);
}

// level-1 with two vectors; y is copied back when `y_out`
template <typename F>
static int v2_host(size_t es, int64_t n, const void *x, int64_t incx, void *y, int64_t incy, bool y_out, F call) {
  HOST_PROLOGUE();
  if (n < 0 || incx <= 0 || incy <= 0) return call(x, y, st);
  void *dx, *dy;
  size_t yb = vec_bytes(n, incy, es);
  CORAL_TRY(put(0, x, vec_bytes(n, incx, es), &dx));
  CORAL_TRY(put(1, y, yb, &dy));
  int s = call(dx ? dx : x, dy ? dy : y, st);
  if (s != CORAL_OK) { finish(); return s; }
  if (y_out) CORAL_TRY(get(y, dy, yb));
  return finish();
}

// level-1 with two vectors both written (swap / rot / rotm)
template <typename F>
static int xy_host(size_t es, int64_t n, void *x, int64_t incx, void *y, int64_t incy, F call) {
  HOST_PROLOGUE();
  if (n < 0 || incx <= 0 || incy <= 0) return call(x, y, st);
  void *dx, *dy;
  const size_t xb = vec_bytes(n, incx, es), yb = vec_bytes(n, incy, es);
  CORAL_TRY(put(0, x, xb, &dx));
  CORAL_TRY(put(1, y, yb, &dy));
  int s = call(dx ? dx : x, dy ? dy : y, st);
  if (s != CORAL_OK) { finish(); return s; }
  CORAL_TRY(get(x, dx, xb));
  CORAL_TRY(get(y, dy, yb));
  return finish();
}

// level-1 with one vector in/out (scal)
template <typename F>
static int x_host(size_t es, int64_t n, void *x, int64_t incx, F call) {
  HOST_PROLOGUE();
  if (n < 0 || incx <= 0) return call(x, st);
  void *dx;
  const size_t xb = vec_bytes(n, incx, es);
  CORAL_TRY(put(0, x, xb, &dx));
  int s = call(dx ? dx : x, st);
  if (s != CORAL_OK) { finish(); return s; }
  CORAL_TRY(get(x, dx, xb));
  return finish();
}

// scalar in/out arguments staged in one device buffer (rotg / rotmg)
template <typename F>
static int scalars_host(void *const *h, const size_t *bytes, int count, F call) {
  HOST_PROLOGUE();
  for (int i = 0; i < count; ++i)
    if (!h[i]) return call((void *const *)h, st);  // null argument: reported without a device copy
  void *d;
  CORAL_TRY(buf(4, 64, &d));
  void *dp[5];
  size_t off = 0;
  for (int i = 0; i < count; ++i) {
    dp[i] = (char *)d + off;
    CORAL_HIP(hipMemcpyAsync(dp[i], h[i], bytes[i], hipMemcpyHostToDevice, g_stream));
    off += 8 * ((bytes[i] + 7) / 8);
  }
  int s = call((void *const *)dp, st);
  if (s != CORAL_OK) { finish(); return s; }
  for (int i = 0; i < count; ++i) CORAL_HIP(hipMemcpyAsync(h[i], dp[i], bytes[i], hipMemcpyDeviceToHost, g_stream));
  return finish();
}

// reductions: result of `rbytes` copied back into *result
template <typename F>
static int red_host(size_t es, int64_t n, const void *x, int64_t incx, const void *y, int64_t incy, void *result,
                    size_t rbytes, F call) {
  HOST_PROLOGUE();
  if (!result) return call(x, y, nullptr, st);
  void *dr;
  CORAL_TRY(buf(5, rbytes, &dr));
  if (n < 0) return call(x, y, dr, st);
  void *dx = nullptr, *dy = nullptr;
  if (incx > 0) CORAL_TRY(put(0, x, vec_bytes(n, incx, es), &dx));
  if (y && incy > 0) CORAL_TRY(put(1, y, vec_bytes(n, incy, es), &dy));
  int s = call(dx ? dx : x, dy ? dy : y, dr, st);
  if (s != CORAL_OK) { finish(); return s; }
  CORAL_TRY(get(result, dr, rbytes));
  return finish();
}

// rank updates: vectors in, matrix in/out
template <typename F>
static int rank_host(size_t es, int64_t rows, int64_t cols, const void *x, int64_t incx, const void *y,
                     int64_t incy, void *a, int64_t lda, bool sane, F call) {
  HOST_PROLOGUE();
  if (!sane || rows < 0 || cols < 0 || incx <= 0 || (y && incy <= 0) || lda < 1) return call(x, y, a, st);
  void *dx, *dy, *da;
  const size_t ab = mat_bytes(rows, cols, lda, es);
  CORAL_TRY(put(0, x, vec_bytes(rows, incx, es), &dx));
  CORAL_TRY(put(1, y, vec_bytes(cols, incy, es), &dy));
  CORAL_TRY(put(2, a, ab, &da));
  int s = call(dx ? dx : x, dy ? dy : y, da ? da : a, st);
  if (s != CORAL_OK) { finish(); return s; }
  CORAL_TRY(get(a, da, ab));
  return finish();
}

}  // namespace host
}  // namespace coral

using namespace coral;
using namespace coral::host;

extern "C" {

// ---- level 3 ----
#define GEMM_HOST(NAME, T, ES, ALPHA_T, BETA_T)                                                                  \
  int coralh_##NAME(int32_t op_a, int32_t op_b, int64_t m, int64_t n, int64_t k, ALPHA_T alpha, const T *a,     \
                    int64_t lda, const T *b, int64_t ldb, BETA_T beta, T *c, int64_t ldc) {                     \
    return gemm_host(ES, op_a, op_b, m, n, k, a, lda, b, ldb, c, ldc,                                            \
                     [&](const void *da, const void *db, void *dc, coral_stream_t st) {                         \
                       return coral_##NAME(op_a, op_b, m, n, k, alpha, (const T *)da, lda, (const T *)db, ldb,  \
                                           beta, (T *)dc, ldc, st);                                              \
                     });                                                                                         \
  }
GEMM_HOST(sgemm, float, 4, float, float)
GEMM_HOST(dgemm, double, 8, double, double)
GEMM_HOST(cgemm, float, 8, const float *, const float *)
GEMM_HOST(zgemm, double, 16, const double *, const double *)
GEMM_HOST(sgemm_bf16x3, float, 4, float, float)
GEMM_HOST(cgemm_3m_bf16x3, float, 8, const float *, const float *)
GEMM_HOST(sgemm_fp16x3, float, 4, float, float)
GEMM_HOST(cgemm_3m_fp16x3, float, 8, const float *, const float *)

// ---- level 2 ----
#define GEMV_HOST(NAME, T, ES, S_T)                                                                              \
  int coralh_##NAME(int32_t trans, int64_t m, int64_t n, S_T alpha, const T *a, int64_t lda, const T *x,         \
                    int64_t incx, S_T beta, T *y, int64_t incy) {                                                \
    const bool nt = trans == CORAL_NO_TRANSPOSE;                                                                 \
    return mv_host(ES, m, n, lda, a, nt ? n : m, x, incx, nt ? m : n, y, incy, valid_op(trans),                 \
                   [&](const void *da, const void *dx, void *dy, coral_stream_t st) {                           \
                     return coral_##NAME(trans, m, n, alpha, (const T *)da, lda, (const T *)dx, incx, beta,     \
                                         (T *)dy, incy, st);                                                     \
                   });                                                                                           \
  }
GEMV_HOST(sgemv, float, 4, float)
GEMV_HOST(dgemv, double, 8, double)
GEMV_HOST(cgemv, float, 8, const float *)
GEMV_HOST(zgemv, double, 16, const double *)

#define SYMV_HOST(NAME, T, ES, S_T)                                                                              \
  int coralh_##NAME(int32_t uplo, int64_t n, S_T alpha, const T *a, int64_t lda, const T *x, int64_t incx,      \
                    S_T beta, T *y, int64_t incy) {                                                              \
    return mv_host(ES, n, n, lda, a, n, x, incx, n, y, incy, valid_uplo(uplo),                                  \
                   [&](const void *da, const void *dx, void *dy, coral_stream_t st) {                           \
                     return coral_##NAME(uplo, n, alpha, (const T *)da, lda, (const T *)dx, incx, beta,         \
                                         (T *)dy, incy, st);                                                     \
                   });                                                                                           \
  }
SYMV_HOST(ssymv, float, 4, float)
SYMV_HOST(dsymv, double, 8, double)
SYMV_HOST(chemv, float, 8, const float *)
SYMV_HOST(zhemv, double, 16, const double *)

#define TR_HOST(NAME, T, ES)                                                                                     \
  int coralh_##NAME(int32_t uplo, int32_t trans, int32_t diag, int64_t n, const T *a, int64_t lda, T *x,        \
                    int64_t incx) {                                                                              \
    return tr_host(ES, n, lda, a, x, incx, [&](const void *da, void *dx, coral_stream_t st) {                   \
      return coral_##NAME(uplo, trans, diag, n, (const T *)da, lda, (T *)dx, incx, st);                         \
    });                                                                                                          \
  }
TR_HOST(strsv, float, 4)
TR_HOST(dtrsv, double, 8)
TR_HOST(ctrsv, float, 8)
TR_HOST(ztrsv, double, 16)
TR_HOST(strmv, float, 4)
TR_HOST(dtrmv, double, 8)
TR_HOST(ctrmv, float, 8)
TR_HOST(ztrmv, double, 16)

// ---- rank updates ----
#define GER_HOST(NAME, T, ES, S_T)                                                                               \
  int coralh_##NAME(int64_t m, int64_t n, S_T alpha, const T *x, int64_t incx, const T *y, int64_t incy, T *a,  \
                    int64_t lda) {                                                                               \
    return rank_host(ES, m, n, x, incx, y, incy, a, lda, true,                                                   \
                     [&](const void *dx, const void *dy, void *da, coral_stream_t st) {                         \
                       return coral_##NAME(m, n, alpha, (const T *)dx, incx, (const T *)dy, incy, (T *)da, lda, \
                                           st);                                                                  \
                     });                                                                                         \
  }
GER_HOST(sger, float, 4, float)
GER_HOST(dger, double, 8, double)
GER_HOST(cgeru, float, 8, const float *)
GER_HOST(cgerc, float, 8, const float *)
GER_HOST(zgeru, double, 16, const double *)
GER_HOST(zgerc, double, 16, const double *)

#define SYR_HOST(NAME, T, ES, S_T)                                                                               \
  int coralh_##NAME(int32_t uplo, int64_t n, S_T alpha, const T *x, int64_t incx, T *a, int64_t lda) {          \
    return rank_host(ES, n, n, x, incx, nullptr, 1, a, lda, valid_uplo(uplo),                                   \
                     [&](const void *dx, const void *, void *da, coral_stream_t st) {                           \
                       return coral_##NAME(uplo, n, alpha, (const T *)dx, incx, (T *)da, lda, st);              \
                     });                                                                                         \
  }
SYR_HOST(ssyr, float, 4, float)
SYR_HOST(dsyr, double, 8, double)
SYR_HOST(cher, float, 8, float)
SYR_HOST(zher, double, 16, double)

#define SYR2_HOST(NAME, T, ES, S_T)                                                                              \
  int coralh_##NAME(int32_t uplo, int64_t n, S_T alpha, const T *x, int64_t incx, const T *y, int64_t incy,     \
                    T *a, int64_t lda) {                                                                         \
    return rank_host(ES, n, n, x, incx, y, incy, a, lda, valid_uplo(uplo),                                       \
                     [&](const void *dx, const void *dy, void *da, coral_stream_t st) {                         \
                       return coral_##NAME(uplo, n, alpha, (const T *)dx, incx, (const T *)dy, incy, (T *)da,   \
                                           lda, st);                                                             \
                     });                                                                                         \
  }
SYR2_HOST(ssyr2, float, 4, float)
SYR2_HOST(dsyr2, double, 8, double)
SYR2_HOST(cher2, float, 8, const float *)
SYR2_HOST(zher2, double, 16, const double *)

// ---- level 1 ----
#define DOT_HOST(NAME, T, ES, RB)                                                                                \
  int coralh_##NAME(int64_t n, const T *x, int64_t incx, const T *y, int64_t incy, T *result) {                 \
    return red_host(ES, n, x, incx, y, incy, result, RB,                                                         \
                    [&](const void *dx, const void *dy, void *dr, coral_stream_t st) {                          \
                      return coral_##NAME(n, (const T *)dx, incx, (const T *)dy, incy, (T *)dr, st);            \
                    });                                                                                          \
  }
DOT_HOST(sdot, float, 4, 4)
DOT_HOST(ddot, double, 8, 8)
DOT_HOST(cdotu, float, 8, 8)
DOT_HOST(cdotc, float, 8, 8)
DOT_HOST(zdotu, double, 16, 16)
DOT_HOST(zdotc, double, 16, 16)

#define AXPY_HOST(NAME, T, ES, S_T)                                                                              \
  int coralh_##NAME(int64_t n, S_T alpha, const T *x, int64_t incx, T *y, int64_t incy) {                       \
    return v2_host(ES, n, x, incx, y, incy, true, [&](const void *dx, void *dy, coral_stream_t st) {            \
      return coral_##NAME(n, alpha, (const T *)dx, incx, (T *)dy, incy, st);                                    \
    });                                                                                                          \
  }
AXPY_HOST(saxpy, float, 4, float)
AXPY_HOST(daxpy, double, 8, double)
AXPY_HOST(caxpy, float, 8, const float *)
AXPY_HOST(zaxpy, double, 16, const double *)

#define NRM2_HOST(NAME, T, ES, R)                                                                                \
  int coralh_##NAME(int64_t n, const T *x, int64_t incx, R *result) {                                           \
    return red_host(ES, n, x, incx, nullptr, 1, result, sizeof(R),                                               \
                    [&](const void *dx, const void *, void *dr, coral_stream_t st) {                            \
                      return coral_##NAME(n, (const T *)dx, incx, (R *)dr, st);                                  \
                    });                                                                                          \
  }
NRM2_HOST(snrm2, float, 4, float)
NRM2_HOST(dnrm2, double, 8, double)
NRM2_HOST(scnrm2, float, 8, float)
NRM2_HOST(dznrm2, double, 16, double)

#define IAMAX_HOST(NAME, T, ES)                                                                                  \
  int coralh_##NAME(int64_t n, const T *x, int64_t incx, int64_t *result) {                                     \
    return red_host(ES, n, x, incx, nullptr, 1, result, sizeof(int64_t),                                        \
                    [&](const void *dx, const void *, void *dr, coral_stream_t st) {                            \
                      return coral_##NAME(n, (const T *)dx, incx, (int64_t *)dr, st);                           \
                    });                                                                                          \
  }
IAMAX_HOST(isamax, float, 4)
IAMAX_HOST(idamax, double, 8)
IAMAX_HOST(icamax, float, 8)
IAMAX_HOST(izamax, double, 16)

#define SCAL_HOST(NAME, T, ES, S_T)                                                                              \
  int coralh_##NAME(int64_t n, S_T alpha, T *x, int64_t incx) {                                                  \
    return x_host(ES, n, x, incx, [&](void *dx, coral_stream_t st) { return coral_##NAME(n, alpha, (T *)dx, incx, st); }); \
  }
SCAL_HOST(sscal, float, 4, float)
SCAL_HOST(dscal, double, 8, double)
SCAL_HOST(cscal, float, 8, const float *)
SCAL_HOST(zscal, double, 16, const double *)
SCAL_HOST(csscal, float, 8, float)
SCAL_HOST(zdscal, double, 16, double)

NRM2_HOST(sasum, float, 4, float)
NRM2_HOST(dasum, double, 8, double)
NRM2_HOST(scasum, float, 8, float)
NRM2_HOST(dzasum, double, 16, double)

#define COPY_HOST(NAME, T, ES)                                                                                   \
  int coralh_##NAME(int64_t n, const T *x, int64_t incx, T *y, int64_t incy) {                                 \
    return v2_host(ES, n, x, incx, y, incy, true, [&](const void *dx, void *dy, coral_stream_t st) {             \
      return coral_##NAME(n, (const T *)dx, incx, (T *)dy, incy, st);                                            \
    });                                                                                                          \
  }
COPY_HOST(scopy, float, 4)
COPY_HOST(dcopy, double, 8)
COPY_HOST(ccopy, float, 8)
COPY_HOST(zcopy, double, 16)

#define SWAP_HOST(NAME, T, ES)                                                                                   \
  int coralh_##NAME(int64_t n, T *x, int64_t incx, T *y, int64_t incy) {                                       \
    return xy_host(ES, n, x, incx, y, incy, [&](void *dx, void *dy, coral_stream_t st) {                         \
      return coral_##NAME(n, (T *)dx, incx, (T *)dy, incy, st);                                                  \
    });                                                                                                          \
  }
SWAP_HOST(sswap, float, 4)
SWAP_HOST(dswap, double, 8)
SWAP_HOST(cswap, float, 8)
SWAP_HOST(zswap, double, 16)

#define ROT_HOST(NAME, T, ES)                                                                                    \
  int coralh_##NAME(int64_t n, T *x, int64_t incx, T *y, int64_t incy, T c, T s) {                             \
    return xy_host(ES, n, x, incx, y, incy, [&](void *dx, void *dy, coral_stream_t st) {                         \
      return coral_##NAME(n, (T *)dx, incx, (T *)dy, incy, c, s, st);                                            \
    });                                                                                                          \
  }
ROT_HOST(srot, float, 4)
ROT_HOST(drot, double, 8)
ROT_HOST(csrot, float, 8)
ROT_HOST(zdrot, double, 16)

#define ROTM_HOST(NAME, T, ES)                                                                                   \
  int coralh_##NAME(int64_t n, T *x, int64_t incx, T *y, int64_t incy, const T param[5]) {                     \
    return xy_host(ES, n, x, incx, y, incy, [&](void *dx, void *dy, coral_stream_t st) {                         \
      return coral_##NAME(n, (T *)dx, incx, (T *)dy, incy, param, st);                                           \
    });                                                                                                          \
  }
ROTM_HOST(srotm, float, 4)
ROTM_HOST(drotm, double, 8)

#define ROTG_HOST(NAME, T)                                                                                       \
  int coralh_##NAME(T *a, T *b, T *c, T *s) {                                                                    \
    void *h[4] = {a, b, c, s};                                                                                   \
    const size_t bytes[4] = {sizeof(T), sizeof(T), sizeof(T), sizeof(T)};                                        \
    return scalars_host(h, bytes, 4, [&](void *const *d, coral_stream_t st) {                                    \
      return coral_##NAME((T *)d[0], (T *)d[1], (T *)d[2], (T *)d[3], st);                                       \
    });                                                                                                          \
  }
ROTG_HOST(srotg, float)
ROTG_HOST(drotg, double)

#define ROTMG_HOST(NAME, T)                                                                                      \
  int coralh_##NAME(T *d1, T *d2, T *x1, T y1, T param[5]) {                                                     \
    void *h[4] = {d1, d2, x1, param};                                                                            \
    const size_t bytes[4] = {sizeof(T), sizeof(T), sizeof(T), 5 * sizeof(T)};                                    \
    return scalars_host(h, bytes, 4, [&](void *const *d, coral_stream_t st) {                                    \
      return coral_##NAME((T *)d[0], (T *)d[1], (T *)d[2], y1, (T *)d[3], st);                                   \
    });                                                                                                          \
  }
ROTMG_HOST(srotmg, float)
ROTMG_HOST(drotmg, double)

}  // extern "C"
