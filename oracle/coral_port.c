/*
 * coral_port.c -- CPU restatement of CORAL's BLAS algorithms.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle: tests/, the smoke check and
 * bench.py's cpu_baseline leg load it as the checker or as the timed CPU baseline.
 * The product (libcoralhip.so) never links, calls or falls back to it.
 *
 * It restates the algorithms of the reference crate coral-aarch64 (paths below are
 * relative to the reference root), which cannot be built here: it is Rust (no
 * cargo/rustc in this image) and every module is cfg-gated to aarch64
 * (coral-aarch64/src/lib.rs:70-82).  NEON vfmaq/vfmsq sites become fma(); Rust's
 * non-fused `a*b + c` sites stay non-fused (build with -ffp-contract=off).
 *
 * Column-major, positive strides, complex interleaved [re, im] with ld/inc in
 * complex elements, 0-based i?amax -- as CORAL.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define NT 0
#define TR 1
#define CT 2

static int64_t rup(int64_t x, int64_t b) { return (x + b - 1) / b * b; }
static int64_t imin(int64_t a, int64_t b) { return a < b ? a : b; }

/* ======================================================================== */
/* Real GEMM: Goto/BLIS 5-loop                                              */
/*   blocking  sgemm MC/NC/KC = 384/576/256, MR x NR = 8 x 12                */
/*             (coral-aarch64/src/level3/sgemm.rs:92-94, f32_packers.rs:1-2) */
/*             dgemm MC/NC/KC = 258/384/256, MR x NR = 6 x 8                  */
/*             (coral-aarch64/src/level3/dgemm.rs:93-95, f64_packers.rs:1-2)  */
/*   driver    sgemm_nn.rs:10-99 (quick return :29-45, beta on first KC panel */
/*             only :64); the T variants differ only in the packers          */
/*   packers   f32_packers.rs:115-333 (zero-padded MR/NR tails :66-86)       */
/*   macro     f32_macro_kernel.rs:12-57 (full tile -> mrxnr, else edge)     */
/*   micro     microkernel/f32_mrxnr.rs:99-325: acc = fma(a, b, acc) over k;  */
/*             beta0 C = alpha*acc, beta1 C = fma(acc, alpha, C),             */
/*             betax C = fma(acc, alpha, C*beta)                             */
/*   edge      microkernel/f32_edge.rs:4-61: acc += a*b (not fused);          */
/*             C = alpha*acc | C += alpha*acc | C = beta*C + alpha*acc       */
/* ======================================================================== */
#define DEFINE_REAL_GEMM(NAME, T, FMA, MC_, NC_, KC_, MR_, NR_)                                              \
  static void NAME##_pack_a(int op, int64_t mc, int64_t kc, const T *a, int64_t lda, T *dst) {             \
    for (int64_t p = 0; p < mc; p += MR_)                                                                    \
      for (int64_t k = 0; k < kc; ++k)                                                                       \
        for (int t = 0; t < MR_; ++t) {                                                                      \
          int64_t i = p + t;                                                                                 \
          *dst++ = (i < mc) ? (op == NT ? a[i + k * lda] : a[k + i * lda]) : (T)0;                           \
        }                                                                                                    \
  }                                                                                                          \
  static void NAME##_pack_b(int op, int64_t kc, int64_t nc, const T *b, int64_t ldb, T *dst) {             \
    for (int64_t p = 0; p < nc; p += NR_)                                                                    \
      for (int64_t k = 0; k < kc; ++k)                                                                       \
        for (int t = 0; t < NR_; ++t) {                                                                      \
          int64_t j = p + t;                                                                                 \
          *dst++ = (j < nc) ? (op == NT ? b[k + j * ldb] : b[j + k * ldb]) : (T)0;                           \
        }                                                                                                    \
  }                                                                                                          \
  static void NAME##_micro(int64_t kc, const T *ap, const T *bp, T *c, int64_t ldc, T alpha, T beta) {      \
    T acc[MR_][NR_];                                                                                         \
    memset(acc, 0, sizeof(acc));                                                                             \
    for (int64_t k = 0; k < kc; ++k, ap += MR_, bp += NR_)                                                   \
      for (int j = 0; j < NR_; ++j)                                                                          \
        for (int r = 0; r < MR_; ++r) acc[r][j] = FMA(ap[r], bp[j], acc[r][j]);                              \
    for (int j = 0; j < NR_; ++j)                                                                            \
      for (int r = 0; r < MR_; ++r) {                                                                        \
        T *cp = c + r + j * ldc;                                                                             \
        if (beta == (T)0) *cp = acc[r][j] * alpha;                                                           \
        else if (beta == (T)1) *cp = FMA(acc[r][j], alpha, *cp);                                             \
        else *cp = FMA(acc[r][j], alpha, *cp * beta);                                                        \
      }                                                                                                      \
  }                                                                                                          \
  static void NAME##_edge(int64_t mr, int64_t nr, int64_t kc, const T *ap, const T *bp, T *c, int64_t ldc, \
                          T alpha, T beta) {                                                                 \
    T acc[MR_][NR_];                                                                                         \
    memset(acc, 0, sizeof(acc));                                                                             \
    for (int64_t k = 0; k < kc; ++k, ap += MR_, bp += NR_)                                                   \
      for (int64_t r = 0; r < mr; ++r)                                                                       \
        for (int64_t j = 0; j < nr; ++j) acc[r][j] += ap[r] * bp[j];                                         \
    for (int64_t j = 0; j < nr; ++j)                                                                         \
      for (int64_t r = 0; r < mr; ++r) {                                                                     \
        T *cp = c + r + j * ldc;                                                                             \
        if (beta == (T)0) *cp = alpha * acc[r][j];                                                           \
        else if (beta == (T)1) *cp += alpha * acc[r][j];                                                     \
        else *cp = beta * *cp + alpha * acc[r][j];                                                           \
      }                                                                                                      \
  }                                                                                                          \
  void port_##NAME(int op_a, int op_b, int64_t m, int64_t n, int64_t k, T alpha, const T *a, int64_t lda,  \
                   const T *b, int64_t ldb, T beta, T *c, int64_t ldc) {                                     \
    if (op_a == CT) op_a = TR; /* real: ConjTranspose == Transpose (sgemm.rs:97-103) */                      \
    if (op_b == CT) op_b = TR;                                                                               \
    if (alpha == (T)0 || k == 0) {                                                                           \
      for (int64_t j = 0; j < n; ++j)                                                                        \
        for (int64_t i = 0; i < m; ++i) {                                                                    \
          if (beta == (T)0) c[i + j * ldc] = 0;                                                              \
          else if (beta != (T)1) c[i + j * ldc] *= beta;                                                     \
        }                                                                                                    \
      return;                                                                                                \
    }                                                                                                        \
    T *abuf = (T *)malloc(sizeof(T) * rup(MC_, MR_) * KC_);                                                  \
    T *bbuf = (T *)malloc(sizeof(T) * KC_ * rup(NC_, NR_));                                                  \
    for (int64_t j0 = 0; j0 < n; j0 += NC_) {                                                                \
      int64_t nc = imin(NC_, n - j0);                                                                        \
      for (int64_t l0 = 0; l0 < k; l0 += KC_) {                                                              \
        int64_t kc = imin(KC_, k - l0);                                                                      \
        NAME##_pack_b(op_b, kc, nc, op_b == NT ? b + l0 + j0 * ldb : b + j0 + l0 * ldb, ldb, bbuf);          \
        T beta_panel = (l0 == 0) ? beta : (T)1;                                                              \
        for (int64_t i0 = 0; i0 < m; i0 += MC_) {                                                            \
          int64_t mc = imin(MC_, m - i0);                                                                    \
          NAME##_pack_a(op_a, mc, kc, op_a == NT ? a + i0 + l0 * lda : a + l0 + i0 * lda, lda, abuf);        \
          T *cb = c + i0 + j0 * ldc;                                                                         \
          for (int64_t jp = 0; jp * NR_ < nc; ++jp) {                                                        \
            int64_t nr = imin(NR_, nc - jp * NR_);                                                           \
            for (int64_t ip = 0; ip * MR_ < mc; ++ip) {                                                      \
              int64_t mr = imin(MR_, mc - ip * MR_);                                                         \
              const T *ap = abuf + ip * kc * MR_, *bp = bbuf + jp * kc * NR_;                                \
              T *cp = cb + ip * MR_ + jp * NR_ * ldc;                                                        \
              if (mr == MR_ && nr == NR_) NAME##_micro(kc, ap, bp, cp, ldc, alpha, beta_panel);              \
              else NAME##_edge(mr, nr, kc, ap, bp, cp, ldc, alpha, beta_panel);                              \
            }                                                                                                \
          }                                                                                                  \
        }                                                                                                    \
      }                                                                                                      \
    }                                                                                                        \
    free(abuf);                                                                                              \
    free(bbuf);                                                                                              \
  }

DEFINE_REAL_GEMM(sgemm, float, fmaf, 384, 576, 256, 8, 12)
DEFINE_REAL_GEMM(dgemm, double, fma, 258, 384, 256, 6, 8)

/* ======================================================================== */
/* Complex GEMM (4M)                                                        */
/*   cgemm MC/NC/KC = 384/576/256, MR x NR = 4 x 8 (cgemm.rs:114-116,        */
/*   c32_packers.rs:1-2); zgemm 256/512/96, 2 x 4 (zgemm.rs:113-115,          */
/*   c64_packers.rs:1-2).  Packing stores [MR re | MR im] per k and negates   */
/*   the imaginary part for ConjTranspose (c32_packers.rs:36-47, :101).       */
/*   k-step (microkernel/c32_mrxnr.rs:63-116): re = fma(ar,br,re);            */
/*   re = re - ai*bi (fused); im = fma(ar,bi,im); im = fma(ai,br,im).         */
/*   alpha (:119-136): out_re = acc_re*ar, out_re -= acc_im*ai (fused);       */
/*   out_im = acc_im*ar, out_im += acc_re*ai (fused); beta likewise (:138-156)*/
/*   edge (microkernel/c32_edge.rs:5-84): not fused.                          */
/* ======================================================================== */
#define DEFINE_CPLX_GEMM(NAME, T, FMA, MC_, NC_, KC_, MR_, NR_)                                              \
  static void NAME##_pack_a(int op, int64_t mc, int64_t kc, const T *a, int64_t lda, T *dst) {             \
    for (int64_t p = 0; p < mc; p += MR_)                                                                    \
      for (int64_t k = 0; k < kc; ++k) {                                                                     \
        for (int t = 0; t < MR_; ++t) {                                                                      \
          int64_t i = p + t;                                                                                 \
          T re = 0, im = 0;                                                                                  \
          if (i < mc) {                                                                                      \
            const T *e = op == NT ? a + 2 * (i + k * lda) : a + 2 * (k + i * lda);                          \
            re = e[0];                                                                                       \
            im = op == CT ? -e[1] : e[1];                                                                    \
          }                                                                                                  \
          dst[t] = re;                                                                                       \
          dst[MR_ + t] = im;                                                                                 \
        }                                                                                                    \
        dst += 2 * MR_;                                                                                      \
      }                                                                                                      \
  }                                                                                                          \
  static void NAME##_pack_b(int op, int64_t kc, int64_t nc, const T *b, int64_t ldb, T *dst) {             \
    for (int64_t p = 0; p < nc; p += NR_)                                                                    \
      for (int64_t k = 0; k < kc; ++k) {                                                                     \
        for (int t = 0; t < NR_; ++t) {                                                                      \
          int64_t j = p + t;                                                                                 \
          T re = 0, im = 0;                                                                                  \
          if (j < nc) {                                                                                      \
            const T *e = op == NT ? b + 2 * (k + j * ldb) : b + 2 * (j + k * ldb);                          \
            re = e[0];                                                                                       \
            im = op == CT ? -e[1] : e[1];                                                                    \
          }                                                                                                  \
          dst[t] = re;                                                                                       \
          dst[NR_ + t] = im;                                                                                 \
        }                                                                                                    \
        dst += 2 * NR_;                                                                                      \
      }                                                                                                      \
  }                                                                                                          \
  static void NAME##_tile(int edge, int64_t mr, int64_t nr, int64_t kc, const T *ap, const T *bp, T *c,     \
                          int64_t ldc, const T *alpha, const T *beta) {                                      \
    T are[MR_][NR_], aim[MR_][NR_];                                                                          \
    memset(are, 0, sizeof(are));                                                                             \
    memset(aim, 0, sizeof(aim));                                                                             \
    for (int64_t k = 0; k < kc; ++k, ap += 2 * MR_, bp += 2 * NR_)                                           \
      for (int64_t j = 0; j < nr; ++j)                                                                       \
        for (int64_t r = 0; r < mr; ++r) {                                                                   \
          T ar = ap[r], ai = ap[MR_ + r], br = bp[j], bi = bp[NR_ + j];                                      \
          if (!edge) {                                                                                       \
            are[r][j] = FMA(ar, br, are[r][j]);                                                              \
            are[r][j] = FMA(-ai, bi, are[r][j]);                                                             \
            aim[r][j] = FMA(ar, bi, aim[r][j]);                                                              \
            aim[r][j] = FMA(ai, br, aim[r][j]);                                                              \
          } else {                                                                                           \
            are[r][j] += ar * br - ai * bi;                                                                  \
            aim[r][j] += ar * bi + ai * br;                                                                  \
          }                                                                                                  \
        }                                                                                                    \
    const T xr = alpha[0], xi = alpha[1], yr = beta[0], yi = beta[1];                                        \
    const int b0 = yr == 0 && yi == 0, b1 = yr == 1 && yi == 0;                                              \
    for (int64_t j = 0; j < nr; ++j)                                                                         \
      for (int64_t r = 0; r < mr; ++r) {                                                                     \
        T *cp = c + 2 * (r + j * ldc);                                                                       \
        T add_re, add_im;                                                                                    \
        if (!edge) {                                                                                         \
          add_re = FMA(-aim[r][j], xi, are[r][j] * xr);                                                      \
          add_im = FMA(are[r][j], xi, aim[r][j] * xr);                                                       \
        } else {                                                                                             \
          add_re = are[r][j] * xr - aim[r][j] * xi;                                                          \
          add_im = are[r][j] * xi + aim[r][j] * xr;                                                          \
        }                                                                                                    \
        if (b0) { cp[0] = add_re; cp[1] = add_im; }                                                          \
        else if (b1) { cp[0] += add_re; cp[1] += add_im; }                                                   \
        else {                                                                                               \
          T cr = cp[0], ci = cp[1], sr, si;                                                                  \
          if (!edge) { sr = FMA(-ci, yi, cr * yr); si = FMA(cr, yi, ci * yr); }                              \
          else { sr = cr * yr - ci * yi; si = cr * yi + ci * yr; }                                           \
          cp[0] = sr + add_re;                                                                               \
          cp[1] = si + add_im;                                                                               \
        }                                                                                                    \
      }                                                                                                      \
  }                                                                                                          \
  void port_##NAME(int op_a, int op_b, int64_t m, int64_t n, int64_t k, const T *alpha, const T *a,        \
                   int64_t lda, const T *b, int64_t ldb, const T *beta, T *c, int64_t ldc) {                 \
    if ((alpha[0] == 0 && alpha[1] == 0) || k == 0) { /* cgemm_nn.rs:42-69 */                                \
      for (int64_t j = 0; j < n; ++j)                                                                        \
        for (int64_t i = 0; i < m; ++i) {                                                                    \
          T *p = c + 2 * (i + j * ldc);                                                                      \
          if (beta[0] == 0 && beta[1] == 0) { p[0] = 0; p[1] = 0; }                                          \
          else if (!(beta[0] == 1 && beta[1] == 0)) {                                                        \
            T re = p[0], im = p[1];                                                                          \
            p[0] = re * beta[0] - im * beta[1];                                                              \
            p[1] = re * beta[1] + im * beta[0];                                                              \
          }                                                                                                  \
        }                                                                                                    \
      return;                                                                                                \
    }                                                                                                        \
    T *abuf = (T *)malloc(sizeof(T) * 2 * rup(MC_, MR_) * KC_);                                              \
    T *bbuf = (T *)malloc(sizeof(T) * 2 * KC_ * rup(NC_, NR_));                                              \
    const T one[2] = {1, 0};                                                                                 \
    for (int64_t j0 = 0; j0 < n; j0 += NC_) {                                                                \
      int64_t nc = imin(NC_, n - j0);                                                                        \
      for (int64_t l0 = 0; l0 < k; l0 += KC_) {                                                              \
        int64_t kc = imin(KC_, k - l0);                                                                      \
        NAME##_pack_b(op_b, kc, nc, op_b == NT ? b + 2 * (l0 + j0 * ldb) : b + 2 * (j0 + l0 * ldb), ldb,     \
                      bbuf);                                                                                 \
        const T *bp_ = (l0 == 0) ? beta : one;                                                               \
        for (int64_t i0 = 0; i0 < m; i0 += MC_) {                                                            \
          int64_t mc = imin(MC_, m - i0);                                                                    \
          NAME##_pack_a(op_a, mc, kc, op_a == NT ? a + 2 * (i0 + l0 * lda) : a + 2 * (l0 + i0 * lda), lda,   \
                        abuf);                                                                               \
          for (int64_t jp = 0; jp * NR_ < nc; ++jp) {                                                        \
            int64_t nr = imin(NR_, nc - jp * NR_);                                                           \
            for (int64_t ip = 0; ip * MR_ < mc; ++ip) {                                                      \
              int64_t mr = imin(MR_, mc - ip * MR_);                                                         \
              NAME##_tile(!(mr == MR_ && nr == NR_), mr, nr, kc, abuf + 2 * ip * kc * MR_,                   \
                          bbuf + 2 * jp * kc * NR_, c + 2 * (i0 + ip * MR_ + (j0 + jp * NR_) * ldc), ldc,    \
                          alpha, bp_);                                                                       \
            }                                                                                                \
          }                                                                                                  \
        }                                                                                                    \
      }                                                                                                      \
    }                                                                                                        \
    free(abuf);                                                                                              \
    free(bbuf);                                                                                              \
  }

DEFINE_CPLX_GEMM(cgemm, float, fmaf, 384, 576, 256, 4, 8)
DEFINE_CPLX_GEMM(zgemm, double, fma, 256, 512, 96, 2, 4)

/* ======================================================================== */
/* GEMV                                                                     */
/*   dgemv.rs:61-79 dispatch; NoTranspose dgemv_notranspose.rs:62-177:       */
/*   quick return (:75-76), y := beta*y by dscal (:88, dscal.rs:30-66 so     */
/*   NaN*0 stays NaN), return if alpha == 0 (:89), xbuf = alpha*x (:93),     */
/*   then daxpyf (daxpyf.rs:45): y_i = fma(a_ij, xbuf_j, y_i), columns in    */
/*   order.  Transpose dgemv_transpose.rs:63-181 -> ddotf (ddotf.rs:37):     */
/*   y_j += sum_i a_ij * xbuf_i (fused accumulation, rows in order here --   */
/*   CORAL's 128-row partial sums change only rounding).                    */
/*   Complex (cgemv.rs:67-84): same with complex mul, conj for C.            */
/* ======================================================================== */
#define DEFINE_REAL_GEMV(NAME, T, FMA)                                                                       \
  void port_##NAME(int trans, int64_t m, int64_t n, T alpha, const T *a, int64_t lda, const T *x,          \
                   int64_t incx, T beta, T *y, int64_t incy) {                                               \
    if (m == 0 || n == 0) return;                                                                            \
    if (alpha == (T)0 && beta == (T)1) return;                                                               \
    const int nt = trans == NT;                                                                              \
    const int64_t ylen = nt ? m : n, xlen = nt ? n : m;                                                      \
    if (beta != (T)1)                                                                                        \
      for (int64_t i = 0; i < ylen; ++i) y[i * incy] = y[i * incy] * beta;                                   \
    if (alpha == (T)0) return;                                                                               \
    T *xb = (T *)malloc(sizeof(T) * xlen);                                                                   \
    for (int64_t i = 0; i < xlen; ++i) xb[i] = x[i * incx] * alpha;                                          \
    if (nt) {                                                                                                \
      for (int64_t j = 0; j < n; ++j)                                                                        \
        for (int64_t i = 0; i < m; ++i) y[i * incy] = FMA(a[i + j * lda], xb[j], y[i * incy]);               \
    } else {                                                                                                 \
      for (int64_t j = 0; j < n; ++j) {                                                                      \
        T s = 0;                                                                                             \
        for (int64_t i = 0; i < m; ++i) s = FMA(a[i + j * lda], xb[i], s);                                   \
        y[j * incy] += s;                                                                                    \
      }                                                                                                      \
    }                                                                                                        \
    free(xb);                                                                                                \
  }
DEFINE_REAL_GEMV(sgemv, float, fmaf)
DEFINE_REAL_GEMV(dgemv, double, fma)

#define DEFINE_CPLX_GEMV(NAME, T)                                                                            \
  void port_##NAME(int trans, int64_t m, int64_t n, const T *alpha, const T *a, int64_t lda, const T *x,   \
                   int64_t incx, const T *beta, T *y, int64_t incy) {                                        \
    if (m == 0 || n == 0) return;                                                                            \
    if (alpha[0] == 0 && alpha[1] == 0 && beta[0] == 1 && beta[1] == 0) return;                             \
    const int nt = trans == NT, cj = trans == CT;                                                            \
    const int64_t ylen = nt ? m : n, xlen = nt ? n : m;                                                      \
    if (!(beta[0] == 1 && beta[1] == 0))                                                                     \
      for (int64_t i = 0; i < ylen; ++i) {                                                                   \
        T *p = y + 2 * i * incy, re = p[0], im = p[1];                                                       \
        p[0] = re * beta[0] - im * beta[1];                                                                  \
        p[1] = re * beta[1] + im * beta[0];                                                                  \
      }                                                                                                      \
    if (alpha[0] == 0 && alpha[1] == 0) return;                                                              \
    T *xb = (T *)malloc(sizeof(T) * 2 * xlen);                                                               \
    for (int64_t i = 0; i < xlen; ++i) {                                                                     \
      const T *p = x + 2 * i * incx;                                                                         \
      xb[2 * i] = p[0] * alpha[0] - p[1] * alpha[1];                                                         \
      xb[2 * i + 1] = p[0] * alpha[1] + p[1] * alpha[0];                                                     \
    }                                                                                                        \
    if (nt) {                                                                                                \
      for (int64_t j = 0; j < n; ++j)                                                                        \
        for (int64_t i = 0; i < m; ++i) {                                                                    \
          const T *e = a + 2 * (i + j * lda);                                                                \
          T *p = y + 2 * i * incy;                                                                           \
          p[0] += e[0] * xb[2 * j] - e[1] * xb[2 * j + 1];                                                   \
          p[1] += e[0] * xb[2 * j + 1] + e[1] * xb[2 * j];                                                   \
        }                                                                                                    \
    } else {                                                                                                 \
      for (int64_t j = 0; j < n; ++j) {                                                                      \
        T sr = 0, si = 0;                                                                                    \
        for (int64_t i = 0; i < m; ++i) {                                                                    \
          const T *e = a + 2 * (i + j * lda);                                                                \
          T er = e[0], ei = cj ? -e[1] : e[1];                                                               \
          sr += er * xb[2 * i] - ei * xb[2 * i + 1];                                                         \
          si += er * xb[2 * i + 1] + ei * xb[2 * i];                                                         \
        }                                                                                                    \
        y[2 * j * incy] += sr;                                                                               \
        y[2 * j * incy + 1] += si;                                                                           \
      }                                                                                                      \
    }                                                                                                        \
    free(xb);                                                                                                \
  }
DEFINE_CPLX_GEMV(cgemv, float)
DEFINE_CPLX_GEMV(zgemv, double)

/* ======================================================================== */
/* SYMV / HEMV: the one-pass per-column loop of chemv.rs:142-231 (upper     */
/* :150-183, lower :186-229): y_i += a_ij*x'_j for the stored i, the        */
/* column's conjugated dot accumulates into y_j with Re(a_jj) (:178-182).   */
/* HEMV zero-fills y when beta == 0 (:135-139); SYMV follows dsymv.rs:99.   */
/* Only the stored triangle is read.                                        */
/* ======================================================================== */
#define DEFINE_SYMV(NAME, T)                                                                                 \
  void port_##NAME(int uplo, int64_t n, T alpha, const T *a, int64_t lda, const T *x, int64_t incx, T beta,\
                   T *y, int64_t incy) {                                                                     \
    if (n == 0 || (alpha == (T)0 && beta == (T)1)) return;                                                   \
    T *xb = (T *)malloc(sizeof(T) * n);                                                                      \
    for (int64_t i = 0; i < n; ++i) xb[i] = x[i * incx] * alpha;                                             \
    for (int64_t i = 0; i < n; ++i) y[i * incy] = (beta == (T)0) ? (T)0 : y[i * incy] * beta;                \
    for (int64_t j = 0; j < n; ++j) {                                                                        \
      T acc = 0;                                                                                             \
      int64_t lo = uplo == 0 ? 0 : j + 1, hi = uplo == 0 ? j : n;                                            \
      for (int64_t i = lo; i < hi; ++i) {                                                                    \
        T aij = a[i + j * lda];                                                                              \
        y[i * incy] += aij * xb[j];                                                                          \
        acc += aij * xb[i];                                                                                  \
      }                                                                                                      \
      y[j * incy] += a[j + j * lda] * xb[j] + acc;                                                           \
    }                                                                                                        \
    free(xb);                                                                                                \
  }
DEFINE_SYMV(ssymv, float)
DEFINE_SYMV(dsymv, double)

#define DEFINE_HEMV(NAME, T)                                                                                 \
  void port_##NAME(int uplo, int64_t n, const T *alpha, const T *a, int64_t lda, const T *x, int64_t incx, \
                   const T *beta, T *y, int64_t incy) {                                                      \
    if (n == 0) return;                                                                                      \
    if (alpha[0] == 0 && alpha[1] == 0 && beta[0] == 1 && beta[1] == 0) return;                             \
    T *xb = (T *)malloc(sizeof(T) * 2 * n);                                                                  \
    for (int64_t i = 0; i < n; ++i) {                                                                        \
      const T *p = x + 2 * i * incx;                                                                         \
      xb[2 * i] = p[0] * alpha[0] - p[1] * alpha[1];                                                         \
      xb[2 * i + 1] = p[0] * alpha[1] + p[1] * alpha[0];                                                     \
    }                                                                                                        \
    for (int64_t i = 0; i < n; ++i) {                                                                        \
      T *p = y + 2 * i * incy;                                                                               \
      if (beta[0] == 0 && beta[1] == 0) { p[0] = 0; p[1] = 0; }                                              \
      else if (!(beta[0] == 1 && beta[1] == 0)) {                                                            \
        T re = p[0], im = p[1];                                                                              \
        p[0] = re * beta[0] - im * beta[1];                                                                  \
        p[1] = re * beta[1] + im * beta[0];                                                                  \
      }                                                                                                      \
    }                                                                                                        \
    for (int64_t j = 0; j < n; ++j) {                                                                        \
      const T xr = xb[2 * j], xi = xb[2 * j + 1];                                                            \
      T accr = 0, acci = 0;                                                                                  \
      int64_t lo = uplo == 0 ? 0 : j + 1, hi = uplo == 0 ? j : n;                                            \
      for (int64_t i = lo; i < hi; ++i) {                                                                    \
        const T ar = a[2 * (i + j * lda)], ai = a[2 * (i + j * lda) + 1];                                    \
        T *p = y + 2 * i * incy;                                                                             \
        p[0] += ar * xr - ai * xi;                                                                           \
        p[1] += ar * xi + ai * xr;                                                                           \
        accr += ar * xb[2 * i] + ai * xb[2 * i + 1];                                                         \
        acci += ar * xb[2 * i + 1] - ai * xb[2 * i];                                                         \
      }                                                                                                      \
      const T ajj = a[2 * (j + j * lda)];                                                                    \
      y[2 * j * incy] += ajj * xr + accr;                                                                    \
      y[2 * j * incy + 1] += ajj * xi + acci;                                                                \
    }                                                                                                        \
    free(xb);                                                                                                \
  }
DEFINE_HEMV(chemv, float)
DEFINE_HEMV(zhemv, double)

/* ======================================================================== */
/* TRSV: lower/NoTrans is dtrlsv.rs:292-347: NB = 8 diagonal blocks solved  */
/* by forward substitution (:65-139: sum += a_ik*x_k not fused, x_i -= sum, */
/* x_i /= a_ii unless unit), then the rows below are updated with daxpyf    */
/* (:230-257: x_tail = fma(a, -x_blk, x_tail)).  The other variants are     */
/* restated as the same substitution on op(A) in the solve order.  Complex  */
/* conj divides by conj(a_ii) with den = re^2 + im^2 (ctrlsv.rs:205-214).   */
/* TRMV: x := op(A) x (dtrmv.rs:58), sums in column order.                  */
/* ======================================================================== */
static inline int tr_forward(int uplo, int trans) { return (uplo == 1) == (trans == NT); }

#define DEFINE_REAL_TR(P, T, FMA)                                                                            \
  static T P##_op(const T *a, int64_t lda, int trans, int64_t i, int64_t j) {                                \
    return trans == NT ? a[i + j * lda] : a[j + i * lda];                                                    \
  }                                                                                                          \
  void port_##P##trsv(int uplo, int trans, int diag, int64_t n, const T *a, int64_t lda, T *x, int64_t incx) \
  {                                                                                                          \
    if (trans == CT) trans = TR;                                                                             \
    const int unit = diag == 0, fwd = tr_forward(uplo, trans);                                               \
    if (fwd && trans == NT) { /* dtrlsv_notranspose */                                                       \
      for (int64_t d = 0; d < n; d += 8) {                                                                   \
        int64_t nb = imin(8, n - d);                                                                         \
        for (int64_t i = d; i < d + nb; ++i) {                                                               \
          T sum = 0;                                                                                         \
          for (int64_t k = d; k < i; ++k) sum += a[i + k * lda] * x[k * incx];                               \
          T xi = x[i * incx] - sum;                                                                          \
          if (!unit) xi /= a[i + i * lda];                                                                   \
          x[i * incx] = xi;                                                                                  \
        }                                                                                                    \
        for (int64_t k = d; k < d + nb; ++k) {                                                               \
          T nx = -x[k * incx];                                                                               \
          for (int64_t i = d + nb; i < n; ++i) x[i * incx] = FMA(a[i + k * lda], nx, x[i * incx]);           \
        }                                                                                                    \
      }                                                                                                      \
      return;                                                                                                \
    }                                                                                                        \
    for (int64_t s = 0; s < n; ++s) {                                                                        \
      int64_t i = fwd ? s : n - 1 - s;                                                                       \
      T sum = 0;                                                                                             \
      for (int64_t t = 0; t < s; ++t) {                                                                      \
        int64_t k = fwd ? t : n - 1 - t;                                                                     \
        sum += P##_op(a, lda, trans, i, k) * x[k * incx];                                                    \
      }                                                                                                      \
      T xi = x[i * incx] - sum;                                                                              \
      if (!unit) xi /= a[i + i * lda];                                                                       \
      x[i * incx] = xi;                                                                                      \
    }                                                                                                        \
  }                                                                                                          \
  void port_##P##trmv(int uplo, int trans, int diag, int64_t n, const T *a, int64_t lda, T *x, int64_t incx) \
  {                                                                                                          \
    if (trans == CT) trans = TR;                                                                             \
    const int unit = diag == 0, fwd = tr_forward(uplo, trans);                                               \
    T *y = (T *)malloc(sizeof(T) * (n ? n : 1));                                                             \
    for (int64_t i = 0; i < n; ++i) {                                                                        \
      T s = unit ? x[i * incx] : P##_op(a, lda, trans, i, i) * x[i * incx];                                  \
      for (int64_t k = 0; k < n; ++k) {                                                                      \
        if (k == i) continue;                                                                                \
        if (fwd ? (k > i) : (k < i)) continue;                                                               \
        s += P##_op(a, lda, trans, i, k) * x[k * incx];                                                      \
      }                                                                                                      \
      y[i] = s;                                                                                              \
    }                                                                                                        \
    for (int64_t i = 0; i < n; ++i) x[i * incx] = y[i];                                                      \
    free(y);                                                                                                 \
  }
DEFINE_REAL_TR(s, float, fmaf)
DEFINE_REAL_TR(d, double, fma)

#define DEFINE_CPLX_TR(P, T)                                                                                 \
  static void P##_cop(const T *a, int64_t lda, int trans, int64_t i, int64_t j, T *re, T *im) {            \
    const T *e = trans == NT ? a + 2 * (i + j * lda) : a + 2 * (j + i * lda);                                \
    *re = e[0];                                                                                              \
    *im = trans == CT ? -e[1] : e[1];                                                                        \
  }                                                                                                          \
  void port_##P##trsv(int uplo, int trans, int diag, int64_t n, const T *a, int64_t lda, T *x, int64_t incx) \
  {                                                                                                          \
    const int unit = diag == 0, fwd = tr_forward(uplo, trans);                                               \
    for (int64_t s = 0; s < n; ++s) {                                                                        \
      int64_t i = fwd ? s : n - 1 - s;                                                                       \
      T sr = 0, si = 0;                                                                                      \
      for (int64_t t = 0; t < s; ++t) {                                                                      \
        int64_t k = fwd ? t : n - 1 - t;                                                                     \
        T ar, ai;                                                                                            \
        P##_cop(a, lda, trans, i, k, &ar, &ai);                                                              \
        T xr = x[2 * k * incx], xi = x[2 * k * incx + 1];                                                    \
        sr += ar * xr - ai * xi;                                                                             \
        si += ar * xi + ai * xr;                                                                             \
      }                                                                                                      \
      T vr = x[2 * i * incx] - sr, vi = x[2 * i * incx + 1] - si;                                            \
      if (!unit) {                                                                                           \
        T dr, di;                                                                                            \
        P##_cop(a, lda, trans, i, i, &dr, &di);                                                              \
        T den = dr * dr + di * di;                                                                           \
        T qr = (vr * dr + vi * di) / den, qi = (vi * dr - vr * di) / den;                                    \
        vr = qr;                                                                                             \
        vi = qi;                                                                                             \
      }                                                                                                      \
      x[2 * i * incx] = vr;                                                                                  \
      x[2 * i * incx + 1] = vi;                                                                              \
    }                                                                                                        \
  }                                                                                                          \
  void port_##P##trmv(int uplo, int trans, int diag, int64_t n, const T *a, int64_t lda, T *x, int64_t incx) \
  {                                                                                                          \
    const int unit = diag == 0, fwd = tr_forward(uplo, trans);                                               \
    T *y = (T *)malloc(sizeof(T) * 2 * (n ? n : 1));                                                         \
    for (int64_t i = 0; i < n; ++i) {                                                                        \
      T sr, si;                                                                                              \
      T xr0 = x[2 * i * incx], xi0 = x[2 * i * incx + 1];                                                    \
      if (unit) { sr = xr0; si = xi0; }                                                                      \
      else {                                                                                                 \
        T dr, di;                                                                                            \
        P##_cop(a, lda, trans, i, i, &dr, &di);                                                              \
        sr = dr * xr0 - di * xi0;                                                                            \
        si = dr * xi0 + di * xr0;                                                                            \
      }                                                                                                      \
      for (int64_t k = 0; k < n; ++k) {                                                                      \
        if (k == i || (fwd ? (k > i) : (k < i))) continue;                                                   \
        T ar, ai;                                                                                            \
        P##_cop(a, lda, trans, i, k, &ar, &ai);                                                              \
        T xr = x[2 * k * incx], xi = x[2 * k * incx + 1];                                                    \
        sr += ar * xr - ai * xi;                                                                             \
        si += ar * xi + ai * xr;                                                                             \
      }                                                                                                      \
      y[2 * i] = sr;                                                                                         \
      y[2 * i + 1] = si;                                                                                     \
    }                                                                                                        \
    for (int64_t i = 0; i < n; ++i) {                                                                        \
      x[2 * i * incx] = y[2 * i];                                                                            \
      x[2 * i * incx + 1] = y[2 * i + 1];                                                                    \
    }                                                                                                        \
    free(y);                                                                                                 \
  }
DEFINE_CPLX_TR(c, float)
DEFINE_CPLX_TR(z, double)

/* ======================================================================== */
/* Level 1                                                                  */
/* ======================================================================== */
/* sdot.rs:42-161: 8 vector accumulators over 32-element chunks, then 16-   */
/* and 4-element steps into acc0, pairwise combine (:132-135), horizontal   */
/* add, scalar tail (not fused).  ddot.rs:43-161 has the same shape with 2   */
/* lanes per vector (16-element chunks).                                    */
#define DEFINE_DOT(NAME, T, FMA, L)                                                                          \
  T port_##NAME(int64_t n, const T *x, int64_t incx, const T *y, int64_t incy) {                            \
    if (n == 0) return 0;                                                                                    \
    if (incx == 1 && incy == 1) {                                                                            \
      T acc[8][L];                                                                                           \
      memset(acc, 0, sizeof(acc));                                                                           \
      int64_t i = 0;                                                                                         \
      for (; i + 8 * L <= n; i += 8 * L)                                                                     \
        for (int v = 0; v < 8; ++v)                                                                          \
          for (int l = 0; l < L; ++l) acc[v][l] = FMA(x[i + v * L + l], y[i + v * L + l], acc[v][l]);      \
      for (; i + 4 * L <= n; i += 4 * L)                                                                     \
        for (int v = 0; v < 4; ++v)                                                                          \
          for (int l = 0; l < L; ++l) acc[v][l] = FMA(x[i + v * L + l], y[i + v * L + l], acc[v][l]);      \
      for (; i + L <= n; i += L)                                                                             \
        for (int l = 0; l < L; ++l) acc[0][l] = FMA(x[i + l], y[i + l], acc[0][l]);                          \
      T v1[L], v2[L], v[L];                                                                                  \
      for (int l = 0; l < L; ++l) {                                                                          \
        v1[l] = (acc[0][l] + acc[1][l]) + (acc[2][l] + acc[3][l]);                                           \
        v2[l] = (acc[4][l] + acc[5][l]) + (acc[6][l] + acc[7][l]);                                           \
        v[l] = v1[l] + v2[l];                                                                                \
      }                                                                                                      \
      T s = (L == 4) ? ((v[0] + v[1]) + (v[2 % L] + v[3 % L])) : (v[0] + v[1 % L]);                          \
      for (; i < n; ++i) s += x[i] * y[i];                                                                   \
      return s;                                                                                              \
    }                                                                                                        \
    T s = 0;                                                                                                 \
    for (int64_t i = 0; i < n; ++i) s += x[i * incx] * y[i * incy];                                          \
    return s;                                                                                                \
  }
DEFINE_DOT(sdot, float, fmaf, 4)
DEFINE_DOT(ddot, double, fma, 2)

/* cdotu.rs:46 / cdotc.rs:46 (conj(x) * y): complex sums in order. */
#define DEFINE_CDOT(NAME, T, CONJ)                                                                           \
  void port_##NAME(int64_t n, const T *x, int64_t incx, const T *y, int64_t incy, T *out) {                 \
    T sr = 0, si = 0;                                                                                        \
    for (int64_t i = 0; i < n; ++i) {                                                                        \
      T xr = x[2 * i * incx], xi = CONJ ? -x[2 * i * incx + 1] : x[2 * i * incx + 1];                        \
      T yr = y[2 * i * incy], yi = y[2 * i * incy + 1];                                                      \
      sr += xr * yr - xi * yi;                                                                               \
      si += xr * yi + xi * yr;                                                                               \
    }                                                                                                        \
    out[0] = sr;                                                                                             \
    out[1] = si;                                                                                             \
  }
DEFINE_CDOT(cdotu, float, 0)
DEFINE_CDOT(cdotc, float, 1)
DEFINE_CDOT(zdotu, double, 0)
DEFINE_CDOT(zdotc, double, 1)

/* daxpy.rs:42-132: quick return n == 0 or alpha == 0 (:51-53); fused body. */
/* How CORAL's axpy family rounds one element update y += c*x:                */
/*   AXPY_UNIT  unit-stride axpy: the NEON body (fused) covers the first      */
/*              floor(len/Q)*Q elements, the scalar tail is `y += c*x` (two   */
/*              roundings).  Q = 8 saxpy (coral-aarch64/src/level1/saxpy.rs:  */
/*              63-112), 4 daxpy (daxpy.rs:63-112), 4 caxpy (caxpy.rs:70-146), */
/*              2 zaxpy (zaxpy.rs:69-145).                                     */
/*   AXPY_PLAIN strided axpy (saxpy.rs:114-126 and alike): two roundings.      */
/*   AXPY_FUSED the inline `x.mul_add(c, a)` loops of the strided real rank    */
/*              updates: fused everywhere.                                    */
/* A fused complex update is yr = fma(ar,xr,yr); yr = fma(-ai,xi,yr);          */
/* yi = fma(ar,xi,yi); yi = fma(ai,xr,yi) (vfmaq/vfmsq order); a plain one is  */
/* yr += ar*xr - ai*xi; yi += ar*xi + ai*xr.                                   */
#define AXPY_PLAIN 0
#define AXPY_UNIT 1
#define AXPY_FUSED 2

#define DEFINE_AXPY_SEG(NAME, T, FMA, Q)                                                                     \
  static void NAME##_seg(int64_t len, T c, const T *x, int64_t incx, T *y, int64_t incy, int mode) {        \
    int64_t nf = mode == AXPY_FUSED ? len : mode == AXPY_UNIT ? len / Q * Q : 0;                             \
    for (int64_t i = 0; i < len; ++i) {                                                                      \
      T *yp = y + i * incy, xv = x[i * incx];                                                                \
      *yp = i < nf ? FMA(xv, c, *yp) : *yp + c * xv;                                                         \
    }                                                                                                        \
  }
DEFINE_AXPY_SEG(saxpy, float, fmaf, 8)
DEFINE_AXPY_SEG(daxpy, double, fma, 4)

#define DEFINE_CAXPY_SEG(NAME, T, FMA, Q)                                                                    \
  static void NAME##_seg(int64_t len, T ar, T ai, const T *x, int64_t incx, T *y, int64_t incy, int mode) { \
    int64_t nf = mode == AXPY_FUSED ? len : mode == AXPY_UNIT ? len / Q * Q : 0;                             \
    for (int64_t i = 0; i < len; ++i) {                                                                      \
      T xr = x[2 * i * incx], xi = x[2 * i * incx + 1];                                                      \
      T yr = y[2 * i * incy], yi = y[2 * i * incy + 1];                                                      \
      if (i < nf) {                                                                                          \
        yr = FMA(ar, xr, yr);                                                                                \
        yr = FMA(-ai, xi, yr);                                                                               \
        yi = FMA(ar, xi, yi);                                                                                \
        yi = FMA(ai, xr, yi);                                                                                \
      } else {                                                                                               \
        yr += ar * xr - ai * xi;                                                                             \
        yi += ar * xi + ai * xr;                                                                             \
      }                                                                                                      \
      y[2 * i * incy] = yr;                                                                                  \
      y[2 * i * incy + 1] = yi;                                                                              \
    }                                                                                                        \
  }
DEFINE_CAXPY_SEG(caxpy, float, fmaf, 4)
DEFINE_CAXPY_SEG(zaxpy, double, fma, 2)

static int axpy_mode(int64_t incx, int64_t incy) { return incx == 1 && incy == 1 ? AXPY_UNIT : AXPY_PLAIN; }

/* level1/saxpy.rs:42-131, daxpy.rs: quick return n == 0 || alpha == 0. */
#define DEFINE_AXPY(NAME, T)                                                                                 \
  void port_##NAME(int64_t n, T alpha, const T *x, int64_t incx, T *y, int64_t incy) {                      \
    if (n == 0 || alpha == (T)0) return;                                                                     \
    NAME##_seg(n, alpha, x, incx, y, incy, axpy_mode(incx, incy));                                           \
  }
DEFINE_AXPY(saxpy, float)
DEFINE_AXPY(daxpy, double)

/* level1/caxpy.rs:45-172, zaxpy.rs: quick return n == 0 || alpha == 0. */
#define DEFINE_CAXPY(NAME, T)                                                                                \
  void port_##NAME(int64_t n, const T *alpha, const T *x, int64_t incx, T *y, int64_t incy) {               \
    if (n == 0 || (alpha[0] == 0 && alpha[1] == 0)) return;                                                  \
    NAME##_seg(n, alpha[0], alpha[1], x, incx, y, incy, axpy_mode(incx, incy));                              \
  }
DEFINE_CAXPY(caxpy, float)
DEFINE_CAXPY(zaxpy, double)

/* ======================================================================== */
/* Level-2 rank updates: per column a coefficient, skipped when zero, then  */
/* an axpy into the stored rows of that column.  uplo 0 = upper (rows 0..j), */
/* 1 = lower (rows j..n-1).                                                  */
/* ======================================================================== */
static int64_t seg_start(int32_t uplo, int64_t j) { return uplo == 0 ? 0 : j; }
static int64_t seg_len(int32_t uplo, int64_t n, int64_t j) { return uplo == 0 ? j + 1 : n - j; }

/* coral-aarch64/src/level2/sger.rs:68-132, dger.rs:69-133: A += x * (alpha*y_j).  Quick return m == 0 ||    */
/* n == 0 || alpha == 0; a column with alpha*y_j == 0 is skipped.  Unit      */
/* strides go through saxpy (AXPY_UNIT), others through mul_add (FUSED).     */
#define DEFINE_GER(NAME, AXPY, T)                                                                            \
  void port_##NAME(int64_t m, int64_t n, T alpha, const T *x, int64_t incx, const T *y, int64_t incy, T *a, \
                   int64_t lda) {                                                                            \
    if (m == 0 || n == 0 || alpha == (T)0) return;                                                           \
    int mode = incx == 1 && incy == 1 ? AXPY_UNIT : AXPY_FUSED;                                              \
    for (int64_t j = 0; j < n; ++j) {                                                                        \
      T c = alpha * y[j * incy];                                                                             \
      if (c != (T)0) AXPY##_seg(m, c, x, incx, a + j * lda, 1, mode);                                        \
    }                                                                                                        \
  }
DEFINE_GER(sger, saxpy, float)
DEFINE_GER(dger, daxpy, double)

/* level2/cgeru.rs:63-144, cgerc.rs:62-143, zgeru.rs:61-142, zgerc.rs:60-141:
   coefficient alpha*y_j (u) or         */
/* alpha*conj(y_j) (c), skipped when zero.  Unit strides go through caxpy    */
/* (AXPY_UNIT); others use an inline non-fused loop (AXPY_PLAIN).           */
#define DEFINE_CGER(NAME, AXPY, T, CONJ)                                                                     \
  void port_##NAME(int64_t m, int64_t n, const T *alpha, const T *x, int64_t incx, const T *y, int64_t incy,\
                   T *a, int64_t lda) {                                                                      \
    if (m == 0 || n == 0 || (alpha[0] == 0 && alpha[1] == 0)) return;                                        \
    int mode = axpy_mode(incx, incy);                                                                        \
    for (int64_t j = 0; j < n; ++j) {                                                                        \
      T yr = y[2 * j * incy], yi = y[2 * j * incy + 1];                                                      \
      T cr = CONJ ? alpha[0] * yr + alpha[1] * yi : alpha[0] * yr - alpha[1] * yi;                           \
      T ci = CONJ ? -alpha[0] * yi + alpha[1] * yr : alpha[0] * yi + alpha[1] * yr;                          \
      if (cr != 0 || ci != 0) AXPY##_seg(m, cr, ci, x, incx, a + 2 * j * lda, 1, mode);                      \
    }                                                                                                        \
  }
DEFINE_CGER(cgeru, caxpy, float, 0)
DEFINE_CGER(cgerc, caxpy, float, 1)
DEFINE_CGER(zgeru, zaxpy, double, 0)
DEFINE_CGER(zgerc, zaxpy, double, 1)

/* level2/ssyr.rs:63-181, dsyr.rs:61-179: aj = alpha*x_j (skipped when zero), then the     */
/* stored rows of column j += x*aj; incx == 1 through saxpy, else mul_add.   */
#define DEFINE_SYR(NAME, AXPY, T)                                                                            \
  void port_##NAME(int32_t uplo, int64_t n, T alpha, const T *x, int64_t incx, T *a, int64_t lda) {         \
    if (n == 0 || alpha == (T)0) return;                                                                     \
    int mode = incx == 1 ? AXPY_UNIT : AXPY_FUSED;                                                           \
    for (int64_t j = 0; j < n; ++j) {                                                                        \
      T aj = alpha * x[j * incx];                                                                            \
      int64_t i0 = seg_start(uplo, j);                                                                       \
      if (aj != (T)0) AXPY##_seg(seg_len(uplo, n, j), aj, x + i0 * incx, incx, a + i0 + j * lda, 1, mode);   \
    }                                                                                                        \
  }
DEFINE_SYR(ssyr, saxpy, float)
DEFINE_SYR(dsyr, daxpy, double)

/* level2/ssyr2.rs:59-197, dsyr2.rs:63-201: aj_y = alpha*y_j, aj_x = alpha*x_j.  Unit     */
/* strides: saxpy of x with aj_y, then of y with aj_x, each skipped when     */
/* zero.  Otherwise, unless both are zero, one mul_add loop applies both     */
/* (a = fma(x_i, aj_y, a); a = fma(y_i, aj_x, a)), zero coefficients too.    */
#define DEFINE_SYR2(NAME, AXPY, T)                                                                           \
  void port_##NAME(int32_t uplo, int64_t n, T alpha, const T *x, int64_t incx, const T *y, int64_t incy,    \
                   T *a, int64_t lda) {                                                                      \
    if (n == 0 || alpha == (T)0) return;                                                                     \
    int unit = incx == 1 && incy == 1;                                                                       \
    for (int64_t j = 0; j < n; ++j) {                                                                        \
      T ajy = alpha * y[j * incy], ajx = alpha * x[j * incx];                                                \
      int64_t i0 = seg_start(uplo, j), len = seg_len(uplo, n, j);                                            \
      T *col = a + i0 + j * lda;                                                                             \
      if (unit) {                                                                                            \
        if (ajy != (T)0) AXPY##_seg(len, ajy, x + i0, 1, col, 1, AXPY_UNIT);                                 \
        if (ajx != (T)0) AXPY##_seg(len, ajx, y + i0, 1, col, 1, AXPY_UNIT);                                 \
      } else if (ajy != (T)0 || ajx != (T)0) {                                                               \
        AXPY##_seg(len, ajy, x + i0 * incx, incx, col, 1, AXPY_FUSED);                                       \
        AXPY##_seg(len, ajx, y + i0 * incy, incy, col, 1, AXPY_FUSED);                                       \
      }                                                                                                      \
    }                                                                                                        \
  }
DEFINE_SYR2(ssyr2, saxpy, float)
DEFINE_SYR2(dsyr2, daxpy, double)

/* level2/cher.rs:62-212, zher.rs:62-212 (alpha real): aj = [alpha*xr_j, -alpha*xi_j],     */
/* skipped when zero; caxpy(x, incx) into the stored rows of column j (fused */
/* body only for incx == 1), then Im(a_jj) = 0 for the updated column.       */
#define DEFINE_HER(NAME, AXPY, T)                                                                            \
  void port_##NAME(int32_t uplo, int64_t n, T alpha, const T *x, int64_t incx, T *a, int64_t lda) {         \
    if (n == 0 || alpha == (T)0) return;                                                                     \
    for (int64_t j = 0; j < n; ++j) {                                                                        \
      T ar = alpha * x[2 * j * incx], ai = -alpha * x[2 * j * incx + 1];                                     \
      if (ar == 0 && ai == 0) continue;                                                                      \
      int64_t i0 = seg_start(uplo, j);                                                                       \
      AXPY##_seg(seg_len(uplo, n, j), ar, ai, x + 2 * i0 * incx, incx, a + 2 * (i0 + j * lda), 1,            \
                 axpy_mode(incx, 1));                                                                        \
      a[2 * (j + j * lda) + 1] = 0;                                                                          \
    }                                                                                                        \
  }
DEFINE_HER(cher, caxpy, float)
DEFINE_HER(zher, zaxpy, double)

/* level2/cher2.rs:65-311, zher2.rs:63-309: aj_y = alpha*conj(y_j), aj_x =                */
/* conj(alpha)*conj(x_j); caxpy(x, incx) with aj_y, then caxpy(y, incy) with */
/* aj_x, each skipped when zero (fused body only for a unit source stride).  */
/* Im(a_jj) = 0: always for unit strides, otherwise only when a coefficient  */
/* is nonzero.                                                               */
#define DEFINE_HER2(NAME, AXPY, T)                                                                           \
  void port_##NAME(int32_t uplo, int64_t n, const T *alpha, const T *x, int64_t incx, const T *y,          \
                   int64_t incy, T *a, int64_t lda) {                                                        \
    if (n == 0 || (alpha[0] == 0 && alpha[1] == 0)) return;                                                  \
    T ar = alpha[0], ai = alpha[1];                                                                          \
    int unit = incx == 1 && incy == 1;                                                                       \
    for (int64_t j = 0; j < n; ++j) {                                                                        \
      T xr = x[2 * j * incx], xi = x[2 * j * incx + 1], yr = y[2 * j * incy], yi = y[2 * j * incy + 1];      \
      T ayr = ar * yr + ai * yi, ayi = -ar * yi + ai * yr;                                                   \
      T axr = ar * xr - ai * xi, axi = -ar * xi - ai * xr;                                                   \
      int any_y = ayr != 0 || ayi != 0, any_x = axr != 0 || axi != 0;                                        \
      int64_t i0 = seg_start(uplo, j), len = seg_len(uplo, n, j);                                            \
      T *col = a + 2 * (i0 + j * lda);                                                                       \
      if (any_y) AXPY##_seg(len, ayr, ayi, x + 2 * i0 * incx, incx, col, 1, axpy_mode(incx, 1));             \
      if (any_x) AXPY##_seg(len, axr, axi, y + 2 * i0 * incy, incy, col, 1, axpy_mode(incy, 1));             \
      if (unit || any_y || any_x) a[2 * (j + j * lda) + 1] = 0;                                              \
    }                                                                                                        \
  }
DEFINE_HER2(cher2, caxpy, float)
DEFINE_HER2(zher2, zaxpy, double)

/* snrm2.rs:45-149 with upd_f32 (nrm2_helpers.rs:15-38): scale = 0, ssq = 1;  */
/* chunks of 16 (f32) / 8 (f64): chunk max, normalise by 1/max, fused ssq,  */
/* merge; scalar tail with the classic update.  scnrm2/dznrm2 treat re and  */
/* im as consecutive entries.                                              */
#define DEFINE_NRM2(NAME, T, FMA, CH)                                                                        \
  static void NAME##_upd(T *scale, T *ssq, T cmax, T cssq) {                                                 \
    if (*scale < cmax) { T r = *scale / cmax; *ssq = *ssq * (r * r) + cssq; *scale = cmax; }                 \
    else if (*scale > 0) { T r = cmax / *scale; *ssq += cssq * (r * r); }                                    \
    else { *scale = cmax; *ssq = cssq; }                                                                     \
  }                                                                                                          \
  static T NAME##_core(int64_t n, const T *x, int64_t inc) {                                                 \
    T scale = 0, ssq = 1;                                                                                    \
    int64_t i = 0;                                                                                           \
    if (inc == 1) {                                                                                          \
      for (; i + CH <= n; i += CH) {                                                                         \
        T mx = 0;                                                                                            \
        for (int t = 0; t < CH; ++t) { T v = fabs(x[i + t]); mx = (v > mx) ? v : mx; }                       \
        if (mx > 0) {                                                                                        \
          T inv = (T)1 / mx, s = 0;                                                                          \
          for (int t = 0; t < CH; ++t) { T q = fabs(x[i + t]) * inv; s = FMA(q, q, s); }                     \
          NAME##_upd(&scale, &ssq, mx, s);                                                                   \
        }                                                                                                    \
      }                                                                                                      \
    }                                                                                                        \
    for (; i < n; ++i) {                                                                                     \
      T xi = x[i * inc];                                                                                     \
      if (xi != 0) {                                                                                         \
        T ax = fabs(xi);                                                                                     \
        if (scale < ax) { T r = scale / ax; ssq = 1 + ssq * (r * r); scale = ax; }                           \
        else if (scale > 0) { T r = ax / scale; ssq += r * r; }                                              \
        else scale = ax;                                                                                     \
      }                                                                                                      \
    }                                                                                                        \
    return scale * (T)sqrt(ssq);                                                                             \
  }                                                                                                          \
  T port_##NAME(int64_t n, const T *x, int64_t incx) {                                                       \
    if (n == 0) return 0;                                                                                    \
    return NAME##_core(n, x, incx);                                                                          \
  }
DEFINE_NRM2(snrm2, float, fmaf, 16)
DEFINE_NRM2(dnrm2, double, fma, 8)

float port_scnrm2(int64_t n, const float *x, int64_t incx) {
  if (n == 0) return 0;
  if (incx == 1) return snrm2_core(2 * n, x, 1);
  float *t = (float *)malloc(sizeof(float) * 2 * n);
  for (int64_t i = 0; i < n; ++i) { t[2 * i] = x[2 * i * incx]; t[2 * i + 1] = x[2 * i * incx + 1]; }
  float r = snrm2_core(2 * n, t, 1);
  free(t);
  return r;
}
double port_dznrm2(int64_t n, const double *x, int64_t incx) {
  if (n == 0) return 0;
  if (incx == 1) return dnrm2_core(2 * n, x, 1);
  double *t = (double *)malloc(sizeof(double) * 2 * n);
  for (int64_t i = 0; i < n; ++i) { t[2 * i] = x[2 * i * incx]; t[2 * i + 1] = x[2 * i * incx + 1]; }
  double r = dnrm2_core(2 * n, t, 1);
  free(t);
  return r;
}

/* isamax.rs:43-151 / icamax.rs:48 (|re| + |im| :79-82): first index of the   */
/* largest magnitude with strict '>', NaN never selected, n == 0 -> 0.       */
#define DEFINE_IAMAX(NAME, T, CPLX)                                                                          \
  int64_t port_##NAME(int64_t n, const T *x, int64_t incx) {                                                \
    if (n == 0 || incx <= 0) return 0;                                                                       \
    T best = -INFINITY;                                                                                      \
    int64_t bi = 0;                                                                                          \
    for (int64_t i = 0; i < n; ++i) {                                                                        \
      T v = CPLX ? (T)(fabs(x[2 * i * incx]) + fabs(x[2 * i * incx + 1])) : (T)fabs(x[i * incx]);          \
      if (v > best) { best = v; bi = i; }                                                                    \
    }                                                                                                        \
    return bi;                                                                                               \
  }
DEFINE_IAMAX(isamax, float, 0)
DEFINE_IAMAX(idamax, double, 0)
DEFINE_IAMAX(icamax, float, 1)
DEFINE_IAMAX(izamax, double, 1)

/* ======================================================================== */
/* Level 1: the rest of the call surface                                    */
/* ======================================================================== */
/* ARM ADDV order of a 4-lane float vector: (l0 + l1) + (l2 + l3).           */
static float hsum4(const float *v) { return (v[0] + v[1]) + (v[2] + v[3]); }

/* sasum.rs:40-107 (real) and scasum.rs:40-108 (complex, over the 2n floats of */
/* a unit-stride vector): four 4-lane accumulators over 16-float chunks,     */
/* 4-float steps into acc0, res += h(acc0) + h(acc1) + h(acc2) + h(acc3),    */
/* scalar tail; strided: res += |x| (complex: res += |re| + |im|).           */
static float sasum_unit(int64_t len, const float *x) {
  float acc[4][4];
  memset(acc, 0, sizeof(acc));
  int64_t i = 0;
  for (; i + 16 <= len; i += 16)
    for (int v = 0; v < 4; ++v)
      for (int l = 0; l < 4; ++l) acc[v][l] += fabsf(x[i + 4 * v + l]);
  for (; i + 4 <= len; i += 4)
    for (int l = 0; l < 4; ++l) acc[0][l] += fabsf(x[i + l]);
  float res = 0;
  res += ((hsum4(acc[0]) + hsum4(acc[1])) + hsum4(acc[2])) + hsum4(acc[3]);
  for (; i < len; ++i) res += fabsf(x[i]);
  return res;
}
float port_sasum(int64_t n, const float *x, int64_t incx) {
  if (n == 0 || incx == 0) return 0;
  if (incx == 1) return sasum_unit(n, x);
  float res = 0;
  for (int64_t i = 0; i < n; ++i) res += fabsf(x[i * incx]);
  return res;
}
float port_scasum(int64_t n, const float *x, int64_t incx) {
  if (n == 0 || incx == 0) return 0;
  if (incx == 1) return sasum_unit(2 * n, x);
  float res = 0;
  for (int64_t i = 0; i < n; ++i) res += fabsf(x[2 * i * incx]) + fabsf(x[2 * i * incx + 1]);
  return res;
}
/* dasum.rs:41-110: four 2-lane accumulators over 8-element chunks, then     */
/* 4-element steps into acc0 and acc1, tail.                                 */
double port_dasum(int64_t n, const double *x, int64_t incx) {
  if (n == 0 || incx == 0) return 0;
  double res = 0;
  if (incx == 1) {
    double acc[4][2];
    memset(acc, 0, sizeof(acc));
    int64_t i = 0;
    for (; i + 8 <= n; i += 8)
      for (int v = 0; v < 4; ++v)
        for (int l = 0; l < 2; ++l) acc[v][l] += fabs(x[i + 2 * v + l]);
    for (; i + 4 <= n; i += 4)
      for (int v = 0; v < 2; ++v)
        for (int l = 0; l < 2; ++l) acc[v][l] += fabs(x[i + 2 * v + l]);
    res += (((acc[0][0] + acc[0][1]) + (acc[1][0] + acc[1][1])) + (acc[2][0] + acc[2][1])) + (acc[3][0] + acc[3][1]);
    for (; i < n; ++i) res += fabs(x[i]);
    return res;
  }
  for (int64_t i = 0; i < n; ++i) res += fabs(x[i * incx]);
  return res;
}
/* dzasum.rs:40-111: over the 2n doubles, 8-double chunks into four 2-lane   */
/* accumulators, then 2-double steps into acc0, tail.                        */
double port_dzasum(int64_t n, const double *x, int64_t incx) {
  if (n == 0 || incx == 0) return 0;
  double res = 0;
  if (incx == 1) {
    const int64_t len = 2 * n;
    double acc[4][2];
    memset(acc, 0, sizeof(acc));
    int64_t i = 0;
    for (; i + 8 <= len; i += 8)
      for (int v = 0; v < 4; ++v)
        for (int l = 0; l < 2; ++l) acc[v][l] += fabs(x[i + 2 * v + l]);
    for (; i + 2 <= len; i += 2)
      for (int l = 0; l < 2; ++l) acc[0][l] += fabs(x[i + l]);
    res += (((acc[0][0] + acc[0][1]) + (acc[1][0] + acc[1][1])) + (acc[2][0] + acc[2][1])) + (acc[3][0] + acc[3][1]);
    for (; i < len; ++i) res += fabs(x[i]);
    return res;
  }
  for (int64_t i = 0; i < n; ++i) res += fabs(x[2 * i * incx]) + fabs(x[2 * i * incx + 1]);
  return res;
}

/* scopy.rs:28-60 .. zcopy.rs, sswap.rs:35-120 .. zswap.rs: element moves    */
/* (W = 1 real, 2 complex floats per element).                               */
#define DEFINE_COPY_SWAP(P, T, W)                                                                            \
  void port_##P##copy(int64_t n, const T *x, int64_t incx, T *y, int64_t incy) {                            \
    for (int64_t i = 0; i < n; ++i)                                                                          \
      for (int w = 0; w < W; ++w) y[W * i * incy + w] = x[W * i * incx + w];                                 \
  }                                                                                                          \
  void port_##P##swap(int64_t n, T *x, int64_t incx, T *y, int64_t incy) {                                  \
    for (int64_t i = 0; i < n; ++i)                                                                          \
      for (int w = 0; w < W; ++w) {                                                                          \
        T t = x[W * i * incx + w];                                                                           \
        x[W * i * incx + w] = y[W * i * incy + w];                                                           \
        y[W * i * incy + w] = t;                                                                             \
      }                                                                                                      \
  }
DEFINE_COPY_SWAP(s, float, 1)
DEFINE_COPY_SWAP(d, double, 1)
DEFINE_COPY_SWAP(c, float, 2)
DEFINE_COPY_SWAP(z, double, 2)

/* srot.rs:51-180 / drot.rs:48-178 (real) and csrot.rs:50-166 / zdrot.rs:     */
/* 51-166 (complex vectors, real c and s; the unit-stride path works on the  */
/* 2n floats).  Unit strides: a NEON body over the first floor(len/Q)*Q      */
/* values, Q = 4 srot / csrot, 2 drot / zdrot, with x' = fma(y, s, c*x) and  */
/* (real) y' = fma(-s, x, c*y) or (complex) y' = c*y - s*x; tail and strided */
/* path: x' = c*x + s*y, y' = c*y - s*x, every product rounded.              */
#define DEFINE_ROT(NAME, T, FMA, Q, W, YFUSED)                                                               \
  void port_##NAME(int64_t n, T *x, int64_t incx, T *y, int64_t incy, T c, T s) {                           \
    if (n == 0) return;                                                                                      \
    const int unit = incx == 1 && incy == 1;                                                                 \
    const int64_t len = W * n, nf = unit ? len / Q * Q : 0;                                                  \
    for (int64_t k = 0; k < len; ++k) {                                                                      \
      const int64_t i = k / W, w = k % W;                                                                    \
      T *px = x + W * i * incx + w, *py = y + W * i * incy + w;                                              \
      const T xv = *px, yv = *py;                                                                            \
      if (k < nf) {                                                                                          \
        *px = FMA(yv, s, c * xv);                                                                            \
        *py = YFUSED ? FMA(-s, xv, c * yv) : c * yv - s * xv;                                                \
      } else {                                                                                               \
        *px = c * xv + s * yv;                                                                               \
        *py = c * yv - s * xv;                                                                               \
      }                                                                                                      \
    }                                                                                                        \
  }
DEFINE_ROT(srot, float, fmaf, 4, 1, 1)
DEFINE_ROT(drot, double, fma, 2, 1, 1)
DEFINE_ROT(csrot, float, fmaf, 4, 2, 0)
DEFINE_ROT(zdrot, double, fma, 2, 2, 0)

/* srotm.rs:47-412 / drotm.rs:47-410: flag = param[0]; -2 is the identity;   */
/* flag < 0: H = [h11 h12; h21 h22], flag == 0: [1 h12; h21 1], otherwise:   */
/* [h11 1; -1 h22].  Unit strides: NEON body over floor(n/Q)*Q elements (Q = */
/* 4 s, 2 d) with x' = fma(y, h12, h11*x), y' = fma(x, h21, h22*y) (flag<0), */
/* x' = fma(y, h12, x), y' = fma(x, h21, y) (flag 0), x' = fma(x, h11, y),    */
/* y' = h22*y - x (otherwise); tail and strided: every product rounded.      */
#define DEFINE_ROTM(NAME, T, FMA, Q)                                                                         \
  void port_##NAME(int64_t n, T *x, int64_t incx, T *y, int64_t incy, const T *param) {                     \
    if (n == 0) return;                                                                                      \
    const T flag = param[0], h11 = param[1], h21 = param[2], h12 = param[3], h22 = param[4];                 \
    if (flag == (T)-2) return;                                                                               \
    const int64_t nf = (incx == 1 && incy == 1) ? n / Q * Q : 0;                                             \
    for (int64_t i = 0; i < n; ++i) {                                                                        \
      T *px = x + i * incx, *py = y + i * incy;                                                              \
      const T xv = *px, yv = *py;                                                                            \
      T xn, yn;                                                                                              \
      if (flag < 0) {                                                                                        \
        xn = i < nf ? FMA(yv, h12, h11 * xv) : h11 * xv + h12 * yv;                                          \
        yn = i < nf ? FMA(xv, h21, h22 * yv) : h21 * xv + h22 * yv;                                          \
      } else if (flag == 0) {                                                                                \
        xn = i < nf ? FMA(yv, h12, xv) : xv + h12 * yv;                                                      \
        yn = i < nf ? FMA(xv, h21, yv) : h21 * xv + yv;                                                      \
      } else {                                                                                               \
        xn = i < nf ? FMA(xv, h11, yv) : h11 * xv + yv;                                                      \
        yn = i < nf ? h22 * yv - xv : -xv + h22 * yv;                                                        \
      }                                                                                                      \
      *py = yn;                                                                                              \
      *px = xn;                                                                                              \
    }                                                                                                        \
  }
DEFINE_ROTM(srotm, float, fmaf, 4)
DEFINE_ROTM(drotm, double, fma, 2)

/* srotg.rs:40-75 / drotg.rs:40-71: r = scale*sqrt((a/scale)^2 + (b/scale)^2) */
/* with scale = |a| + |b| and the sign of the larger of a, b; c = a/r,       */
/* s = b/r; z = s if |a| > |b|, 1/c if |b| >= |a| and c != 0, else 1;        */
/* a := r, b := z.  scale == 0 gives c = 1, s = 0, a = b = 0.                */
#define DEFINE_ROTG(NAME, T, ABS, SQRT)                                                                      \
  void port_##NAME(T *a, T *b, T *c, T *s) {                                                                 \
    const T roe = ABS(*a) > ABS(*b) ? *a : *b;                                                               \
    const T scale = ABS(*a) + ABS(*b);                                                                       \
    if (scale == 0) {                                                                                        \
      *c = 1;                                                                                                \
      *s = 0;                                                                                                \
      *a = 0;                                                                                                \
      *b = 0;                                                                                                \
      return;                                                                                                \
    }                                                                                                        \
    const T as = *a / scale, bs = *b / scale;                                                                \
    T r = scale * SQRT(as * as + bs * bs);                                                                   \
    if (roe < 0) r = -r;                                                                                     \
    *c = *a / r;                                                                                             \
    *s = *b / r;                                                                                             \
    T z = 1;                                                                                                 \
    if (ABS(*a) > ABS(*b)) z = *s;                                                                           \
    if (ABS(*b) >= ABS(*a) && *c != 0) z = 1 / *c;                                                           \
    *a = r;                                                                                                  \
    *b = z;                                                                                                  \
  }
DEFINE_ROTG(srotg, float, fabsf, sqrtf)
DEFINE_ROTG(drotg, double, fabs, sqrt)

/* srotmg.rs:42-199 / drotmg.rs:41-195: the modified Givens construction,    */
/* GAM = 4096 rescaling of d1 and d2 into [1/GAM^2, GAM^2].  The rescaling   */
/* loops run only on finite d1/d2 (CORAL's loops would not end on an         */
/* infinite one).                                                            */
#define DEFINE_ROTMG(NAME, T, ABS, ISFIN)                                                                    \
  void port_##NAME(T *d1, T *d2, T *x1, T y1, T *param) {                                                    \
    const T gam = 4096, gamsq = gam * gam, rgamsq = 1 / gamsq;                                               \
    T flag, h11 = 0, h12 = 0, h21 = 0, h22 = 0;                                                              \
    if (*d1 < 0) {                                                                                           \
      flag = -1;                                                                                             \
      *d1 = 0;                                                                                               \
      *d2 = 0;                                                                                               \
      *x1 = 0;                                                                                               \
    } else {                                                                                                 \
      const T p2 = *d2 * y1;                                                                                 \
      if (p2 == 0) {                                                                                         \
        param[0] = -2;                                                                                       \
        return;                                                                                              \
      }                                                                                                      \
      const T p1 = *d1 * *x1, q2 = p2 * y1, q1 = p1 * *x1;                                                   \
      if (ABS(q1) > ABS(q2)) {                                                                               \
        h21 = -y1 / *x1;                                                                                     \
        h12 = p2 / p1;                                                                                       \
        const T u = 1 - h12 * h21;                                                                           \
        if (u <= 0) {                                                                                        \
          flag = -1;                                                                                         \
          *d1 = 0;                                                                                           \
          *d2 = 0;                                                                                           \
          *x1 = 0;                                                                                           \
          h11 = h12 = h21 = h22 = 0;                                                                         \
        } else {                                                                                             \
          flag = 0;                                                                                          \
          *d1 /= u;                                                                                          \
          *d2 /= u;                                                                                          \
          *x1 *= u;                                                                                          \
        }                                                                                                    \
      } else if (q2 < 0) {                                                                                   \
        flag = -1;                                                                                           \
        *d1 = 0;                                                                                             \
        *d2 = 0;                                                                                             \
        *x1 = 0;                                                                                             \
        h11 = h12 = h21 = h22 = 0;                                                                           \
      } else {                                                                                               \
        flag = 1;                                                                                            \
        h11 = p1 / p2;                                                                                       \
        h22 = *x1 / y1;                                                                                      \
        const T u = 1 + h11 * h22;                                                                           \
        const T t = *d2 / u;                                                                                 \
        *d2 = *d1 / u;                                                                                       \
        *d1 = t;                                                                                             \
        *x1 = y1 * u;                                                                                        \
      }                                                                                                      \
      if (*d1 != 0 && ISFIN(*d1)) {                                                                          \
        while (*d1 <= rgamsq || *d1 >= gamsq) {                                                              \
          if (flag == 0) {                                                                                   \
            h11 = 1;                                                                                         \
            h22 = 1;                                                                                         \
          } else {                                                                                           \
            h21 = -1;                                                                                        \
            h12 = 1;                                                                                         \
          }                                                                                                  \
          flag = -1;                                                                                         \
          if (*d1 <= rgamsq) {                                                                               \
            *d1 *= gamsq;                                                                                    \
            *x1 /= gam;                                                                                      \
            h11 /= gam;                                                                                      \
            h12 /= gam;                                                                                      \
          } else {                                                                                           \
            *d1 /= gamsq;                                                                                    \
            *x1 *= gam;                                                                                      \
            h11 *= gam;                                                                                      \
            h12 *= gam;                                                                                      \
          }                                                                                                  \
        }                                                                                                    \
      }                                                                                                      \
      if (*d2 != 0 && ISFIN(*d2)) {                                                                          \
        while (ABS(*d2) <= rgamsq || ABS(*d2) >= gamsq) {                                                    \
          if (flag == 0) {                                                                                   \
            h11 = 1;                                                                                         \
            h22 = 1;                                                                                         \
          } else {                                                                                           \
            h21 = -1;                                                                                        \
            h12 = 1;                                                                                         \
          }                                                                                                  \
          flag = -1;                                                                                         \
          if (ABS(*d2) <= rgamsq) {                                                                          \
            *d2 *= gamsq;                                                                                    \
            h21 /= gam;                                                                                      \
            h22 /= gam;                                                                                      \
          } else {                                                                                           \
            *d2 /= gamsq;                                                                                    \
            h21 *= gam;                                                                                      \
            h22 *= gam;                                                                                      \
          }                                                                                                  \
        }                                                                                                    \
      }                                                                                                      \
    }                                                                                                        \
    param[1] = param[2] = param[3] = param[4] = 0;                                                           \
    if (flag < 0) {                                                                                          \
      param[1] = h11;                                                                                        \
      param[2] = h21;                                                                                        \
      param[3] = h12;                                                                                        \
      param[4] = h22;                                                                                        \
    } else if (flag == 0) {                                                                                  \
      param[2] = h21;                                                                                        \
      param[3] = h12;                                                                                        \
    } else {                                                                                                 \
      param[1] = h11;                                                                                        \
      param[4] = h22;                                                                                        \
    }                                                                                                        \
    param[0] = flag;                                                                                         \
  }
DEFINE_ROTMG(srotmg, float, fabsf, isfinite)
DEFINE_ROTMG(drotmg, double, fabs, isfinite)

/* sscal.rs:30-66 / dscal.rs:30-66 (x := alpha*x, one rounding),             */
/* csscal.rs:39-88 / zdscal.rs:38-88 (real alpha, each part one rounding),   */
/* cscal.rs:43-111 / zscal.rs:41-109: unit stride, NEON body over floor(n/Q)*Q */
/* elements (Q = 4 c, 2 z) with re' = fma(-xi, ai, xr*ar), im' = fma(xr, ai,  */
/* xi*ar); tail and strided re' = ar*xr - ai*xi, im' = ar*xi + ai*xr.        */
/* Quick return: n == 0 or incx == 0.                                        */
#define DEFINE_RSCAL(NAME, T, W)                                                                             \
  void port_##NAME(int64_t n, T alpha, T *x, int64_t incx) {                                                 \
    if (n == 0 || incx == 0) return;                                                                         \
    for (int64_t i = 0; i < n; ++i)                                                                          \
      for (int w = 0; w < W; ++w) x[W * i * incx + w] = x[W * i * incx + w] * alpha;                        \
  }
DEFINE_RSCAL(sscal, float, 1)
DEFINE_RSCAL(dscal, double, 1)
DEFINE_RSCAL(csscal, float, 2)
DEFINE_RSCAL(zdscal, double, 2)
#define DEFINE_CSCAL(NAME, T, FMA, Q)                                                                        \
  void port_##NAME(int64_t n, const T *alpha, T *x, int64_t incx) {                                          \
    if (n == 0 || incx == 0) return;                                                                         \
    const T ar = alpha[0], ai = alpha[1];                                                                    \
    const int64_t nf = incx == 1 ? n / Q * Q : 0;                                                            \
    for (int64_t i = 0; i < n; ++i) {                                                                        \
      T *p = x + 2 * i * incx;                                                                               \
      const T xr = p[0], xi = p[1];                                                                          \
      if (i < nf) {                                                                                          \
        p[0] = FMA(-xi, ai, xr * ar);                                                                        \
        p[1] = FMA(xr, ai, xi * ar);                                                                         \
      } else {                                                                                               \
        p[0] = ar * xr - ai * xi;                                                                            \
        p[1] = ar * xi + ai * xr;                                                                            \
      }                                                                                                      \
    }                                                                                                        \
  }
DEFINE_CSCAL(cscal, float, fmaf, 4)
DEFINE_CSCAL(zscal, double, fma, 2)
