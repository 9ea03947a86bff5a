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
#define DEFINE_AXPY(NAME, T, FMA)                                                                            \
  void port_##NAME(int64_t n, T alpha, const T *x, int64_t incx, T *y, int64_t incy) {                      \
    if (n == 0 || alpha == (T)0) return;                                                                     \
    for (int64_t i = 0; i < n; ++i) y[i * incy] = FMA(alpha, x[i * incx], y[i * incy]);                      \
  }
DEFINE_AXPY(saxpy, float, fmaf)
DEFINE_AXPY(daxpy, double, fma)

#define DEFINE_CAXPY(NAME, T)                                                                                \
  void port_##NAME(int64_t n, const T *alpha, const T *x, int64_t incx, T *y, int64_t incy) {               \
    if (n == 0 || (alpha[0] == 0 && alpha[1] == 0)) return;                                                  \
    for (int64_t i = 0; i < n; ++i) {                                                                        \
      T xr = x[2 * i * incx], xi = x[2 * i * incx + 1];                                                      \
      y[2 * i * incy] += alpha[0] * xr - alpha[1] * xi;                                                      \
      y[2 * i * incy + 1] += alpha[0] * xi + alpha[1] * xr;                                                  \
    }                                                                                                        \
  }
DEFINE_CAXPY(caxpy, float)
DEFINE_CAXPY(zaxpy, double)

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
