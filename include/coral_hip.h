/*
 * coral_hip.h -- C ABI of libcoralhip.so, an MI355X (gfx950) implementation of
 * CORAL's BLAS call surface.
 *
 * Every routine below replaces one public Rust function of the reference crate
 * `coral-aarch64` (paths relative to the reference root); the citation names the
 * function it stands in for.  Argument order and meaning are CORAL's:
 * column-major storage, positive strides, leading dimensions in elements
 * (complex routines count ld/inc in complex elements and store interleaved
 * [re, im] pairs), complex scalars as two-element arrays [re, im].
 *
 * Two forms of every routine:
 *   coral_<name>(..., coral_stream_t stream)
 *       Device pointers (A, B, C, x, y and result pointers live in GPU memory),
 *       stream ordered, no implicit synchronisation.  The performance path.
 *   coralh_<name>(...)
 *       Host pointers with CORAL's synchronous semantics: operands are copied to
 *       the GPU, the device routine runs, outputs are copied back before return.
 *       This is the drop-in form a Rust/C caller of CORAL binds to.
 *
 * Return value: CORAL_OK (0); -i when argument i (1-based, CORAL's argument
 * order, stream excluded) is invalid; a positive CORAL_ERR_* on a runtime
 * failure.  CORAL itself has no runtime errors (invalid input is undefined
 * behaviour there, `debug_assert!` in debug builds, e.g.
 * coral-aarch64/src/level2/dgemv_notranspose.rs:78-85); this library checks the
 * same length/ld/stride contracts (coral-aarch64/src/level2/assert_length_helpers.rs:4-33)
 * and reports them instead.
 *
 * Quick returns follow CORAL: GEMM with alpha==0 or k==0 computes C := beta*C
 * and never reads C when beta==0 (coral-aarch64/src/level3/sgemm_nn.rs:29-45);
 * GEMV returns when m==0, n==0 or (alpha==0 && beta==1)
 * (coral-aarch64/src/level2/dgemv_notranspose.rs:75-76).  Deviation: with
 * beta==0 GEMV/SYMV overwrite y (reference BLAS semantics) where CORAL's dscal
 * multiplies NaN/Inf in y by zero (coral-aarch64/src/level1/dscal.rs:30-66).
 */
#ifndef CORAL_HIP_H
#define CORAL_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- enums: coral-aarch64/src/enums.rs:1-18, codes in declaration order ---- */
enum { CORAL_NO_TRANSPOSE = 0, CORAL_TRANSPOSE = 1, CORAL_CONJUGATE_TRANSPOSE = 2 };
enum { CORAL_UPPER_TRIANGULAR = 0, CORAL_LOWER_TRIANGULAR = 1 };
enum { CORAL_UNIT_DIAGONAL = 0, CORAL_NON_UNIT_DIAGONAL = 1 };

enum {
  CORAL_OK = 0,
  CORAL_ERR_HIP = 1,        /* a HIP runtime call or kernel launch failed     */
  CORAL_ERR_NO_DEVICE = 2,  /* no usable gfx950 device                        */
  CORAL_ERR_ALLOC = 3,      /* workspace allocation failed                    */
  CORAL_ERR_UNSUPPORTED = 4,/* argument combination not supported             */
  CORAL_ERR_IO = 5,         /* autotuner table file could not be read/written */
  CORAL_ERR_WAIT_BOUND = 6, /* a device-side wait gave up (coral_stream_check): the x of the
                               triangular solve that set it is invalid                       */
  CORAL_ERR_PROBLEM_ARG = 7 /* a batched call skipped problems whose per-problem dimensions or
                               leading dimensions are invalid (coral_stream_check)           */
};

typedef void *coral_stream_t; /* a hipStream_t; NULL = the library's default stream */

/* ---- library / device utilities (no reference counterpart: device plumbing) ---- */
int coral_init(void);
const char *coral_status_string(int status);
const char *coral_last_error(void);
int coral_device_name(char *buf, size_t len);
int coral_malloc(void **ptr, size_t bytes);
int coral_free(void *ptr);
/* The copies and the fill are blocking: they return once the device memory holds the
 * result, so work queued afterwards on any stream sees it. */
int coral_memcpy_htod(void *dst, const void *src, size_t bytes);
int coral_memcpy_dtoh(void *dst, const void *src, size_t bytes);
int coral_memcpy_dtod(void *dst, const void *src, size_t bytes);
int coral_memset(void *dst, int value, size_t bytes);
int coral_synchronize(void);
int coral_stream_create(coral_stream_t *stream);
int coral_stream_destroy(coral_stream_t stream);
int coral_stream_synchronize(coral_stream_t stream);
/* Synchronises the stream and reports, then clears, the errors its kernels recorded
 * since the last check: CORAL_ERR_WAIT_BOUND (a TRSV handoff wait gave up),
 * CORAL_ERR_PROBLEM_ARG (a batched call skipped problems with invalid per-problem
 * arguments), else CORAL_OK.  The device-pointer routines cannot report these when
 * they return (their kernels have not run yet); the host-pointer forms check them
 * before returning. */
int coral_stream_check(coral_stream_t stream);
/* Test hook: bound on the failed polls of a TRSV handoff wait (0 restores the
 * default, about seconds of waiting; -1 makes every wait give up at once, found or
 * not), so the error path can be exercised while every kernel still exits
 * normally. */
int coral_testing_set_wait_limit(int64_t spins);
int coral_event_create(void **event);
int coral_event_destroy(void *event);
int coral_event_record(void *event, coral_stream_t stream);
int coral_event_elapsed_ms(float *ms, void *start, void *stop);
int coral_event_synchronize(void *event);
/* hipGraph capture of stream work (e.g. a batched call) and replay. */
int coral_graph_begin_capture(coral_stream_t stream);
int coral_graph_end_capture(coral_stream_t stream, void **graph_exec);
int coral_graph_launch(void *graph_exec, coral_stream_t stream);
int coral_graph_destroy(void *graph_exec);
/* Seeded synthetic data on the device (counter-based splitmix64; identical
 * bits to the host generator in coral_amd/synthetic.py for the uniform kinds).
 * kind: 0 = uniform[-1,1), 1 = normal(0,1) (Box-Muller in f64). */
int coral_fill_random_f32(float *dst, int64_t count, uint64_t seed, int32_t kind, coral_stream_t stream);
int coral_fill_random_f64(double *dst, int64_t count, uint64_t seed, int32_t kind, coral_stream_t stream);
/* Kernel timing for roofline reporting (no reference counterpart).  When enabled,
 * the library brackets its dominant kernel launches with HIP events on the launch
 * stream.  kind: 0 = GEMM MFMA kernel, 1 = GEMV streaming kernel.
 * coral_timing_read synchronises on the recorded events and returns the summed
 * duration and launch count since the last reset. */
int coral_timing_enable(int32_t on);
int coral_timing_reset(void);
int coral_timing_read(int32_t kind, double *total_ms, int64_t *launches);
/* Scale row i of column-major M (rows x cols, ld) by 2^row_exp[i] and column j
 * by 2^col_exp[j] (either exponent pointer may be NULL).  Exponents are device int32. */
int coral_scale_pow2_f32(float *m, int64_t rows, int64_t cols, int64_t ld,
                         const int32_t *row_exp, const int32_t *col_exp, coral_stream_t stream);

/* ======================================================================== */
/* Level 3                                                                  */
/* ======================================================================== */

/* coral-aarch64/src/level3/sgemm.rs:106-130 (sgemm): C := alpha*op(A)*op(B) + beta*C, FP32 MFMA. */
int coral_sgemm(int32_t op_a, int32_t op_b, int64_t m, int64_t n, int64_t k,
                float alpha, const float *a, int64_t lda, const float *b, int64_t ldb,
                float beta, float *c, int64_t ldc, coral_stream_t stream);
/* coral-aarch64/src/level3/dgemm.rs:107 (dgemm): FP64 MFMA. */
int coral_dgemm(int32_t op_a, int32_t op_b, int64_t m, int64_t n, int64_t k,
                double alpha, const double *a, int64_t lda, const double *b, int64_t ldb,
                double beta, double *c, int64_t ldc, coral_stream_t stream);
/* coral-aarch64/src/level3/cgemm.rs:127-157 (cgemm, 9 op pairs): 4M on FP32 MFMA. */
int coral_cgemm(int32_t op_a, int32_t op_b, int64_t m, int64_t n, int64_t k,
                const float alpha[2], const float *a, int64_t lda, const float *b, int64_t ldb,
                const float beta[2], float *c, int64_t ldc, coral_stream_t stream);
/* coral-aarch64/src/level3/zgemm.rs:123 (zgemm): 4M on FP64 MFMA. */
int coral_zgemm(int32_t op_a, int32_t op_b, int64_t m, int64_t n, int64_t k,
                const double alpha[2], const double *a, int64_t lda, const double *b, int64_t ldb,
                const double beta[2], double *c, int64_t ldc, coral_stream_t stream);

/* Split-precision forms (no reference counterpart; same argument list as coral_sgemm /
 * coral_cgemm).  Each FP32 operand is scaled by powers of two per row of op(A) and
 * per column of op(B), split into BF16 hi + lo, and three BF16 MFMA products
 * (hi*hi + hi*lo + lo*hi) are accumulated in FP32.  CGEMM uses 3M (Gauss) with
 * each real product in that form. */
int coral_sgemm_bf16x3(int32_t op_a, int32_t op_b, int64_t m, int64_t n, int64_t k,
                       float alpha, const float *a, int64_t lda, const float *b, int64_t ldb,
                       float beta, float *c, int64_t ldc, coral_stream_t stream);
int coral_cgemm_3m_bf16x3(int32_t op_a, int32_t op_b, int64_t m, int64_t n, int64_t k,
                          const float alpha[2], const float *a, int64_t lda, const float *b, int64_t ldb,
                          const float beta[2], float *c, int64_t ldc, coral_stream_t stream);
/* The same split with FP16 halves (11 + 11 significand bits, three FP16 MFMAs, FP32
 * accumulation; rows/columns scaled so their maximum lies in [2^14, 2^15)).  Same
 * MFMA rate as the BF16 form; the 22-bit operand representation brings the error
 * to that of plain FP32 MFMA.  No reference counterpart. */
int coral_sgemm_fp16x3(int32_t op_a, int32_t op_b, int64_t m, int64_t n, int64_t k,
                       float alpha, const float *a, int64_t lda, const float *b, int64_t ldb,
                       float beta, float *c, int64_t ldc, coral_stream_t stream);
int coral_cgemm_3m_fp16x3(int32_t op_a, int32_t op_b, int64_t m, int64_t n, int64_t k,
                          const float alpha[2], const float *a, int64_t lda, const float *b, int64_t ldb,
                          const float beta[2], float *c, int64_t ldc, coral_stream_t stream);

/* Batched uneven GEMM (no reference counterpart; equivalent to a loop of
 * coral-aarch64/src/level3/sgemm.rs:106 / dgemm.rs:107 calls).  All arrays are
 * DEVICE arrays of length `batch` (at most 2^30): pointer arrays a/b/c and
 * per-problem m, n, k, lda, ldb, ldc.  One op pair and one alpha/beta for the whole
 * batch.  Runs as one persistent work-queue kernel; capturable in a hipGraph after
 * one eager call of the same batch size on the capturing stream (which sizes the
 * stream's workspace; a graph keeps the workspace it captured alive, but graphs
 * captured from one stream share it, so replays of two such graphs must not
 * overlap).  A problem with m, n or k < 0, a leading dimension below its rows
 * (max(1, rows)), or lda or ldb above 2^24 is skipped, its C left unchanged, and
 * recorded for coral_stream_check (CORAL_ERR_PROBLEM_ARG). */
int coral_sgemm_batched(int32_t op_a, int32_t op_b, int64_t batch,
                        const int32_t *m, const int32_t *n, const int32_t *k, float alpha,
                        const float *const *a, const int32_t *lda,
                        const float *const *b, const int32_t *ldb, float beta,
                        float *const *c, const int32_t *ldc, coral_stream_t stream);
int coral_dgemm_batched(int32_t op_a, int32_t op_b, int64_t batch,
                        const int32_t *m, const int32_t *n, const int32_t *k, double alpha,
                        const double *const *a, const int32_t *lda,
                        const double *const *b, const int32_t *ldb, double beta,
                        double *const *c, const int32_t *ldc, coral_stream_t stream);

/* Autotuner (no reference counterpart; stands in for CORAL's compile-time
 * MC/NC/KC/MR/NR constants, coral-aarch64/src/level3/sgemm.rs:92-94).
 * Tile choice is keyed by (routine, op pair, size bucket); the table persists
 * as a text file.  coral_autotune_run times every candidate tile for the bucket
 * of (m,n,k) on the current device and stores the fastest. */
int coral_autotune_run(const char *routine, int32_t op_a, int32_t op_b,
                       int64_t m, int64_t n, int64_t k, int32_t *chosen_tile);
int coral_autotune_save(const char *path);
int coral_autotune_load(const char *path);
int coral_autotune_get(const char *routine, int32_t op_a, int32_t op_b,
                       int64_t m, int64_t n, int64_t k, int32_t *tile);
int coral_autotune_set(const char *routine, int32_t op_a, int32_t op_b,
                       int64_t m, int64_t n, int64_t k, int32_t tile);
int coral_autotune_clear(void);

/* ======================================================================== */
/* Level 2                                                                  */
/* ======================================================================== */

/* coral-aarch64/src/level2/sgemv.rs:63 / dgemv.rs:61 / cgemv.rs:67 / zgemv.rs:63 */
int coral_sgemv(int32_t trans, int64_t m, int64_t n, float alpha, const float *a, int64_t lda,
                const float *x, int64_t incx, float beta, float *y, int64_t incy, coral_stream_t stream);
int coral_dgemv(int32_t trans, int64_t m, int64_t n, double alpha, const double *a, int64_t lda,
                const double *x, int64_t incx, double beta, double *y, int64_t incy, coral_stream_t stream);
int coral_cgemv(int32_t trans, int64_t m, int64_t n, const float alpha[2], const float *a, int64_t lda,
                const float *x, int64_t incx, const float beta[2], float *y, int64_t incy, coral_stream_t stream);
int coral_zgemv(int32_t trans, int64_t m, int64_t n, const double alpha[2], const double *a, int64_t lda,
                const double *x, int64_t incx, const double beta[2], double *y, int64_t incy, coral_stream_t stream);

/* Allocation-free DGEMV: no counterpart in the reference.  For callers that cannot
 * allocate device memory (code under hipGraph capture, tight serving loops, hosts
 * that forbid allocation): the same arguments, kernels and results as coral_dgemv,
 * with the partial sums in a fixed 4 MiB device array of the library (one per GPU,
 * held until the process ends) instead of a workspace; it calls no hipMalloc.  y
 * may hold at most 2^19 elements (m for NoTranspose, n otherwise); larger calls
 * return -2 (m) or -3 (n).  Every call on one GPU shares the array: calls that
 * overlap on one GPU (on different streams) race. */
int coral_dgemv_noalloc(int32_t trans, int64_t m, int64_t n, double alpha, const double *a, int64_t lda,
                        const double *x, int64_t incx, double beta, double *y, int64_t incy, coral_stream_t stream);

/* coral-aarch64/src/level2/ssymv.rs:99 / dsymv.rs:99 */
int coral_ssymv(int32_t uplo, int64_t n, float alpha, const float *a, int64_t lda,
                const float *x, int64_t incx, float beta, float *y, int64_t incy, coral_stream_t stream);
int coral_dsymv(int32_t uplo, int64_t n, double alpha, const double *a, int64_t lda,
                const double *x, int64_t incx, double beta, double *y, int64_t incy, coral_stream_t stream);
/* coral-aarch64/src/level2/chemv.rs:96 / zhemv.rs:97 */
int coral_chemv(int32_t uplo, int64_t n, const float alpha[2], const float *a, int64_t lda,
                const float *x, int64_t incx, const float beta[2], float *y, int64_t incy, coral_stream_t stream);
int coral_zhemv(int32_t uplo, int64_t n, const double alpha[2], const double *a, int64_t lda,
                const double *x, int64_t incx, const double beta[2], double *y, int64_t incy, coral_stream_t stream);

/* coral-aarch64/src/level2/strsv.rs:60 / dtrsv.rs:62 / ctrsv.rs:59 / ztrsv.rs:60 */
int coral_strsv(int32_t uplo, int32_t trans, int32_t diag, int64_t n, const float *a, int64_t lda,
                float *x, int64_t incx, coral_stream_t stream);
int coral_dtrsv(int32_t uplo, int32_t trans, int32_t diag, int64_t n, const double *a, int64_t lda,
                double *x, int64_t incx, coral_stream_t stream);
int coral_ctrsv(int32_t uplo, int32_t trans, int32_t diag, int64_t n, const float *a, int64_t lda,
                float *x, int64_t incx, coral_stream_t stream);
int coral_ztrsv(int32_t uplo, int32_t trans, int32_t diag, int64_t n, const double *a, int64_t lda,
                double *x, int64_t incx, coral_stream_t stream);

/* coral-aarch64/src/level2/strmv.rs:59 / dtrmv.rs:58 / ctrmv.rs:58 / ztrmv.rs:58 */
int coral_strmv(int32_t uplo, int32_t trans, int32_t diag, int64_t n, const float *a, int64_t lda,
                float *x, int64_t incx, coral_stream_t stream);
int coral_dtrmv(int32_t uplo, int32_t trans, int32_t diag, int64_t n, const double *a, int64_t lda,
                double *x, int64_t incx, coral_stream_t stream);
int coral_ctrmv(int32_t uplo, int32_t trans, int32_t diag, int64_t n, const float *a, int64_t lda,
                float *x, int64_t incx, coral_stream_t stream);
int coral_ztrmv(int32_t uplo, int32_t trans, int32_t diag, int64_t n, const double *a, int64_t lda,
                double *x, int64_t incx, coral_stream_t stream);

/* Batched DTRSV (no reference counterpart; a loop of dtrsv.rs:62 calls).  Device
 * arrays of length batch (at most 2^30): n, lda, matrix pointers, x pointers,
 * incx = 1.  Any n is solved (n <= 4096 with the solution staged in LDS, larger n
 * in place in x).  A problem with n < 0 or lda < n is skipped, x left unchanged,
 * and recorded for coral_stream_check (CORAL_ERR_PROBLEM_ARG). */
int coral_dtrsv_batched(int32_t uplo, int32_t trans, int32_t diag, int64_t batch,
                        const int32_t *n, const double *const *a, const int32_t *lda,
                        double *const *x, coral_stream_t stream);

/* Rank updates.  Results are bit-identical to CORAL's: the same per-column
 * coefficient and skip rules, and the same rounding per element (fused where
 * CORAL's NEON axpy body fuses, two roundings in its scalar tails and strided
 * loops).  HER/HER2 set Im(a_jj) to zero as CORAL does.
 * coral-aarch64/src/level2/sger.rs:68 / dger.rs:69: A += alpha * x * y^T */
int coral_sger(int64_t m, int64_t n, float alpha, const float *x, int64_t incx, const float *y, int64_t incy,
               float *a, int64_t lda, coral_stream_t stream);
int coral_dger(int64_t m, int64_t n, double alpha, const double *x, int64_t incx, const double *y, int64_t incy,
               double *a, int64_t lda, coral_stream_t stream);
/* coral-aarch64/src/level2/cgeru.rs:63 / cgerc.rs:62 / zgeru.rs:61 / zgerc.rs:60:
 * A += alpha * x * y^T (u) or alpha * x * y^H (c) */
int coral_cgeru(int64_t m, int64_t n, const float alpha[2], const float *x, int64_t incx, const float *y,
                int64_t incy, float *a, int64_t lda, coral_stream_t stream);
int coral_cgerc(int64_t m, int64_t n, const float alpha[2], const float *x, int64_t incx, const float *y,
                int64_t incy, float *a, int64_t lda, coral_stream_t stream);
int coral_zgeru(int64_t m, int64_t n, const double alpha[2], const double *x, int64_t incx, const double *y,
                int64_t incy, double *a, int64_t lda, coral_stream_t stream);
int coral_zgerc(int64_t m, int64_t n, const double alpha[2], const double *x, int64_t incx, const double *y,
                int64_t incy, double *a, int64_t lda, coral_stream_t stream);
/* coral-aarch64/src/level2/ssyr.rs:63 / dsyr.rs:61: A += alpha * x * x^T (stored triangle) */
int coral_ssyr(int32_t uplo, int64_t n, float alpha, const float *x, int64_t incx, float *a, int64_t lda,
               coral_stream_t stream);
int coral_dsyr(int32_t uplo, int64_t n, double alpha, const double *x, int64_t incx, double *a, int64_t lda,
               coral_stream_t stream);
/* coral-aarch64/src/level2/ssyr2.rs:59 / dsyr2.rs:63: A += alpha * (x * y^T + y * x^T) */
int coral_ssyr2(int32_t uplo, int64_t n, float alpha, const float *x, int64_t incx, const float *y, int64_t incy,
                float *a, int64_t lda, coral_stream_t stream);
int coral_dsyr2(int32_t uplo, int64_t n, double alpha, const double *x, int64_t incx, const double *y,
                int64_t incy, double *a, int64_t lda, coral_stream_t stream);
/* coral-aarch64/src/level2/cher.rs:62 / zher.rs:62: A += alpha * x * x^H, alpha real */
int coral_cher(int32_t uplo, int64_t n, float alpha, const float *x, int64_t incx, float *a, int64_t lda,
               coral_stream_t stream);
int coral_zher(int32_t uplo, int64_t n, double alpha, const double *x, int64_t incx, double *a, int64_t lda,
               coral_stream_t stream);
/* coral-aarch64/src/level2/cher2.rs:65 / zher2.rs:63: A += alpha * x * y^H + conj(alpha) * y * x^H */
int coral_cher2(int32_t uplo, int64_t n, const float alpha[2], const float *x, int64_t incx, const float *y,
                int64_t incy, float *a, int64_t lda, coral_stream_t stream);
int coral_zher2(int32_t uplo, int64_t n, const double alpha[2], const double *x, int64_t incx, const double *y,
                int64_t incy, double *a, int64_t lda, coral_stream_t stream);

/* ======================================================================== */
/* Level 1 (results of reductions are written to DEVICE memory)             */
/* ======================================================================== */

/* coral-aarch64/src/level1/sdot.rs:42 / ddot.rs:43 / cdotu.rs:46 / cdotc.rs:46 / zdotu.rs / zdotc.rs:45 */
int coral_sdot(int64_t n, const float *x, int64_t incx, const float *y, int64_t incy, float *result, coral_stream_t stream);
int coral_ddot(int64_t n, const double *x, int64_t incx, const double *y, int64_t incy, double *result, coral_stream_t stream);
int coral_cdotu(int64_t n, const float *x, int64_t incx, const float *y, int64_t incy, float *result, coral_stream_t stream);
int coral_cdotc(int64_t n, const float *x, int64_t incx, const float *y, int64_t incy, float *result, coral_stream_t stream);
int coral_zdotu(int64_t n, const double *x, int64_t incx, const double *y, int64_t incy, double *result, coral_stream_t stream);
int coral_zdotc(int64_t n, const double *x, int64_t incx, const double *y, int64_t incy, double *result, coral_stream_t stream);

/* coral-aarch64/src/level1/saxpy.rs:42 / daxpy.rs:42 / caxpy.rs:45 / zaxpy.rs:44 */
int coral_saxpy(int64_t n, float alpha, const float *x, int64_t incx, float *y, int64_t incy, coral_stream_t stream);
int coral_daxpy(int64_t n, double alpha, const double *x, int64_t incx, double *y, int64_t incy, coral_stream_t stream);
int coral_caxpy(int64_t n, const float alpha[2], const float *x, int64_t incx, float *y, int64_t incy, coral_stream_t stream);
int coral_zaxpy(int64_t n, const double alpha[2], const double *x, int64_t incx, double *y, int64_t incy, coral_stream_t stream);

/* coral-aarch64/src/level1/snrm2.rs:45 / dnrm2.rs:45 / scnrm2.rs:44 / dznrm2.rs:43 */
int coral_snrm2(int64_t n, const float *x, int64_t incx, float *result, coral_stream_t stream);
int coral_dnrm2(int64_t n, const double *x, int64_t incx, double *result, coral_stream_t stream);
int coral_scnrm2(int64_t n, const float *x, int64_t incx, float *result, coral_stream_t stream);
int coral_dznrm2(int64_t n, const double *x, int64_t incx, double *result, coral_stream_t stream);

/* coral-aarch64/src/level1/isamax.rs:43 / idamax.rs:43 / icamax.rs:48 / izamax.rs:33
 * 0-based index of the first max |x| (complex: |re|+|im|), NaN ignored, n==0 -> 0. */
int coral_isamax(int64_t n, const float *x, int64_t incx, int64_t *result, coral_stream_t stream);
int coral_idamax(int64_t n, const double *x, int64_t incx, int64_t *result, coral_stream_t stream);
int coral_icamax(int64_t n, const float *x, int64_t incx, int64_t *result, coral_stream_t stream);
int coral_izamax(int64_t n, const double *x, int64_t incx, int64_t *result, coral_stream_t stream);

/* coral-aarch64/src/level1/sscal.rs / dscal.rs:30 / cscal.rs:43 / zscal.rs (CORAL uses these
 * internally for the beta pre-scaling of y). */
int coral_sscal(int64_t n, float alpha, float *x, int64_t incx, coral_stream_t stream);
int coral_dscal(int64_t n, double alpha, double *x, int64_t incx, coral_stream_t stream);
int coral_cscal(int64_t n, const float alpha[2], float *x, int64_t incx, coral_stream_t stream);
int coral_zscal(int64_t n, const double alpha[2], double *x, int64_t incx, coral_stream_t stream);
/* coral-aarch64/src/level1/csscal.rs:39 / zdscal.rs:38: complex x times a real alpha.
 * All six scal routines round as CORAL does (complex scal: fused in the NEON body of
 * a unit-stride call, every product rounded elsewhere), so results are bit-identical. */
int coral_csscal(int64_t n, float alpha, float *x, int64_t incx, coral_stream_t stream);
int coral_zdscal(int64_t n, double alpha, double *x, int64_t incx, coral_stream_t stream);

/* coral-aarch64/src/level1/sasum.rs:40 / dasum.rs:41 / scasum.rs:40 / dzasum.rs:40:
 * sum of |x_i| (complex |re| + |im|); n == 0 -> 0.  Result in device memory. */
int coral_sasum(int64_t n, const float *x, int64_t incx, float *result, coral_stream_t stream);
int coral_dasum(int64_t n, const double *x, int64_t incx, double *result, coral_stream_t stream);
int coral_scasum(int64_t n, const float *x, int64_t incx, float *result, coral_stream_t stream);
int coral_dzasum(int64_t n, const double *x, int64_t incx, double *result, coral_stream_t stream);

/* coral-aarch64/src/level1/scopy.rs:28 / dcopy.rs:28 / ccopy.rs:29 / zcopy.rs:28: y := x */
int coral_scopy(int64_t n, const float *x, int64_t incx, float *y, int64_t incy, coral_stream_t stream);
int coral_dcopy(int64_t n, const double *x, int64_t incx, double *y, int64_t incy, coral_stream_t stream);
int coral_ccopy(int64_t n, const float *x, int64_t incx, float *y, int64_t incy, coral_stream_t stream);
int coral_zcopy(int64_t n, const double *x, int64_t incx, double *y, int64_t incy, coral_stream_t stream);
/* coral-aarch64/src/level1/sswap.rs:35 / dswap.rs:35 / cswap.rs:36 / zswap.rs:35: x <-> y */
int coral_sswap(int64_t n, float *x, int64_t incx, float *y, int64_t incy, coral_stream_t stream);
int coral_dswap(int64_t n, double *x, int64_t incx, double *y, int64_t incy, coral_stream_t stream);
int coral_cswap(int64_t n, float *x, int64_t incx, float *y, int64_t incy, coral_stream_t stream);
int coral_zswap(int64_t n, double *x, int64_t incx, double *y, int64_t incy, coral_stream_t stream);

/* coral-aarch64/src/level1/srot.rs:51 / drot.rs:48 / csrot.rs:50 / zdrot.rs:51:
 * x' = c*x + s*y, y' = c*y - s*x (complex vectors, real c and s); bit-identical to CORAL. */
int coral_srot(int64_t n, float *x, int64_t incx, float *y, int64_t incy, float c, float s, coral_stream_t stream);
int coral_drot(int64_t n, double *x, int64_t incx, double *y, int64_t incy, double c, double s,
               coral_stream_t stream);
int coral_csrot(int64_t n, float *x, int64_t incx, float *y, int64_t incy, float c, float s, coral_stream_t stream);
int coral_zdrot(int64_t n, double *x, int64_t incx, double *y, int64_t incy, double c, double s,
                coral_stream_t stream);
/* coral-aarch64/src/level1/srotm.rs:47 / drotm.rs:47: modified rotation; param is a HOST
 * array [flag, h11, h21, h12, h22] (the flag selects the kernel); bit-identical to CORAL. */
int coral_srotm(int64_t n, float *x, int64_t incx, float *y, int64_t incy, const float param[5],
                coral_stream_t stream);
int coral_drotm(int64_t n, double *x, int64_t incx, double *y, int64_t incy, const double param[5],
                coral_stream_t stream);
/* coral-aarch64/src/level1/srotg.rs:40 / drotg.rs:40: Givens construction on DEVICE
 * scalars: a := r, b := z, c, s written. */
int coral_srotg(float *a, float *b, float *c, float *s, coral_stream_t stream);
int coral_drotg(double *a, double *b, double *c, double *s, coral_stream_t stream);
/* coral-aarch64/src/level1/srotmg.rs:42 / drotmg.rs:41: modified Givens construction on
 * DEVICE scalars d1, d2, x1 (in/out) and param[5] (out); y1 by value.  The GAM rescaling
 * runs only on finite d1/d2 (CORAL's loop does not end on an infinite one). */
int coral_srotmg(float *d1, float *d2, float *x1, float y1, float *param, coral_stream_t stream);
int coral_drotmg(double *d1, double *d2, double *x1, double y1, double *param, coral_stream_t stream);

/* ======================================================================== */
/* Host-pointer drop-in forms: same arguments, host memory, synchronous.    */
/* Reductions return their value through a host pointer.                    */
/* ======================================================================== */
int coralh_sgemm(int32_t op_a, int32_t op_b, int64_t m, int64_t n, int64_t k, float alpha,
                 const float *a, int64_t lda, const float *b, int64_t ldb, float beta, float *c, int64_t ldc);
int coralh_dgemm(int32_t op_a, int32_t op_b, int64_t m, int64_t n, int64_t k, double alpha,
                 const double *a, int64_t lda, const double *b, int64_t ldb, double beta, double *c, int64_t ldc);
int coralh_cgemm(int32_t op_a, int32_t op_b, int64_t m, int64_t n, int64_t k, const float alpha[2],
                 const float *a, int64_t lda, const float *b, int64_t ldb, const float beta[2], float *c, int64_t ldc);
int coralh_zgemm(int32_t op_a, int32_t op_b, int64_t m, int64_t n, int64_t k, const double alpha[2],
                 const double *a, int64_t lda, const double *b, int64_t ldb, const double beta[2], double *c, int64_t ldc);
int coralh_sgemm_bf16x3(int32_t op_a, int32_t op_b, int64_t m, int64_t n, int64_t k, float alpha,
                        const float *a, int64_t lda, const float *b, int64_t ldb, float beta, float *c, int64_t ldc);
int coralh_cgemm_3m_bf16x3(int32_t op_a, int32_t op_b, int64_t m, int64_t n, int64_t k, const float alpha[2],
                           const float *a, int64_t lda, const float *b, int64_t ldb, const float beta[2], float *c, int64_t ldc);
int coralh_sgemm_fp16x3(int32_t op_a, int32_t op_b, int64_t m, int64_t n, int64_t k, float alpha,
                        const float *a, int64_t lda, const float *b, int64_t ldb, float beta, float *c, int64_t ldc);
int coralh_cgemm_3m_fp16x3(int32_t op_a, int32_t op_b, int64_t m, int64_t n, int64_t k, const float alpha[2],
                           const float *a, int64_t lda, const float *b, int64_t ldb, const float beta[2], float *c, int64_t ldc);
int coralh_sgemv(int32_t trans, int64_t m, int64_t n, float alpha, const float *a, int64_t lda,
                 const float *x, int64_t incx, float beta, float *y, int64_t incy);
int coralh_dgemv(int32_t trans, int64_t m, int64_t n, double alpha, const double *a, int64_t lda,
                 const double *x, int64_t incx, double beta, double *y, int64_t incy);
int coralh_cgemv(int32_t trans, int64_t m, int64_t n, const float alpha[2], const float *a, int64_t lda,
                 const float *x, int64_t incx, const float beta[2], float *y, int64_t incy);
int coralh_zgemv(int32_t trans, int64_t m, int64_t n, const double alpha[2], const double *a, int64_t lda,
                 const double *x, int64_t incx, const double beta[2], double *y, int64_t incy);
int coralh_ssymv(int32_t uplo, int64_t n, float alpha, const float *a, int64_t lda,
                 const float *x, int64_t incx, float beta, float *y, int64_t incy);
int coralh_dsymv(int32_t uplo, int64_t n, double alpha, const double *a, int64_t lda,
                 const double *x, int64_t incx, double beta, double *y, int64_t incy);
int coralh_chemv(int32_t uplo, int64_t n, const float alpha[2], const float *a, int64_t lda,
                 const float *x, int64_t incx, const float beta[2], float *y, int64_t incy);
int coralh_zhemv(int32_t uplo, int64_t n, const double alpha[2], const double *a, int64_t lda,
                 const double *x, int64_t incx, const double beta[2], double *y, int64_t incy);
int coralh_strsv(int32_t uplo, int32_t trans, int32_t diag, int64_t n, const float *a, int64_t lda, float *x, int64_t incx);
int coralh_dtrsv(int32_t uplo, int32_t trans, int32_t diag, int64_t n, const double *a, int64_t lda, double *x, int64_t incx);
int coralh_ctrsv(int32_t uplo, int32_t trans, int32_t diag, int64_t n, const float *a, int64_t lda, float *x, int64_t incx);
int coralh_ztrsv(int32_t uplo, int32_t trans, int32_t diag, int64_t n, const double *a, int64_t lda, double *x, int64_t incx);
int coralh_strmv(int32_t uplo, int32_t trans, int32_t diag, int64_t n, const float *a, int64_t lda, float *x, int64_t incx);
int coralh_dtrmv(int32_t uplo, int32_t trans, int32_t diag, int64_t n, const double *a, int64_t lda, double *x, int64_t incx);
int coralh_ctrmv(int32_t uplo, int32_t trans, int32_t diag, int64_t n, const float *a, int64_t lda, float *x, int64_t incx);
int coralh_ztrmv(int32_t uplo, int32_t trans, int32_t diag, int64_t n, const double *a, int64_t lda, double *x, int64_t incx);
int coralh_sger(int64_t m, int64_t n, float alpha, const float *x, int64_t incx, const float *y, int64_t incy,
                float *a, int64_t lda);
int coralh_dger(int64_t m, int64_t n, double alpha, const double *x, int64_t incx, const double *y, int64_t incy,
                double *a, int64_t lda);
int coralh_cgeru(int64_t m, int64_t n, const float alpha[2], const float *x, int64_t incx, const float *y,
                 int64_t incy, float *a, int64_t lda);
int coralh_cgerc(int64_t m, int64_t n, const float alpha[2], const float *x, int64_t incx, const float *y,
                 int64_t incy, float *a, int64_t lda);
int coralh_zgeru(int64_t m, int64_t n, const double alpha[2], const double *x, int64_t incx, const double *y,
                 int64_t incy, double *a, int64_t lda);
int coralh_zgerc(int64_t m, int64_t n, const double alpha[2], const double *x, int64_t incx, const double *y,
                 int64_t incy, double *a, int64_t lda);
int coralh_ssyr(int32_t uplo, int64_t n, float alpha, const float *x, int64_t incx, float *a, int64_t lda);
int coralh_dsyr(int32_t uplo, int64_t n, double alpha, const double *x, int64_t incx, double *a, int64_t lda);
int coralh_ssyr2(int32_t uplo, int64_t n, float alpha, const float *x, int64_t incx, const float *y, int64_t incy,
                 float *a, int64_t lda);
int coralh_dsyr2(int32_t uplo, int64_t n, double alpha, const double *x, int64_t incx, const double *y,
                 int64_t incy, double *a, int64_t lda);
int coralh_cher(int32_t uplo, int64_t n, float alpha, const float *x, int64_t incx, float *a, int64_t lda);
int coralh_zher(int32_t uplo, int64_t n, double alpha, const double *x, int64_t incx, double *a, int64_t lda);
int coralh_cher2(int32_t uplo, int64_t n, const float alpha[2], const float *x, int64_t incx, const float *y,
                 int64_t incy, float *a, int64_t lda);
int coralh_zher2(int32_t uplo, int64_t n, const double alpha[2], const double *x, int64_t incx, const double *y,
                 int64_t incy, double *a, int64_t lda);
int coralh_sdot(int64_t n, const float *x, int64_t incx, const float *y, int64_t incy, float *result);
int coralh_ddot(int64_t n, const double *x, int64_t incx, const double *y, int64_t incy, double *result);
int coralh_cdotu(int64_t n, const float *x, int64_t incx, const float *y, int64_t incy, float *result);
int coralh_cdotc(int64_t n, const float *x, int64_t incx, const float *y, int64_t incy, float *result);
int coralh_zdotu(int64_t n, const double *x, int64_t incx, const double *y, int64_t incy, double *result);
int coralh_zdotc(int64_t n, const double *x, int64_t incx, const double *y, int64_t incy, double *result);
int coralh_saxpy(int64_t n, float alpha, const float *x, int64_t incx, float *y, int64_t incy);
int coralh_daxpy(int64_t n, double alpha, const double *x, int64_t incx, double *y, int64_t incy);
int coralh_caxpy(int64_t n, const float alpha[2], const float *x, int64_t incx, float *y, int64_t incy);
int coralh_zaxpy(int64_t n, const double alpha[2], const double *x, int64_t incx, double *y, int64_t incy);
int coralh_snrm2(int64_t n, const float *x, int64_t incx, float *result);
int coralh_dnrm2(int64_t n, const double *x, int64_t incx, double *result);
int coralh_scnrm2(int64_t n, const float *x, int64_t incx, float *result);
int coralh_dznrm2(int64_t n, const double *x, int64_t incx, double *result);
int coralh_isamax(int64_t n, const float *x, int64_t incx, int64_t *result);
int coralh_idamax(int64_t n, const double *x, int64_t incx, int64_t *result);
int coralh_icamax(int64_t n, const float *x, int64_t incx, int64_t *result);
int coralh_izamax(int64_t n, const double *x, int64_t incx, int64_t *result);
int coralh_sscal(int64_t n, float alpha, float *x, int64_t incx);
int coralh_dscal(int64_t n, double alpha, double *x, int64_t incx);
int coralh_cscal(int64_t n, const float alpha[2], float *x, int64_t incx);
int coralh_zscal(int64_t n, const double alpha[2], double *x, int64_t incx);
int coralh_csscal(int64_t n, float alpha, float *x, int64_t incx);
int coralh_zdscal(int64_t n, double alpha, double *x, int64_t incx);
int coralh_sasum(int64_t n, const float *x, int64_t incx, float *result);
int coralh_dasum(int64_t n, const double *x, int64_t incx, double *result);
int coralh_scasum(int64_t n, const float *x, int64_t incx, float *result);
int coralh_dzasum(int64_t n, const double *x, int64_t incx, double *result);
int coralh_scopy(int64_t n, const float *x, int64_t incx, float *y, int64_t incy);
int coralh_dcopy(int64_t n, const double *x, int64_t incx, double *y, int64_t incy);
int coralh_ccopy(int64_t n, const float *x, int64_t incx, float *y, int64_t incy);
int coralh_zcopy(int64_t n, const double *x, int64_t incx, double *y, int64_t incy);
int coralh_sswap(int64_t n, float *x, int64_t incx, float *y, int64_t incy);
int coralh_dswap(int64_t n, double *x, int64_t incx, double *y, int64_t incy);
int coralh_cswap(int64_t n, float *x, int64_t incx, float *y, int64_t incy);
int coralh_zswap(int64_t n, double *x, int64_t incx, double *y, int64_t incy);
int coralh_srot(int64_t n, float *x, int64_t incx, float *y, int64_t incy, float c, float s);
int coralh_drot(int64_t n, double *x, int64_t incx, double *y, int64_t incy, double c, double s);
int coralh_csrot(int64_t n, float *x, int64_t incx, float *y, int64_t incy, float c, float s);
int coralh_zdrot(int64_t n, double *x, int64_t incx, double *y, int64_t incy, double c, double s);
int coralh_srotm(int64_t n, float *x, int64_t incx, float *y, int64_t incy, const float param[5]);
int coralh_drotm(int64_t n, double *x, int64_t incx, double *y, int64_t incy, const double param[5]);
int coralh_srotg(float *a, float *b, float *c, float *s);
int coralh_drotg(double *a, double *b, double *c, double *s);
int coralh_srotmg(float *d1, float *d2, float *x1, float y1, float param[5]);
int coralh_drotmg(double *d1, double *d2, double *x1, double y1, double param[5]);

#ifdef __cplusplus
}
#endif

#endif /* CORAL_HIP_H */
