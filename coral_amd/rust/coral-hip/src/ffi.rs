//! Raw bindings of libcoralhip.so's host-pointer entry points (include/coral_hip.h).
//!
//! Every `coralh_*` function takes host buffers, stages them to the GPU, runs the
//! gfx950 kernels and copies the outputs back before returning.  Each returns 0 on
//! success, -i when argument i (1-based, CORAL's order) is invalid, or a positive
//! runtime status whose text `coral_last_error` gives.  Dimensions and strides are
//! i64, op codes i32 in the declaration order of coral_aarch64's enums, complex
//! scalars pointers to [re, im].

use std::os::raw::c_char;

#[link(name = "coralhip")]
extern "C" {
    pub fn coral_init() -> i32;
    pub fn coral_last_error() -> *const c_char;
    pub fn coral_status_string(status: i32) -> *const c_char;
    pub fn coralh_sgemm(
        op_a: i32, op_b: i32, m: i64, n: i64, k: i64, alpha: f32, a: *const f32, lda: i64, b: *const f32,
        ldb: i64, beta: f32, c: *mut f32, ldc: i64
    ) -> i32;
    pub fn coralh_dgemm(
        op_a: i32, op_b: i32, m: i64, n: i64, k: i64, alpha: f64, a: *const f64, lda: i64, b: *const f64,
        ldb: i64, beta: f64, c: *mut f64, ldc: i64
    ) -> i32;
    pub fn coralh_cgemm(
        op_a: i32, op_b: i32, m: i64, n: i64, k: i64, alpha: *const f32, a: *const f32, lda: i64,
        b: *const f32, ldb: i64, beta: *const f32, c: *mut f32, ldc: i64
    ) -> i32;
    pub fn coralh_zgemm(
        op_a: i32, op_b: i32, m: i64, n: i64, k: i64, alpha: *const f64, a: *const f64, lda: i64,
        b: *const f64, ldb: i64, beta: *const f64, c: *mut f64, ldc: i64
    ) -> i32;
    pub fn coralh_sgemm_bf16x3(
        op_a: i32, op_b: i32, m: i64, n: i64, k: i64, alpha: f32, a: *const f32, lda: i64, b: *const f32,
        ldb: i64, beta: f32, c: *mut f32, ldc: i64
    ) -> i32;
    pub fn coralh_cgemm_3m_bf16x3(
        op_a: i32, op_b: i32, m: i64, n: i64, k: i64, alpha: *const f32, a: *const f32, lda: i64,
        b: *const f32, ldb: i64, beta: *const f32, c: *mut f32, ldc: i64
    ) -> i32;
    pub fn coralh_sgemm_fp16x3(
        op_a: i32, op_b: i32, m: i64, n: i64, k: i64, alpha: f32, a: *const f32, lda: i64, b: *const f32,
        ldb: i64, beta: f32, c: *mut f32, ldc: i64
    ) -> i32;
    pub fn coralh_cgemm_3m_fp16x3(
        op_a: i32, op_b: i32, m: i64, n: i64, k: i64, alpha: *const f32, a: *const f32, lda: i64,
        b: *const f32, ldb: i64, beta: *const f32, c: *mut f32, ldc: i64
    ) -> i32;
    pub fn coralh_sgemv(
        trans: i32, m: i64, n: i64, alpha: f32, a: *const f32, lda: i64, x: *const f32, incx: i64, beta: f32,
        y: *mut f32, incy: i64
    ) -> i32;
    pub fn coralh_dgemv(
        trans: i32, m: i64, n: i64, alpha: f64, a: *const f64, lda: i64, x: *const f64, incx: i64, beta: f64,
        y: *mut f64, incy: i64
    ) -> i32;
    pub fn coralh_cgemv(
        trans: i32, m: i64, n: i64, alpha: *const f32, a: *const f32, lda: i64, x: *const f32, incx: i64,
        beta: *const f32, y: *mut f32, incy: i64
    ) -> i32;
    pub fn coralh_zgemv(
        trans: i32, m: i64, n: i64, alpha: *const f64, a: *const f64, lda: i64, x: *const f64, incx: i64,
        beta: *const f64, y: *mut f64, incy: i64
    ) -> i32;
    pub fn coralh_ssymv(
        uplo: i32, n: i64, alpha: f32, a: *const f32, lda: i64, x: *const f32, incx: i64, beta: f32,
        y: *mut f32, incy: i64
    ) -> i32;
    pub fn coralh_dsymv(
        uplo: i32, n: i64, alpha: f64, a: *const f64, lda: i64, x: *const f64, incx: i64, beta: f64,
        y: *mut f64, incy: i64
    ) -> i32;
    pub fn coralh_chemv(
        uplo: i32, n: i64, alpha: *const f32, a: *const f32, lda: i64, x: *const f32, incx: i64,
        beta: *const f32, y: *mut f32, incy: i64
    ) -> i32;
    pub fn coralh_zhemv(
        uplo: i32, n: i64, alpha: *const f64, a: *const f64, lda: i64, x: *const f64, incx: i64,
        beta: *const f64, y: *mut f64, incy: i64
    ) -> i32;
    pub fn coralh_strsv(
        uplo: i32, trans: i32, diag: i32, n: i64, a: *const f32, lda: i64, x: *mut f32, incx: i64
    ) -> i32;
    pub fn coralh_dtrsv(
        uplo: i32, trans: i32, diag: i32, n: i64, a: *const f64, lda: i64, x: *mut f64, incx: i64
    ) -> i32;
    pub fn coralh_ctrsv(
        uplo: i32, trans: i32, diag: i32, n: i64, a: *const f32, lda: i64, x: *mut f32, incx: i64
    ) -> i32;
    pub fn coralh_ztrsv(
        uplo: i32, trans: i32, diag: i32, n: i64, a: *const f64, lda: i64, x: *mut f64, incx: i64
    ) -> i32;
    pub fn coralh_strmv(
        uplo: i32, trans: i32, diag: i32, n: i64, a: *const f32, lda: i64, x: *mut f32, incx: i64
    ) -> i32;
    pub fn coralh_dtrmv(
        uplo: i32, trans: i32, diag: i32, n: i64, a: *const f64, lda: i64, x: *mut f64, incx: i64
    ) -> i32;
    pub fn coralh_ctrmv(
        uplo: i32, trans: i32, diag: i32, n: i64, a: *const f32, lda: i64, x: *mut f32, incx: i64
    ) -> i32;
    pub fn coralh_ztrmv(
        uplo: i32, trans: i32, diag: i32, n: i64, a: *const f64, lda: i64, x: *mut f64, incx: i64
    ) -> i32;
    pub fn coralh_sger(
        m: i64, n: i64, alpha: f32, x: *const f32, incx: i64, y: *const f32, incy: i64, a: *mut f32, lda: i64
    ) -> i32;
    pub fn coralh_dger(
        m: i64, n: i64, alpha: f64, x: *const f64, incx: i64, y: *const f64, incy: i64, a: *mut f64, lda: i64
    ) -> i32;
    pub fn coralh_cgeru(
        m: i64, n: i64, alpha: *const f32, x: *const f32, incx: i64, y: *const f32, incy: i64, a: *mut f32,
        lda: i64
    ) -> i32;
    pub fn coralh_cgerc(
        m: i64, n: i64, alpha: *const f32, x: *const f32, incx: i64, y: *const f32, incy: i64, a: *mut f32,
        lda: i64
    ) -> i32;
    pub fn coralh_zgeru(
        m: i64, n: i64, alpha: *const f64, x: *const f64, incx: i64, y: *const f64, incy: i64, a: *mut f64,
        lda: i64
    ) -> i32;
    pub fn coralh_zgerc(
        m: i64, n: i64, alpha: *const f64, x: *const f64, incx: i64, y: *const f64, incy: i64, a: *mut f64,
        lda: i64
    ) -> i32;
    pub fn coralh_ssyr(uplo: i32, n: i64, alpha: f32, x: *const f32, incx: i64, a: *mut f32, lda: i64) -> i32;
    pub fn coralh_dsyr(uplo: i32, n: i64, alpha: f64, x: *const f64, incx: i64, a: *mut f64, lda: i64) -> i32;
    pub fn coralh_ssyr2(
        uplo: i32, n: i64, alpha: f32, x: *const f32, incx: i64, y: *const f32, incy: i64, a: *mut f32,
        lda: i64
    ) -> i32;
    pub fn coralh_dsyr2(
        uplo: i32, n: i64, alpha: f64, x: *const f64, incx: i64, y: *const f64, incy: i64, a: *mut f64,
        lda: i64
    ) -> i32;
    pub fn coralh_cher(uplo: i32, n: i64, alpha: f32, x: *const f32, incx: i64, a: *mut f32, lda: i64) -> i32;
    pub fn coralh_zher(uplo: i32, n: i64, alpha: f64, x: *const f64, incx: i64, a: *mut f64, lda: i64) -> i32;
    pub fn coralh_cher2(
        uplo: i32, n: i64, alpha: *const f32, x: *const f32, incx: i64, y: *const f32, incy: i64, a: *mut f32,
        lda: i64
    ) -> i32;
    pub fn coralh_zher2(
        uplo: i32, n: i64, alpha: *const f64, x: *const f64, incx: i64, y: *const f64, incy: i64, a: *mut f64,
        lda: i64
    ) -> i32;
    pub fn coralh_sdot(n: i64, x: *const f32, incx: i64, y: *const f32, incy: i64, result: *mut f32) -> i32;
    pub fn coralh_ddot(n: i64, x: *const f64, incx: i64, y: *const f64, incy: i64, result: *mut f64) -> i32;
    pub fn coralh_cdotu(n: i64, x: *const f32, incx: i64, y: *const f32, incy: i64, result: *mut f32) -> i32;
    pub fn coralh_cdotc(n: i64, x: *const f32, incx: i64, y: *const f32, incy: i64, result: *mut f32) -> i32;
    pub fn coralh_zdotu(n: i64, x: *const f64, incx: i64, y: *const f64, incy: i64, result: *mut f64) -> i32;
    pub fn coralh_zdotc(n: i64, x: *const f64, incx: i64, y: *const f64, incy: i64, result: *mut f64) -> i32;
    pub fn coralh_saxpy(n: i64, alpha: f32, x: *const f32, incx: i64, y: *mut f32, incy: i64) -> i32;
    pub fn coralh_daxpy(n: i64, alpha: f64, x: *const f64, incx: i64, y: *mut f64, incy: i64) -> i32;
    pub fn coralh_caxpy(n: i64, alpha: *const f32, x: *const f32, incx: i64, y: *mut f32, incy: i64) -> i32;
    pub fn coralh_zaxpy(n: i64, alpha: *const f64, x: *const f64, incx: i64, y: *mut f64, incy: i64) -> i32;
    pub fn coralh_snrm2(n: i64, x: *const f32, incx: i64, result: *mut f32) -> i32;
    pub fn coralh_dnrm2(n: i64, x: *const f64, incx: i64, result: *mut f64) -> i32;
    pub fn coralh_scnrm2(n: i64, x: *const f32, incx: i64, result: *mut f32) -> i32;
    pub fn coralh_dznrm2(n: i64, x: *const f64, incx: i64, result: *mut f64) -> i32;
    pub fn coralh_isamax(n: i64, x: *const f32, incx: i64, result: *mut i64) -> i32;
    pub fn coralh_idamax(n: i64, x: *const f64, incx: i64, result: *mut i64) -> i32;
    pub fn coralh_icamax(n: i64, x: *const f32, incx: i64, result: *mut i64) -> i32;
    pub fn coralh_izamax(n: i64, x: *const f64, incx: i64, result: *mut i64) -> i32;
    pub fn coralh_sscal(n: i64, alpha: f32, x: *mut f32, incx: i64) -> i32;
    pub fn coralh_dscal(n: i64, alpha: f64, x: *mut f64, incx: i64) -> i32;
    pub fn coralh_cscal(n: i64, alpha: *const f32, x: *mut f32, incx: i64) -> i32;
    pub fn coralh_zscal(n: i64, alpha: *const f64, x: *mut f64, incx: i64) -> i32;
    pub fn coralh_csscal(n: i64, alpha: f32, x: *mut f32, incx: i64) -> i32;
    pub fn coralh_zdscal(n: i64, alpha: f64, x: *mut f64, incx: i64) -> i32;
    pub fn coralh_sasum(n: i64, x: *const f32, incx: i64, result: *mut f32) -> i32;
    pub fn coralh_dasum(n: i64, x: *const f64, incx: i64, result: *mut f64) -> i32;
    pub fn coralh_scasum(n: i64, x: *const f32, incx: i64, result: *mut f32) -> i32;
    pub fn coralh_dzasum(n: i64, x: *const f64, incx: i64, result: *mut f64) -> i32;
    pub fn coralh_scopy(n: i64, x: *const f32, incx: i64, y: *mut f32, incy: i64) -> i32;
    pub fn coralh_dcopy(n: i64, x: *const f64, incx: i64, y: *mut f64, incy: i64) -> i32;
    pub fn coralh_ccopy(n: i64, x: *const f32, incx: i64, y: *mut f32, incy: i64) -> i32;
    pub fn coralh_zcopy(n: i64, x: *const f64, incx: i64, y: *mut f64, incy: i64) -> i32;
    pub fn coralh_sswap(n: i64, x: *mut f32, incx: i64, y: *mut f32, incy: i64) -> i32;
    pub fn coralh_dswap(n: i64, x: *mut f64, incx: i64, y: *mut f64, incy: i64) -> i32;
    pub fn coralh_cswap(n: i64, x: *mut f32, incx: i64, y: *mut f32, incy: i64) -> i32;
    pub fn coralh_zswap(n: i64, x: *mut f64, incx: i64, y: *mut f64, incy: i64) -> i32;
    pub fn coralh_srot(n: i64, x: *mut f32, incx: i64, y: *mut f32, incy: i64, c: f32, s: f32) -> i32;
    pub fn coralh_drot(n: i64, x: *mut f64, incx: i64, y: *mut f64, incy: i64, c: f64, s: f64) -> i32;
    pub fn coralh_csrot(n: i64, x: *mut f32, incx: i64, y: *mut f32, incy: i64, c: f32, s: f32) -> i32;
    pub fn coralh_zdrot(n: i64, x: *mut f64, incx: i64, y: *mut f64, incy: i64, c: f64, s: f64) -> i32;
    pub fn coralh_srotm(n: i64, x: *mut f32, incx: i64, y: *mut f32, incy: i64, param: *const f32) -> i32;
    pub fn coralh_drotm(n: i64, x: *mut f64, incx: i64, y: *mut f64, incy: i64, param: *const f64) -> i32;
    pub fn coralh_srotg(a: *mut f32, b: *mut f32, c: *mut f32, s: *mut f32) -> i32;
    pub fn coralh_drotg(a: *mut f64, b: *mut f64, c: *mut f64, s: *mut f64) -> i32;
    pub fn coralh_srotmg(d1: *mut f32, d2: *mut f32, x1: *mut f32, y1: f32, param: *mut f32) -> i32;
    pub fn coralh_drotmg(d1: *mut f64, d2: *mut f64, x1: *mut f64, y1: f64, param: *mut f64) -> i32;
}
