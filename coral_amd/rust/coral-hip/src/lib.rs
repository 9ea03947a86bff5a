//! coral-hip: CORAL's BLAS call surface on one MI355X (gfx950).
//!
//! The modules, names, argument orders and types are those of `coral_aarch64`
//! (coral-aarch64/src/lib.rs:70-82, level1/mod.rs, level2/mod.rs:70-112,
//! level3/mod.rs:47-50, enums.rs:1-18), so a caller switches with
//! `use coral_hip as coral_aarch64;`.  Each function checks CORAL's slice-length
//! contracts (level1/assert_length_helpers.rs:8-26, level2/assert_length_helpers.rs:4-33),
//! which CORAL itself only `debug_assert!`s, then calls the synchronous host-pointer
//! entry point of libcoralhip.so (`ffi`), which stages the operands to the GPU, runs
//! the gfx950 kernels and copies the outputs back.  A bad argument panics, as
//! CORAL's debug build does; a runtime failure panics with the library's message.
//!
//! Beyond CORAL's surface, `split` holds the split-precision GEMMs (BF16x3 / FP16x3).

pub mod ffi;

pub mod enums {
    //! Operation selectors, in CORAL's declaration order (coral-aarch64/src/enums.rs:1-18);
    //! the discriminants are the C ABI's codes.

    #[derive(Debug, Copy, Clone)]
    pub enum CoralTranspose {
        NoTranspose,
        Transpose,
        ConjugateTranspose,
    }

    #[derive(Debug, Copy, Clone)]
    pub enum CoralTriangular {
        UpperTriangular,
        LowerTriangular,
    }

    #[derive(Debug, Copy, Clone)]
    pub enum CoralDiagonal {
        UnitDiagonal,
        NonUnitDiagonal,
    }
}

mod check {
    use crate::ffi;

    /// Panics with the library's message on a non-zero status.
    pub(crate) fn status(name: &str, status: i32) {
        if status != 0 {
            let detail = unsafe { std::ffi::CStr::from_ptr(ffi::coral_last_error()) }
                .to_string_lossy()
                .into_owned();
            panic!("{name}: status {status} ({detail})");
        }
    }

    /// `len >= (n-1)*inc + 1` in elements (complex: twice that, plus one more value).
    pub(crate) fn vec_ok(len: usize, n: usize, inc: usize, cplx: bool) -> bool {
        if n == 0 {
            return true;
        }
        let w = if cplx { 2 } else { 1 };
        len >= (n - 1).saturating_mul(inc.saturating_mul(w)).saturating_add(w)
    }

    /// `len >= (n_cols-1)*lda + n_rows` (complex: twice that); empty matrices pass.
    pub(crate) fn mat_ok(len: usize, n_rows: usize, n_cols: usize, lda: usize, cplx: bool) -> bool {
        if n_rows == 0 || n_cols == 0 {
            return true;
        }
        let w = if cplx { 2 } else { 1 };
        len >= (n_cols - 1).saturating_mul(lda).saturating_add(n_rows).saturating_mul(w)
    }

    pub(crate) fn vec(name: &str, what: &str, len: usize, n: usize, inc: usize, cplx: bool) {
        assert!(vec_ok(len, n, inc, cplx), "{name}: {what} is too short for n = {n}, inc = {inc}");
    }

    pub(crate) fn mat(name: &str, len: usize, n_rows: usize, n_cols: usize, lda: usize, cplx: bool) {
        assert!(lda >= n_rows.max(1), "{name}: lda = {lda} is smaller than n_rows = {n_rows}");
        assert!(
            mat_ok(len, n_rows, n_cols, lda, cplx),
            "{name}: matrix is too short for {n_rows} x {n_cols}, lda = {lda}"
        );
    }
}

macro_rules! call {
    ($name:ident, $ffi:ident ( $($arg:expr),* $(,)? )) => {
        crate::check::status(stringify!($name), unsafe { crate::ffi::$ffi($($arg),*) })
    };
}

pub mod level1 {
    //! Level 1 (coral-aarch64/src/level1/*.rs).  Complex vectors are interleaved
    //! [re, im] slices with strides counted in complex elements.

    macro_rules! reduce {
        ($name:ident, $ffi:ident, $t:ty, $cplx:expr) => {
            #[doc = concat!("`", stringify!($name), "` (coral-aarch64/src/level1/", stringify!($name), ".rs).")]
            pub fn $name(n: usize, x: &[$t], incx: usize) -> $t {
                crate::check::vec(stringify!($name), "x", x.len(), n, incx, $cplx);
                let mut r: $t = 0.0;
                call!($name, $ffi(n as i64, x.as_ptr(), incx as i64, &mut r));
                r
            }
        };
    }
    reduce!(sasum, coralh_sasum, f32, false);
    reduce!(dasum, coralh_dasum, f64, false);
    reduce!(scasum, coralh_scasum, f32, true);
    reduce!(dzasum, coralh_dzasum, f64, true);
    reduce!(snrm2, coralh_snrm2, f32, false);
    reduce!(dnrm2, coralh_dnrm2, f64, false);
    reduce!(scnrm2, coralh_scnrm2, f32, true);
    reduce!(dznrm2, coralh_dznrm2, f64, true);

    macro_rules! iamax {
        ($name:ident, $ffi:ident, $t:ty, $cplx:expr) => {
            #[doc = concat!("`", stringify!($name), "`: 0-based first index of the largest |x_i| (complex |re|+|im|), NaN skipped, 0 for n == 0.")]
            pub fn $name(n: usize, x: &[$t], incx: usize) -> usize {
                crate::check::vec(stringify!($name), "x", x.len(), n, incx, $cplx);
                let mut r: i64 = 0;
                call!($name, $ffi(n as i64, x.as_ptr(), incx as i64, &mut r));
                r as usize
            }
        };
    }
    iamax!(isamax, coralh_isamax, f32, false);
    iamax!(idamax, coralh_idamax, f64, false);
    iamax!(icamax, coralh_icamax, f32, true);
    iamax!(izamax, coralh_izamax, f64, true);

    macro_rules! scal {
        ($name:ident, $ffi:ident, $t:ty, $a:ty, $cplx:expr, |$al:ident| $pass:expr) => {
            #[doc = concat!("`", stringify!($name), "`: x := alpha x.")]
            pub fn $name(n: usize, alpha: $a, x: &mut [$t], incx: usize) {
                crate::check::vec(stringify!($name), "x", x.len(), n, incx, $cplx);
                let $al = &alpha;
                call!($name, $ffi(n as i64, $pass, x.as_mut_ptr(), incx as i64));
            }
        };
    }
    scal!(sscal, coralh_sscal, f32, f32, false, |a| *a);
    scal!(dscal, coralh_dscal, f64, f64, false, |a| *a);
    scal!(cscal, coralh_cscal, f32, [f32; 2], true, |a| a.as_ptr());
    scal!(zscal, coralh_zscal, f64, [f64; 2], true, |a| a.as_ptr());
    scal!(csscal, coralh_csscal, f32, f32, true, |a| *a);
    scal!(zdscal, coralh_zdscal, f64, f64, true, |a| *a);

    macro_rules! copy {
        ($name:ident, $ffi:ident, $t:ty, $cplx:expr) => {
            #[doc = concat!("`", stringify!($name), "`: y := x.")]
            pub fn $name(n: usize, x: &[$t], incx: usize, y: &mut [$t], incy: usize) {
                crate::check::vec(stringify!($name), "x", x.len(), n, incx, $cplx);
                crate::check::vec(stringify!($name), "y", y.len(), n, incy, $cplx);
                call!($name, $ffi(n as i64, x.as_ptr(), incx as i64, y.as_mut_ptr(), incy as i64));
            }
        };
    }
    copy!(scopy, coralh_scopy, f32, false);
    copy!(dcopy, coralh_dcopy, f64, false);
    copy!(ccopy, coralh_ccopy, f32, true);
    copy!(zcopy, coralh_zcopy, f64, true);

    macro_rules! swap {
        ($name:ident, $ffi:ident, $t:ty, $cplx:expr) => {
            #[doc = concat!("`", stringify!($name), "`: x <-> y.")]
            pub fn $name(n: usize, x: &mut [$t], incx: usize, y: &mut [$t], incy: usize) {
                crate::check::vec(stringify!($name), "x", x.len(), n, incx, $cplx);
                crate::check::vec(stringify!($name), "y", y.len(), n, incy, $cplx);
                call!($name, $ffi(n as i64, x.as_mut_ptr(), incx as i64, y.as_mut_ptr(), incy as i64));
            }
        };
    }
    swap!(sswap, coralh_sswap, f32, false);
    swap!(dswap, coralh_dswap, f64, false);
    swap!(cswap, coralh_cswap, f32, true);
    swap!(zswap, coralh_zswap, f64, true);

    macro_rules! axpy {
        ($name:ident, $ffi:ident, $t:ty, $a:ty, $cplx:expr, |$al:ident| $pass:expr) => {
            #[doc = concat!("`", stringify!($name), "`: y := alpha x + y, rounded as CORAL rounds it.")]
            pub fn $name(n: usize, alpha: $a, x: &[$t], incx: usize, y: &mut [$t], incy: usize) {
                crate::check::vec(stringify!($name), "x", x.len(), n, incx, $cplx);
                crate::check::vec(stringify!($name), "y", y.len(), n, incy, $cplx);
                let $al = &alpha;
                call!($name, $ffi(n as i64, $pass, x.as_ptr(), incx as i64, y.as_mut_ptr(), incy as i64));
            }
        };
    }
    axpy!(saxpy, coralh_saxpy, f32, f32, false, |a| *a);
    axpy!(daxpy, coralh_daxpy, f64, f64, false, |a| *a);
    axpy!(caxpy, coralh_caxpy, f32, [f32; 2], true, |a| a.as_ptr());
    axpy!(zaxpy, coralh_zaxpy, f64, [f64; 2], true, |a| a.as_ptr());

    macro_rules! dot {
        ($name:ident, $ffi:ident, $t:ty) => {
            #[doc = concat!("`", stringify!($name), "`: sum x_i y_i.")]
            pub fn $name(n: usize, x: &[$t], incx: usize, y: &[$t], incy: usize) -> $t {
                crate::check::vec(stringify!($name), "x", x.len(), n, incx, false);
                crate::check::vec(stringify!($name), "y", y.len(), n, incy, false);
                let mut r: $t = 0.0;
                call!($name, $ffi(n as i64, x.as_ptr(), incx as i64, y.as_ptr(), incy as i64, &mut r));
                r
            }
        };
        ($name:ident, $ffi:ident, $t:ty, complex) => {
            #[doc = concat!("`", stringify!($name), "`: complex dot product as [re, im] (dotc conjugates x).")]
            pub fn $name(n: usize, x: &[$t], incx: usize, y: &[$t], incy: usize) -> [$t; 2] {
                crate::check::vec(stringify!($name), "x", x.len(), n, incx, true);
                crate::check::vec(stringify!($name), "y", y.len(), n, incy, true);
                let mut r: [$t; 2] = [0.0; 2];
                call!($name, $ffi(n as i64, x.as_ptr(), incx as i64, y.as_ptr(), incy as i64, r.as_mut_ptr()));
                r
            }
        };
    }
    dot!(sdot, coralh_sdot, f32);
    dot!(ddot, coralh_ddot, f64);
    dot!(cdotu, coralh_cdotu, f32, complex);
    dot!(zdotu, coralh_zdotu, f64, complex);
    dot!(cdotc, coralh_cdotc, f32, complex);
    dot!(zdotc, coralh_zdotc, f64, complex);

    macro_rules! rot {
        ($name:ident, $ffi:ident, $t:ty, $r:ty, $cplx:expr) => {
            #[doc = concat!("`", stringify!($name), "`: (x, y) := (c x + s y, c y - s x).")]
            pub fn $name(n: usize, x: &mut [$t], incx: usize, y: &mut [$t], incy: usize, c: $r, s: $r) {
                crate::check::vec(stringify!($name), "x", x.len(), n, incx, $cplx);
                crate::check::vec(stringify!($name), "y", y.len(), n, incy, $cplx);
                call!($name, $ffi(n as i64, x.as_mut_ptr(), incx as i64, y.as_mut_ptr(), incy as i64, c, s));
            }
        };
    }
    rot!(srot, coralh_srot, f32, f32, false);
    rot!(drot, coralh_drot, f64, f64, false);
    rot!(csrot, coralh_csrot, f32, f32, true);
    rot!(zdrot, coralh_zdrot, f64, f64, true);

    macro_rules! rotm {
        ($name:ident, $ffi:ident, $t:ty) => {
            #[doc = concat!("`", stringify!($name), "`: modified rotation with param = [flag, h11, h21, h12, h22].")]
            pub fn $name(n: usize, x: &mut [$t], incx: usize, y: &mut [$t], incy: usize, param: &[$t; 5]) {
                crate::check::vec(stringify!($name), "x", x.len(), n, incx, false);
                crate::check::vec(stringify!($name), "y", y.len(), n, incy, false);
                call!($name, $ffi(n as i64, x.as_mut_ptr(), incx as i64, y.as_mut_ptr(), incy as i64,
                                  param.as_ptr()));
            }
        };
    }
    rotm!(srotm, coralh_srotm, f32);
    rotm!(drotm, coralh_drotm, f64);

    macro_rules! rotg {
        ($name:ident, $ffi:ident, $t:ty) => {
            #[doc = concat!("`", stringify!($name), "`: Givens rotation; a := r, b := z, c and s written.")]
            pub fn $name(a: &mut $t, b: &mut $t, c: &mut $t, s: &mut $t) {
                call!($name, $ffi(a, b, c, s));
            }
        };
    }
    rotg!(srotg, coralh_srotg, f32);
    rotg!(drotg, coralh_drotg, f64);

    macro_rules! rotmg {
        ($name:ident, $ffi:ident, $t:ty) => {
            #[doc = concat!("`", stringify!($name), "`: modified Givens rotation setup.")]
            pub fn $name(sd1: &mut $t, sd2: &mut $t, sx1: &mut $t, sy1: $t, param: &mut [$t; 5]) {
                call!($name, $ffi(sd1, sd2, sx1, sy1, param.as_mut_ptr()));
            }
        };
    }
    rotmg!(srotmg, coralh_srotmg, f32);
    rotmg!(drotmg, coralh_drotmg, f64);
}

pub mod level2 {
    //! Level 2 (coral-aarch64/src/level2/*.rs).
    use crate::enums::{CoralDiagonal, CoralTranspose, CoralTriangular};

    macro_rules! gemv {
        ($name:ident, $ffi:ident, $t:ty, $a:ty, $cplx:expr, |$al:ident, $be:ident| ($pa:expr, $pb:expr)) => {
            #[doc = concat!("`", stringify!($name), "`: y := alpha op(A) x + beta y.")]
            pub fn $name(
                trans: CoralTranspose, n_rows: usize, n_cols: usize, alpha: $a, matrix: &[$t], lda: usize,
                x: &[$t], incx: usize, beta: $a, y: &mut [$t], incy: usize,
            ) {
                let (nx, ny) = match trans {
                    CoralTranspose::NoTranspose => (n_cols, n_rows),
                    _ => (n_rows, n_cols),
                };
                crate::check::mat(stringify!($name), matrix.len(), n_rows, n_cols, lda, $cplx);
                crate::check::vec(stringify!($name), "x", x.len(), nx, incx, $cplx);
                crate::check::vec(stringify!($name), "y", y.len(), ny, incy, $cplx);
                let ($al, $be) = (&alpha, &beta);
                call!($name, $ffi(trans as i32, n_rows as i64, n_cols as i64, $pa, matrix.as_ptr(), lda as i64,
                                  x.as_ptr(), incx as i64, $pb, y.as_mut_ptr(), incy as i64));
            }
        };
    }
    gemv!(sgemv, coralh_sgemv, f32, f32, false, |a, b| (*a, *b));
    gemv!(dgemv, coralh_dgemv, f64, f64, false, |a, b| (*a, *b));
    gemv!(cgemv, coralh_cgemv, f32, [f32; 2], true, |a, b| (a.as_ptr(), b.as_ptr()));
    gemv!(zgemv, coralh_zgemv, f64, [f64; 2], true, |a, b| (a.as_ptr(), b.as_ptr()));

    macro_rules! tri {
        ($name:ident, $ffi:ident, $t:ty, $cplx:expr, $what:expr) => {
            #[doc = concat!("`", stringify!($name), "`: ", $what)]
            pub fn $name(
                uplo: CoralTriangular, transpose: CoralTranspose, diagonal: CoralDiagonal, n: usize,
                matrix: &[$t], lda: usize, x: &mut [$t], incx: usize,
            ) {
                crate::check::mat(stringify!($name), matrix.len(), n, n, lda, $cplx);
                crate::check::vec(stringify!($name), "x", x.len(), n, incx, $cplx);
                call!($name, $ffi(uplo as i32, transpose as i32, diagonal as i32, n as i64, matrix.as_ptr(),
                                  lda as i64, x.as_mut_ptr(), incx as i64));
            }
        };
    }
    tri!(strsv, coralh_strsv, f32, false, "solve op(A) x = b in place, A triangular.");
    tri!(dtrsv, coralh_dtrsv, f64, false, "solve op(A) x = b in place, A triangular.");
    tri!(ctrsv, coralh_ctrsv, f32, true, "solve op(A) x = b in place, A triangular.");
    tri!(ztrsv, coralh_ztrsv, f64, true, "solve op(A) x = b in place, A triangular.");
    tri!(strmv, coralh_strmv, f32, false, "x := op(A) x, A triangular.");
    tri!(dtrmv, coralh_dtrmv, f64, false, "x := op(A) x, A triangular.");
    tri!(ctrmv, coralh_ctrmv, f32, true, "x := op(A) x, A triangular.");
    tri!(ztrmv, coralh_ztrmv, f64, true, "x := op(A) x, A triangular.");

    macro_rules! symv {
        ($name:ident, $ffi:ident, $t:ty, $a:ty, $cplx:expr, |$al:ident, $be:ident| ($pa:expr, $pb:expr)) => {
            #[doc = concat!("`", stringify!($name), "`: y := alpha A x + beta y, A symmetric / Hermitian from `uplo`.")]
            pub fn $name(
                uplo: CoralTriangular, n: usize, alpha: $a, matrix: &[$t], lda: usize, x: &[$t], incx: usize,
                beta: $a, y: &mut [$t], incy: usize,
            ) {
                crate::check::mat(stringify!($name), matrix.len(), n, n, lda, $cplx);
                crate::check::vec(stringify!($name), "x", x.len(), n, incx, $cplx);
                crate::check::vec(stringify!($name), "y", y.len(), n, incy, $cplx);
                let ($al, $be) = (&alpha, &beta);
                call!($name, $ffi(uplo as i32, n as i64, $pa, matrix.as_ptr(), lda as i64, x.as_ptr(),
                                  incx as i64, $pb, y.as_mut_ptr(), incy as i64));
            }
        };
    }
    symv!(ssymv, coralh_ssymv, f32, f32, false, |a, b| (*a, *b));
    symv!(dsymv, coralh_dsymv, f64, f64, false, |a, b| (*a, *b));
    symv!(chemv, coralh_chemv, f32, [f32; 2], true, |a, b| (a.as_ptr(), b.as_ptr()));
    symv!(zhemv, coralh_zhemv, f64, [f64; 2], true, |a, b| (a.as_ptr(), b.as_ptr()));

    macro_rules! ger {
        ($name:ident, $ffi:ident, $t:ty, $a:ty, $cplx:expr, |$al:ident| $pa:expr) => {
            #[doc = concat!("`", stringify!($name), "`: A := alpha x y^T (y^H for gerc) + A.")]
            pub fn $name(
                n_rows: usize, n_cols: usize, alpha: $a, x: &[$t], incx: usize, y: &[$t], incy: usize,
                matrix: &mut [$t], lda: usize,
            ) {
                crate::check::vec(stringify!($name), "x", x.len(), n_rows, incx, $cplx);
                crate::check::vec(stringify!($name), "y", y.len(), n_cols, incy, $cplx);
                crate::check::mat(stringify!($name), matrix.len(), n_rows, n_cols, lda, $cplx);
                let $al = &alpha;
                call!($name, $ffi(n_rows as i64, n_cols as i64, $pa, x.as_ptr(), incx as i64, y.as_ptr(),
                                  incy as i64, matrix.as_mut_ptr(), lda as i64));
            }
        };
    }
    ger!(sger, coralh_sger, f32, f32, false, |a| *a);
    ger!(dger, coralh_dger, f64, f64, false, |a| *a);
    ger!(cgeru, coralh_cgeru, f32, [f32; 2], true, |a| a.as_ptr());
    ger!(cgerc, coralh_cgerc, f32, [f32; 2], true, |a| a.as_ptr());
    ger!(zgeru, coralh_zgeru, f64, [f64; 2], true, |a| a.as_ptr());
    ger!(zgerc, coralh_zgerc, f64, [f64; 2], true, |a| a.as_ptr());

    macro_rules! syr {
        ($name:ident, $ffi:ident, $t:ty, $cplx:expr) => {
            #[doc = concat!("`", stringify!($name), "`: A := alpha x x^T (x^H) + A on the `uplo` triangle; alpha real.")]
            pub fn $name(uplo: CoralTriangular, n: usize, alpha: $t, x: &[$t], incx: usize, matrix: &mut [$t],
                         lda: usize) {
                crate::check::vec(stringify!($name), "x", x.len(), n, incx, $cplx);
                crate::check::mat(stringify!($name), matrix.len(), n, n, lda, $cplx);
                call!($name, $ffi(uplo as i32, n as i64, alpha, x.as_ptr(), incx as i64, matrix.as_mut_ptr(),
                                  lda as i64));
            }
        };
    }
    syr!(ssyr, coralh_ssyr, f32, false);
    syr!(dsyr, coralh_dsyr, f64, false);
    syr!(cher, coralh_cher, f32, true);
    syr!(zher, coralh_zher, f64, true);

    macro_rules! syr2 {
        ($name:ident, $ffi:ident, $t:ty, $a:ty, $cplx:expr, |$al:ident| $pa:expr) => {
            #[doc = concat!("`", stringify!($name), "`: A := alpha x y^T + alpha y x^T (Hermitian form for her2) + A.")]
            pub fn $name(
                uplo: CoralTriangular, n: usize, alpha: $a, x: &[$t], incx: usize, y: &[$t], incy: usize,
                matrix: &mut [$t], lda: usize,
            ) {
                crate::check::vec(stringify!($name), "x", x.len(), n, incx, $cplx);
                crate::check::vec(stringify!($name), "y", y.len(), n, incy, $cplx);
                crate::check::mat(stringify!($name), matrix.len(), n, n, lda, $cplx);
                let $al = &alpha;
                call!($name, $ffi(uplo as i32, n as i64, $pa, x.as_ptr(), incx as i64, y.as_ptr(), incy as i64,
                                  matrix.as_mut_ptr(), lda as i64));
            }
        };
    }
    syr2!(ssyr2, coralh_ssyr2, f32, f32, false, |a| *a);
    syr2!(dsyr2, coralh_dsyr2, f64, f64, false, |a| *a);
    syr2!(cher2, coralh_cher2, f32, [f32; 2], true, |a| a.as_ptr());
    syr2!(zher2, coralh_zher2, f64, [f64; 2], true, |a| a.as_ptr());
}

macro_rules! gemm {
    ($(#[$doc:meta])* $name:ident, $ffi:ident, $t:ty, $a:ty, |$al:ident, $be:ident| ($pa:expr, $pb:expr)) => {
        $(#[$doc])*
        #[allow(clippy::too_many_arguments, clippy::not_unsafe_ptr_arg_deref)]
        pub fn $name(
            op_a: CoralTranspose, op_b: CoralTranspose, m: usize, n: usize, k: usize, alpha: $a,
            a: *const $t, lda: usize, b: *const $t, ldb: usize, beta: $a, c: *mut $t, ldc: usize,
        ) {
            let ($al, $be) = (&alpha, &beta);
            call!($name, $ffi(op_a as i32, op_b as i32, m as i64, n as i64, k as i64, $pa, a, lda as i64, b,
                              ldb as i64, $pb, c, ldc as i64));
        }
    };
}

pub mod level3 {
    //! Level 3 (coral-aarch64/src/level3/{s,d,c,z}gemm.rs): C := alpha op(A) op(B) + beta C,
    //! column-major, raw pointers as CORAL takes them (lengths cannot be checked).
    use crate::enums::CoralTranspose;

    gemm!(
        /// `sgemm` (coral-aarch64/src/level3/sgemm.rs:106), FP32 MFMA.
        sgemm, coralh_sgemm, f32, f32, |a, b| (*a, *b));
    gemm!(
        /// `dgemm` (coral-aarch64/src/level3/dgemm.rs:107), FP64 MFMA.
        dgemm, coralh_dgemm, f64, f64, |a, b| (*a, *b));
    gemm!(
        /// `cgemm` (coral-aarch64/src/level3/cgemm.rs:127), FP32 MFMA, 4M.
        cgemm, coralh_cgemm, f32, [f32; 2], |a, b| (a.as_ptr(), b.as_ptr()));
    gemm!(
        /// `zgemm` (coral-aarch64/src/level3/zgemm.rs:123), FP64 MFMA, 4M.
        zgemm, coralh_zgemm, f64, [f64; 2], |a, b| (a.as_ptr(), b.as_ptr()));
}

pub mod split {
    //! Split-precision GEMMs, no CORAL counterpart: each FP32 operand is split into
    //! two BF16 (or FP16) halves after power-of-two row/column scaling and three MFMA
    //! products are accumulated in FP32 (complex: 3M).  Same arguments as `sgemm` /
    //! `cgemm`.
    use crate::enums::CoralTranspose;

    gemm!(
        /// SGEMM from three BF16 products (16 significand bits per operand).
        sgemm_bf16x3, coralh_sgemm_bf16x3, f32, f32, |a, b| (*a, *b));
    gemm!(
        /// SGEMM from three FP16 products (22 significand bits per operand).
        sgemm_fp16x3, coralh_sgemm_fp16x3, f32, f32, |a, b| (*a, *b));
    gemm!(
        /// CGEMM as three real BF16x3 GEMMs (3M).
        cgemm_3m_bf16x3, coralh_cgemm_3m_bf16x3, f32, [f32; 2], |a, b| (a.as_ptr(), b.as_ptr()));
    gemm!(
        /// CGEMM as three real FP16x3 GEMMs (3M).
        cgemm_3m_fp16x3, coralh_cgemm_3m_fp16x3, f32, [f32; 2], |a, b| (a.as_ptr(), b.as_ptr()));
}

/// Initialises the library on the current gfx950 device; the routines call it on
/// first use, so callers only need it to check for a device up front.
pub fn init() -> Result<(), String> {
    let status = unsafe { ffi::coral_init() };
    if status == 0 {
        Ok(())
    } else {
        Err(unsafe { std::ffi::CStr::from_ptr(ffi::coral_last_error()) }.to_string_lossy().into_owned())
    }
}

#[cfg(test)]
mod tests {
    use crate::check::{mat_ok, vec_ok};

    #[test]
    fn length_contracts() {
        // level1/assert_length_helpers.rs:8-26
        assert!(vec_ok(0, 0, 5, false));
        assert!(vec_ok(7, 3, 3, false) && !vec_ok(6, 3, 3, false));
        assert!(vec_ok(14, 3, 3, true) && !vec_ok(13, 3, 3, true));
        // level2/assert_length_helpers.rs:4-33
        assert!(mat_ok(0, 0, 4, 1, false));
        assert!(mat_ok(5 * 7 + 4, 4, 6, 7, false) && !mat_ok(5 * 7 + 3, 4, 6, 7, false));
        assert!(mat_ok(2 * (5 * 7 + 4), 4, 6, 7, true) && !mat_ok(2 * (5 * 7 + 4) - 1, 4, 6, 7, true));
    }
}
