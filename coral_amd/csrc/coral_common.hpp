// Shared host/device helpers for libcoralhip: status handling, per-stream
// workspaces, launch helpers and small device utilities.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#include "../../include/coral_hip.h"

namespace coral {

// ---------------------------------------------------------------------------
// status / errors
// ---------------------------------------------------------------------------
void set_last_error(const char *fmt, ...);
int hip_fail(hipError_t e, const char *what, const char *file, int line);

#define CORAL_HIP(call)                                                        \
  do {                                                                         \
    hipError_t _e = (call);                                                    \
    if (_e != hipSuccess) return ::coral::hip_fail(_e, #call, __FILE__, __LINE__); \
  } while (0)

#define CORAL_LAUNCH_CHECK()                                                   \
  do {                                                                         \
    hipError_t _e = hipGetLastError();                                         \
    if (_e != hipSuccess) return ::coral::hip_fail(_e, "kernel launch", __FILE__, __LINE__); \
  } while (0)

#define CORAL_TRY(expr)                                                        \
  do {                                                                         \
    int _s = (expr);                                                           \
    if (_s != CORAL_OK) return _s;                                             \
  } while (0)

int ensure_init();                 // selects the device, caches properties
int num_cus();                     // compute units of the current device
hipStream_t resolve(coral_stream_t s);

// Per-stream scratch buffers that grow on demand.  Slot ids separate the
// buffers one routine needs at the same time.  A workspace belongs to one
// stream; calls on one stream are ordered, so reuse across calls is safe.  A slot
// cannot grow while its stream is being captured (the call fails with
// CORAL_ERR_UNSUPPORTED); a graph keeps every buffer its captured work uses.
// WS_ERROR is the stream's error word (zeroed when allocated; kernels set
// ERRF_* bits; coral_stream_check reads and clears it).
enum WsSlot { WS_PACK_A = 0, WS_PACK_B = 1, WS_AUX0 = 2, WS_AUX1 = 3, WS_AUX2 = 4, WS_SMALL = 5, WS_ERROR = 6,
              WS_BATCH = 7, WS_NSLOTS = 8 };
int workspace(hipStream_t stream, int slot, size_t bytes, void **ptr);
enum ErrFlag { ERRF_WAIT_BOUND = 1, ERRF_PROBLEM_ARG = 2 };
int error_word(hipStream_t stream, int **ptr);
// Spin bound of the device-side waits that have one (TRSV's staged handoff);
// coral_testing_set_wait_limit lowers it so tests can make a wait give up.
uint32_t wait_limit();

// Optional event bracketing of dominant kernels (coral_timing_enable).
// timing_begin returns a slot (or -1 when timing is off); timing_end closes it.
enum TimingKind { TK_GEMM = 0, TK_GEMV = 1, TK_NKINDS = 2 };
int timing_begin(int kind, hipStream_t s);
void timing_end(int slot, hipStream_t s);

// Autotuned tile id for (routine, op pair, size bucket); -1 when there is no entry.
int tuned_tile(const char *routine, int opa, int opb, int64_t m, int64_t n, int64_t k);
int tile_count(const char *routine);
int gemm_menu_size(const char *routine);     // gemm.hip
int batched_menu_size(const char *routine);  // batched.hip

// ---------------------------------------------------------------------------
// argument contracts (coral-aarch64/src/level1/assert_length_helpers.rs:8-26,
// coral-aarch64/src/level2/assert_length_helpers.rs:4-33): CORAL takes usize
// dims, positive strides and ld >= rows.
// ---------------------------------------------------------------------------
inline bool valid_op(int32_t op) { return op == CORAL_NO_TRANSPOSE || op == CORAL_TRANSPOSE || op == CORAL_CONJUGATE_TRANSPOSE; }
inline bool valid_uplo(int32_t u) { return u == CORAL_UPPER_TRIANGULAR || u == CORAL_LOWER_TRIANGULAR; }
inline bool valid_diag(int32_t d) { return d == CORAL_UNIT_DIAGONAL || d == CORAL_NON_UNIT_DIAGONAL; }
__host__ __device__ inline int64_t imax64(int64_t a, int64_t b) { return a > b ? a : b; }
__host__ __device__ inline int64_t imin64(int64_t a, int64_t b) { return a < b ? a : b; }
__host__ __device__ inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------
template <typename T> struct cplx { T re, im; };

__device__ __forceinline__ float  dev_abs(float v)  { return fabsf(v); }
__device__ __forceinline__ double dev_abs(double v) { return fabs(v); }

// Wavefront (64-lane) sum via DPP-backed shuffles.
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

}  // namespace coral

namespace coral {
// ---------------------------------------------------------------------------
// element arithmetic shared by the level-1/2 kernels.  E is float, double,
// float2 or double2 (interleaved complex [re, im]).
// ---------------------------------------------------------------------------
template <typename E> struct elem_traits;
template <> struct elem_traits<float>   { using real = float;  static constexpr bool cplx = false; };
template <> struct elem_traits<double>  { using real = double; static constexpr bool cplx = false; };
template <> struct elem_traits<float2>  { using real = float;  static constexpr bool cplx = true; };
template <> struct elem_traits<double2> { using real = double; static constexpr bool cplx = true; };

template <typename E> __host__ __device__ __forceinline__ E ezero() { return E{}; }
template <typename E> __host__ __device__ __forceinline__ E make_e(double re, double im) {
  if constexpr (elem_traits<E>::cplx) { E v; v.x = (typename elem_traits<E>::real)re; v.y = (typename elem_traits<E>::real)im; return v; }
  else { (void)im; return (E)re; }
}
// acc + a*x  (CONJ: acc + conj(a)*x)
template <bool CONJ = false>
__device__ __forceinline__ float emad(float acc, float a, float x) { return fmaf(a, x, acc); }
template <bool CONJ = false>
__device__ __forceinline__ double emad(double acc, double a, double x) { return fma(a, x, acc); }
template <bool CONJ = false>
__device__ __forceinline__ float2 emad(float2 acc, float2 a, float2 x) {
  float ai = CONJ ? -a.y : a.y;
  acc.x = fmaf(a.x, x.x, fmaf(-ai, x.y, acc.x));
  acc.y = fmaf(a.x, x.y, fmaf(ai, x.x, acc.y));
  return acc;
}
template <bool CONJ = false>
__device__ __forceinline__ double2 emad(double2 acc, double2 a, double2 x) {
  double ai = CONJ ? -a.y : a.y;
  acc.x = fma(a.x, x.x, fma(-ai, x.y, acc.x));
  acc.y = fma(a.x, x.y, fma(ai, x.x, acc.y));
  return acc;
}
__device__ __forceinline__ float eadd(float a, float b) { return a + b; }
__device__ __forceinline__ double eadd(double a, double b) { return a + b; }
__device__ __forceinline__ float2 eadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 eadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float emul(float a, float b) { return a * b; }
__device__ __forceinline__ double emul(double a, double b) { return a * b; }
__device__ __forceinline__ float2 emul(float2 a, float2 b) { return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ __forceinline__ double2 emul(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ __forceinline__ float econj(float a) { return a; }
__device__ __forceinline__ double econj(double a) { return a; }
__device__ __forceinline__ float2 econj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ double2 econj(double2 a) { return make_double2(a.x, -a.y); }
template <typename E> __host__ __device__ __forceinline__ bool eiszero(E a) {
  if constexpr (elem_traits<E>::cplx) return a.x == 0 && a.y == 0; else return a == 0;
}
template <typename E> __host__ __device__ __forceinline__ bool eiszero_h(E a) { return eiszero(a); }
template <typename E> __host__ __device__ __forceinline__ bool eisone(E a) {
  if constexpr (elem_traits<E>::cplx) return a.x == 1 && a.y == 0; else return a == 1;
}
template <typename E> __device__ __forceinline__ E eshfl_xor(E v, int off) {
  if constexpr (elem_traits<E>::cplx) { v.x = __shfl_xor(v.x, off, 64); v.y = __shfl_xor(v.y, off, 64); return v; }
  else return __shfl_xor(v, off, 64);
}
// CORAL's two roundings of one axpy element update y += c*x.
//  fused: the NEON body of the unit-stride axpy (vfmaq / vfmsq):
//         y = fma(x, c, y); complex yr = fma(cr, xr, yr); yr = fma(-ci, xi, yr);
//         yi = fma(cr, xi, yi); yi = fma(ci, xr, yi)  (caxpy.rs:86-89)
//  plain: its scalar tail and strided loop, two roundings per product and sum:
//         y = y + c*x; complex yr += cr*xr - ci*xi; yi += cr*xi + ci*xr  (caxpy.rs:144-145)
// The body covers the first floor(len/Q)*Q elements of a unit-stride call
// (coral-aarch64/src/level1/saxpy.rs:63-112 Q = 8, daxpy.rs Q = 4, caxpy.rs:70-132 Q = 4,
// zaxpy.rs:69-131 Q = 2).
// fused multiply-add in the element's own precision (no promotion of float to double)
__device__ __forceinline__ float efma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double efma(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <typename E> struct axpy_body;
template <> struct axpy_body<float>   { static constexpr int64_t Q = 8; };
template <> struct axpy_body<double>  { static constexpr int64_t Q = 4; };
template <> struct axpy_body<float2>  { static constexpr int64_t Q = 4; };
template <> struct axpy_body<double2> { static constexpr int64_t Q = 2; };

__device__ __forceinline__ float axpy_upd(float y, float c, float x, bool fused) {
#pragma clang fp contract(off)
  const float f = __builtin_fmaf(x, c, y);
  const float p = y + c * x;
  return fused ? f : p;
}
__device__ __forceinline__ double axpy_upd(double y, double c, double x, bool fused) {
#pragma clang fp contract(off)
  const double f = __builtin_fma(x, c, y);
  const double p = y + c * x;
  return fused ? f : p;
}
template <typename E>
__device__ __forceinline__ E axpy_upd_c(E y, E c, E x, bool fused) {
#pragma clang fp contract(off)
  using R = typename elem_traits<E>::real;
  R fr = efma(c.x, x.x, y.x);
  fr = efma(-c.y, x.y, fr);
  R fi = efma(c.x, x.y, y.y);
  fi = efma(c.y, x.x, fi);
  const R pr = y.x + (c.x * x.x - c.y * x.y);
  const R pi = y.y + (c.x * x.y + c.y * x.x);
  E r;
  r.x = fused ? fr : pr;
  r.y = fused ? fi : pi;
  return r;
}
__device__ __forceinline__ float2 axpy_upd(float2 y, float2 c, float2 x, bool f) { return axpy_upd_c(y, c, x, f); }
__device__ __forceinline__ double2 axpy_upd(double2 y, double2 c, double2 x, bool f) { return axpy_upd_c(y, c, x, f); }

template <typename E> __device__ __forceinline__ E wave_esum(E v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = eadd(v, eshfl_xor(v, off));
  return v;
}
// y := alpha*s + beta*y, with beta == 0 overwriting y without reading it.
template <typename E> __device__ __forceinline__ E axpby(E alpha, E s, E beta, const E *y, bool beta_zero) {
  E o = emul(alpha, s);
  if (!beta_zero) o = eadd(o, emul(beta, *y));
  return o;
}
}  // namespace coral
