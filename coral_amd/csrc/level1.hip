// Level-1 BLAS for gfx950: dot/dotu/dotc, axpy, nrm2, i?amax, scal, asum, copy, swap,
// rot, rotm, rotg, rotmg.
//
// Reductions are deterministic: a fixed grid (a function of n only) writes one
// partial per workgroup, and a single workgroup combines the partials in a fixed
// tree order.  Results stay in device memory (stream ordered).
#include "coral_common.hpp"

#include <algorithm>
#include <type_traits>

namespace coral {
namespace l1 {

constexpr int kThreads = 256;

static int64_t reduce_grid(int64_t n) { return std::max<int64_t>(1, std::min<int64_t>(cdiv(n, kThreads * 8), 1024)); }

typedef uint32_t l1v4u __attribute__((ext_vector_type(4)));

// Visits elements [0, n) of x (and of y when Y) in a per-thread order that depends
// only on n and the grid.  Contiguous 16-byte-aligned inputs are read with 16-byte
// nontemporal loads, kU vectors in flight per thread, each wave instruction covering
// one contiguous 1 KiB run; other inputs one element per thread per step.  Each
// thread visits its elements in increasing index order.  f(i, x_i, y_i).
//
// The vector loops run every step that lies wholly inside the vectors without
// per-vector guards: a guarded load is a branch, and at each branch join the
// compiler waits for every load issued so far, which left one load in flight per
// thread instead of kU.  Only a workgroup's last, partial step is guarded.
constexpr int kU = 4;
template <typename E, bool Y, class F>
__device__ __forceinline__ void stream_visit(int64_t n, const E *__restrict__ x, int64_t incx,
                                             const E *__restrict__ y, int64_t incy, bool vec, F &&f) {
  constexpr int EV = 16 / sizeof(E);
  const int64_t t = threadIdx.x, nt = (int64_t)gridDim.x * kThreads;
  int64_t done = 0;
  if (vec) {
    const int64_t nv = n / EV;  // whole vectors
    const l1v4u *xv = reinterpret_cast<const l1v4u *>(x);
    const l1v4u *yv = reinterpret_cast<const l1v4u *>(y);
    auto step = [&](int64_t base, auto guarded) {
      constexpr bool G = decltype(guarded)::value;
      l1v4u a[kU], b[kU];
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        const int64_t v = base + u * kThreads + t;
        if (!G || v < nv) {
          a[u] = __builtin_nontemporal_load(xv + v);
          if constexpr (Y) b[u] = __builtin_nontemporal_load(yv + v);
        }
      }
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        const int64_t v = base + u * kThreads + t;
        if (!G || v < nv) {
          E ea[EV], eb[EV];
          __builtin_memcpy(ea, &a[u], 16);
          if constexpr (Y) __builtin_memcpy(eb, &b[u], 16);
#pragma unroll
          for (int e = 0; e < EV; ++e) f(v * EV + e, ea[e], Y ? eb[e] : ea[e]);
        }
      }
    };
    for (int64_t base = (int64_t)blockIdx.x * kThreads * kU; base < nv; base += nt * kU) {
      if (base + kU * kThreads <= nv) step(base, std::false_type{});  // whole step in range: no guards
      else step(base, std::true_type{});
    }
    done = nv * EV;
  }
  for (int64_t i = done + (int64_t)blockIdx.x * kThreads + t; i < n; i += nt) {
    const E a = x[i * incx];
    f(i, a, Y ? y[i * incy] : a);
  }
}

template <typename E>
static bool contiguous16(const E *p, int64_t inc) {
  return inc == 1 && ((uintptr_t)p % 16 == 0);
}

// block-wide sum of one value per thread (256 threads = 4 wavefronts)
template <typename A>
__device__ __forceinline__ A block_sum(A v, A *sh) {
  v = wave_esum(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  A r = ezero<A>();
  if (threadIdx.x == 0) {
    r = eadd(eadd(sh[0], sh[1]), eadd(sh[2], sh[3]));
  }
  return r;
}

// ---------------------------------------------------------------------------
// dot: sum_i op(x_i) * y_i, op = conj for dotc
// (coral-aarch64/src/level1/sdot.rs:42, ddot.rs:43, cdotu.rs:46, cdotc.rs:46)
// ---------------------------------------------------------------------------
template <typename E, bool CONJ>
__global__ __launch_bounds__(kThreads) void dot_partial_kernel(int64_t n, const E *__restrict__ x, int64_t incx,
                                                              const E *__restrict__ y, int64_t incy, E *part,
                                                              int vec) {
  __shared__ E sh[4];
  E acc = ezero<E>();
  stream_visit<E, true>(n, x, incx, y, incy, vec, [&](int64_t, E a, E b) { acc = emad<CONJ>(acc, a, b); });
  E r = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = r;
}

template <typename E>
__global__ __launch_bounds__(kThreads) void sum_final_kernel(const E *part, int64_t np, E *result) {
  __shared__ E sh[4];
  E acc = ezero<E>();
  for (int64_t i = threadIdx.x; i < np; i += kThreads) acc = eadd(acc, part[i]);
  E r = block_sum(acc, sh);
  if (threadIdx.x == 0) *result = r;
}

template <typename E>
__global__ void store_zero_kernel(E *result) { *result = ezero<E>(); }

template <typename E, bool CONJ>
static int dot_impl(int64_t n, const E *x, int64_t incx, const E *y, int64_t incy, E *result, coral_stream_t stream) {
  // CORAL argument order: n=1 x=2 incx=3 y=4 incy=5 (result=6)
  if (n < 0) return -1;
  if (incx <= 0) return -3;
  if (incy <= 0) return -5;
  if (!result) return -6;
  if (n > 0 && !x) return -2;
  if (n > 0 && !y) return -4;
  CORAL_TRY(ensure_init());
  hipStream_t s = resolve(stream);
  if (n == 0) {  // coral-aarch64/src/level1/sdot.rs: n == 0 returns 0
    hipLaunchKernelGGL(store_zero_kernel<E>, dim3(1), dim3(1), 0, s, result);
    CORAL_LAUNCH_CHECK();
    return CORAL_OK;
  }
  const int64_t g = reduce_grid(n);
  void *part = nullptr;
  CORAL_TRY(workspace(s, WS_SMALL, sizeof(E) * (size_t)g, &part));
  hipLaunchKernelGGL((dot_partial_kernel<E, CONJ>), dim3((unsigned)g), dim3(kThreads), 0, s, n, x, incx, y, incy,
                     (E *)part, (int)(contiguous16(x, incx) && contiguous16(y, incy)));
  CORAL_LAUNCH_CHECK();
  hipLaunchKernelGGL(sum_final_kernel<E>, dim3(1), dim3(kThreads), 0, s, (const E *)part, g, result);
  CORAL_LAUNCH_CHECK();
  return CORAL_OK;
}

// ---------------------------------------------------------------------------
// axpy: y += alpha*x (coral-aarch64/src/level1/daxpy.rs:42-132); quick return when
// n == 0 or alpha == 0 (:51-53).  Bit-identical to CORAL: element i gets the fused
// update when i < nf (the NEON body of a unit-stride call, nf = floor(n/Q)*Q) and
// the twice-rounded one otherwise (its tail, and every element of a strided call).
// ---------------------------------------------------------------------------
template <typename E>
__global__ __launch_bounds__(kThreads) void axpy_kernel(int64_t n, int64_t nf, E alpha, const E *__restrict__ x,
                                                       int64_t incx, E *__restrict__ y, int64_t incy) {
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
    E *yp = y + i * incy;
    *yp = axpy_upd(*yp, alpha, x[i * incx], i < nf);
  }
}

// contiguous, 16-byte aligned x and y: kU vectors of each in flight per thread
template <typename E>
__global__ __launch_bounds__(kThreads) void axpy_vec_kernel(int64_t n, int64_t nf, E alpha, const E *__restrict__ x,
                                                           E *__restrict__ y) {
  constexpr int EV = 16 / sizeof(E);
  const int64_t nv = n / EV, nt = (int64_t)gridDim.x * kThreads;
  const l1v4u *xv = reinterpret_cast<const l1v4u *>(x);
  l1v4u *yv = reinterpret_cast<l1v4u *>(y);
  auto step = [&](int64_t base, auto guarded) {  // guards only on a partial step (see stream_visit)
    constexpr bool G = decltype(guarded)::value;
    l1v4u a[kU], b[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int64_t v = base + u * kThreads + threadIdx.x;
      if (!G || v < nv) { a[u] = __builtin_nontemporal_load(xv + v); b[u] = yv[v]; }
    }
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int64_t v = base + u * kThreads + threadIdx.x;
      if (!G || v < nv) {
        E ea[EV], eb[EV];
        __builtin_memcpy(ea, &a[u], 16);
        __builtin_memcpy(eb, &b[u], 16);
#pragma unroll
        for (int e = 0; e < EV; ++e) eb[e] = axpy_upd(eb[e], alpha, ea[e], v * EV + e < nf);
        __builtin_memcpy(&b[u], eb, 16);
        yv[v] = b[u];
      }
    }
  };
  for (int64_t base = (int64_t)blockIdx.x * kThreads * kU; base < nv; base += nt * kU) {
    if (base + kU * kThreads <= nv) step(base, std::false_type{});
    else step(base, std::true_type{});
  }
  for (int64_t i = nv * EV + (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += nt)
    y[i] = axpy_upd(y[i], alpha, x[i], i < nf);
}

template <typename E>
static int axpy_impl(int64_t n, E alpha, const E *x, int64_t incx, E *y, int64_t incy, coral_stream_t stream) {
  // n=1 alpha=2 x=3 incx=4 y=5 incy=6
  if (n < 0) return -1;
  if (incx <= 0) return -4;
  if (incy <= 0) return -6;
  if (n == 0 || eiszero_h(alpha)) return CORAL_OK;
  if (!x) return -3;
  if (!y) return -5;
  CORAL_TRY(ensure_init());
  const int64_t nf = (incx == 1 && incy == 1) ? n / axpy_body<E>::Q * axpy_body<E>::Q : 0;
  if (contiguous16(x, incx) && contiguous16((const E *)y, incy)) {
    const int64_t g = std::max<int64_t>(1, std::min<int64_t>(cdiv(n, kThreads * 16), 2048));
    hipLaunchKernelGGL(axpy_vec_kernel<E>, dim3((unsigned)g), dim3(kThreads), 0, resolve(stream), n, nf, alpha, x, y);
    CORAL_LAUNCH_CHECK();
    return CORAL_OK;
  }
  const int64_t g = std::max<int64_t>(1, std::min<int64_t>(cdiv(n, kThreads * 4), 8192));
  hipLaunchKernelGGL(axpy_kernel<E>, dim3((unsigned)g), dim3(kThreads), 0, resolve(stream), n, nf, alpha, x, incx, y,
                     incy);
  CORAL_LAUNCH_CHECK();
  return CORAL_OK;
}

// ---------------------------------------------------------------------------
// scal: x *= alpha (coral-aarch64/src/level1/dscal.rs:30-66 multiplies, so alpha == 0
// keeps NaN/Inf as NaN).  Real x: one rounding.  Complex x (cscal.rs:43-111,
// zscal.rs:41-109): element i < nf (the NEON body of a unit-stride call, nf =
// floor(n/Q)*Q, Q = 4 c, 2 z) gets re' = fma(-xi, ai, xr*ar), im' = fma(xr, ai, xi*ar);
// the others re' = ar*xr - ai*xi, im' = ar*xi + ai*xr with every product rounded.
// Real alpha on complex x (csscal.rs:39-88, zdscal.rs:38-88): each part one rounding.
// ---------------------------------------------------------------------------
template <typename E> struct scal_body;
template <> struct scal_body<float2>  { static constexpr int64_t Q = 4; };
template <> struct scal_body<double2> { static constexpr int64_t Q = 2; };

__device__ __forceinline__ float scal_upd(float x, float a, bool) { return x * a; }
__device__ __forceinline__ double scal_upd(double x, double a, bool) { return x * a; }
template <typename E>
__device__ __forceinline__ E scal_upd_c(E x, E a, bool fused) {
#pragma clang fp contract(off)
  E r;
  const auto fr = efma(-x.y, a.y, x.x * a.x), fi = efma(x.x, a.y, x.y * a.x);
  const auto pr = a.x * x.x - a.y * x.y, pi = a.x * x.y + a.y * x.x;
  r.x = fused ? fr : pr;
  r.y = fused ? fi : pi;
  return r;
}
__device__ __forceinline__ float2 scal_upd(float2 x, float2 a, bool f) { return scal_upd_c(x, a, f); }
__device__ __forceinline__ double2 scal_upd(double2 x, double2 a, bool f) { return scal_upd_c(x, a, f); }
// real alpha on a complex element
__device__ __forceinline__ float2 scal_upd(float2 x, float a, bool) { return make_float2(x.x * a, x.y * a); }
__device__ __forceinline__ double2 scal_upd(double2 x, double a, bool) { return make_double2(x.x * a, x.y * a); }

template <typename E, typename A>
__global__ __launch_bounds__(kThreads) void scal_kernel(int64_t n, int64_t nf, A alpha, E *x, int64_t incx) {
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
    E *p = x + i * incx;
    *p = scal_upd(*p, alpha, i < nf);
  }
}

template <typename E, typename A>
__global__ __launch_bounds__(kThreads) void scal_vec_kernel(int64_t n, int64_t nf, A alpha, E *__restrict__ x) {
  constexpr int EV = 16 / sizeof(E);
  const int64_t nv = n / EV, nt = (int64_t)gridDim.x * kThreads;
  l1v4u *xv = reinterpret_cast<l1v4u *>(x);
  auto step = [&](int64_t base, auto guarded) {  // guards only on a partial step (see stream_visit)
    constexpr bool G = decltype(guarded)::value;
    l1v4u a[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int64_t v = base + u * kThreads + threadIdx.x;
      if (!G || v < nv) a[u] = __builtin_nontemporal_load(xv + v);
    }
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int64_t v = base + u * kThreads + threadIdx.x;
      if (!G || v < nv) {
        E ea[EV];
        __builtin_memcpy(ea, &a[u], 16);
#pragma unroll
        for (int e = 0; e < EV; ++e) ea[e] = scal_upd(ea[e], alpha, v * EV + e < nf);
        __builtin_memcpy(&a[u], ea, 16);
        __builtin_nontemporal_store(a[u], xv + v);
      }
    }
  };
  for (int64_t base = (int64_t)blockIdx.x * kThreads * kU; base < nv; base += nt * kU) {
    if (base + kU * kThreads <= nv) step(base, std::false_type{});
    else step(base, std::true_type{});
  }
  for (int64_t i = nv * EV + (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += nt)
    x[i] = scal_upd(x[i], alpha, i < nf);
}

template <typename E, typename A>
static int scal_impl(int64_t n, A alpha, E *x, int64_t incx, coral_stream_t stream) {
  // n=1 alpha=2 x=3 incx=4
  if (n < 0) return -1;
  if (incx <= 0) return -4;
  if (n == 0) return CORAL_OK;
  if (!x) return -3;
  CORAL_TRY(ensure_init());
  int64_t nf = 0;
  if constexpr (elem_traits<E>::cplx && std::is_same<A, E>::value)
    nf = incx == 1 ? n / scal_body<E>::Q * scal_body<E>::Q : 0;
  if (contiguous16((const E *)x, incx)) {
    const int64_t g = std::max<int64_t>(1, std::min<int64_t>(cdiv(n, kThreads * 16), 2048));
    hipLaunchKernelGGL((scal_vec_kernel<E, A>), dim3((unsigned)g), dim3(kThreads), 0, resolve(stream), n, nf, alpha,
                       x);
    CORAL_LAUNCH_CHECK();
    return CORAL_OK;
  }
  const int64_t g = std::max<int64_t>(1, std::min<int64_t>(cdiv(n, kThreads * 4), 8192));
  hipLaunchKernelGGL((scal_kernel<E, A>), dim3((unsigned)g), dim3(kThreads), 0, resolve(stream), n, nf, alpha, x,
                     incx);
  CORAL_LAUNCH_CHECK();
  return CORAL_OK;
}

// ---------------------------------------------------------------------------
// nrm2.  Single precision: squares summed in double (no overflow/underflow is
// possible for binary32 inputs), then sqrt and one rounding -- the one use of FP64
// in a single-precision routine: the kernel streams x at the HBM rate, so the FP64
// FMAs cost no time, and it needs neither scaling nor a second accumulator.  Double precision:
// Blue's three-accumulator scaled sum of squares (one pass, no division; the
// thresholds are those of Anderson, "A Fresh Look at Blue's Algorithm", 2017).
// NaN anywhere gives NaN and Inf gives Inf (IEEE), where CORAL's chunked scaling
// (coral-aarch64/src/level1/snrm2.rs:64-120, nrm2_helpers.rs:15-65) can skip a NaN.
// ---------------------------------------------------------------------------
struct Blue { double sml, med, big; };
__device__ __forceinline__ Blue blue_add(Blue a, Blue b) { return Blue{a.sml + b.sml, a.med + b.med, a.big + b.big}; }

constexpr double kTsml = 0x1.0p-511, kTbig = 0x1.0p+486, kSsml = 0x1.0p+537, kSbig = 0x1.0p-538;

__device__ __forceinline__ void blue_acc(Blue &b, double v) {
  double ax = fabs(v);
  if (ax > kTbig) { double t = ax * kSbig; b.big = fma(t, t, b.big); }
  else if (ax < kTsml) { double t = ax * kSsml; b.sml = fma(t, t, b.sml); }
  else b.med = fma(ax, ax, b.med);
}

__device__ double blue_finish(Blue b) {
  double scl, sumsq;
  if (b.big > 0) {
    if (b.med > 0 || b.med != b.med) b.big += (b.med * kSbig) * kSbig;
    scl = 1.0 / kSbig;
    sumsq = b.big;
  } else if (b.sml > 0) {
    if (b.med > 0 || b.med != b.med) {
      double ymed = sqrt(b.med), ysml = sqrt(b.sml) / kSsml;
      double ymin = ysml > ymed ? ymed : ysml, ymax = ysml > ymed ? ysml : ymed;
      scl = 1.0;
      double q = ymin / ymax;
      sumsq = ymax * ymax * (1.0 + q * q);
    } else {
      scl = 1.0 / kSsml;
      sumsq = b.sml;
    }
  } else {
    scl = 1.0;
    sumsq = b.med;
  }
  return scl * sqrt(sumsq);
}

template <typename E>
__global__ __launch_bounds__(kThreads) void nrm2_partial_kernel(int64_t n, const E *__restrict__ x, int64_t incx,
                                                               Blue *part, int vec) {
  __shared__ Blue sh[kThreads];
  Blue b{0, 0, 0};
  stream_visit<E, false>(n, x, incx, x, incx, vec, [&](int64_t, E v, E) {
    if constexpr (sizeof(typename elem_traits<E>::real) == 4) {  // binary32: plain double sum of squares
      if constexpr (elem_traits<E>::cplx) { b.med = fma((double)v.x, (double)v.x, b.med); b.med = fma((double)v.y, (double)v.y, b.med); }
      else b.med = fma((double)v, (double)v, b.med);
    } else {
      if constexpr (elem_traits<E>::cplx) { blue_acc(b, v.x); blue_acc(b, v.y); }
      else blue_acc(b, v);
    }
  });
  sh[threadIdx.x] = b;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] = blue_add(sh[threadIdx.x], sh[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

template <typename R>
__global__ __launch_bounds__(kThreads) void nrm2_final_kernel(const Blue *part, int64_t np, R *result, int single) {
  __shared__ Blue sh[kThreads];
  Blue b{0, 0, 0};
  for (int64_t i = threadIdx.x; i < np; i += kThreads) b = blue_add(b, part[i]);
  sh[threadIdx.x] = b;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] = blue_add(sh[threadIdx.x], sh[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *result = single ? (R)sqrt(sh[0].med) : (R)blue_finish(sh[0]);
}

template <typename E, typename R>
static int nrm2_impl(int64_t n, const E *x, int64_t incx, R *result, coral_stream_t stream) {
  // n=1 x=2 incx=3 (result=4)
  if (n < 0) return -1;
  if (incx <= 0) return -3;
  if (!result) return -4;
  if (n > 0 && !x) return -2;
  CORAL_TRY(ensure_init());
  hipStream_t s = resolve(stream);
  if (n == 0) {
    hipLaunchKernelGGL(store_zero_kernel<R>, dim3(1), dim3(1), 0, s, result);
    CORAL_LAUNCH_CHECK();
    return CORAL_OK;
  }
  const int64_t g = reduce_grid(n);
  void *part = nullptr;
  CORAL_TRY(workspace(s, WS_SMALL, sizeof(Blue) * (size_t)g, &part));
  hipLaunchKernelGGL(nrm2_partial_kernel<E>, dim3((unsigned)g), dim3(kThreads), 0, s, n, x, incx, (Blue *)part,
                     (int)contiguous16(x, incx));
  CORAL_LAUNCH_CHECK();
  hipLaunchKernelGGL(nrm2_final_kernel<R>, dim3(1), dim3(kThreads), 0, s, (const Blue *)part, g, result,
                     (int)(sizeof(R) == 4));
  CORAL_LAUNCH_CHECK();
  return CORAL_OK;
}

// ---------------------------------------------------------------------------
// i?amax: 0-based index of the first maximum of |x_i| (complex: |re| + |im| in the
// element precision), NaN never selected, n == 0 -> 0 (coral-aarch64/src/level1/
// isamax.rs:49-55,72-76,96-147, icamax.rs:79-82).  The (value, index) reduction
// keeps the larger value and, on equal values, the smaller index, which makes the
// result independent of the reduction order.
// ---------------------------------------------------------------------------
// Magnitudes stay in the element's precision (float for s/c, double for d/z).
template <typename R> struct ValIdx { R v; int64_t i; };

template <typename R>
__device__ __forceinline__ ValIdx<R> vi_best(ValIdx<R> a, ValIdx<R> b) {
  if (b.v > a.v || (b.v == a.v && b.i < a.i)) return b;
  return a;
}

template <typename E>
__device__ __forceinline__ typename elem_traits<E>::real mag(E v) {
  if constexpr (elem_traits<E>::cplx) return fabs(v.x) + fabs(v.y);  // rounded in element precision
  else return fabs(v);
}

template <typename E>
__global__ __launch_bounds__(kThreads) void iamax_partial_kernel(int64_t n, const E *__restrict__ x, int64_t incx,
                                                                ValIdx<typename elem_traits<E>::real> *part, int vec) {
  using R = typename elem_traits<E>::real;
  __shared__ ValIdx<R> sh[kThreads];
  ValIdx<R> b{R(-1), INT64_MAX};  // -1: nothing seen (magnitudes are >= 0)
  // each thread visits its indices in increasing order, so the strict '>' keeps the
  // first index of a maximum per thread
  stream_visit<E, false>(n, x, incx, x, incx, vec, [&](int64_t i, E v, E) {
    const R m = mag(v);
    if (m == m && m > b.v) { b.v = m; b.i = i; }
  });
  sh[threadIdx.x] = b;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] = vi_best(sh[threadIdx.x], sh[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

template <typename R>
__global__ __launch_bounds__(kThreads) void iamax_final_kernel(const ValIdx<R> *part, int64_t np, int64_t *result) {
  __shared__ ValIdx<R> sh[kThreads];
  ValIdx<R> b{R(-1), INT64_MAX};
  for (int64_t i = threadIdx.x; i < np; i += kThreads) b = vi_best(b, part[i]);
  sh[threadIdx.x] = b;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] = vi_best(sh[threadIdx.x], sh[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *result = (sh[0].v < R(0)) ? 0 : sh[0].i;
}

template <typename E>
static int iamax_impl(int64_t n, const E *x, int64_t incx, int64_t *result, coral_stream_t stream) {
  // n=1 x=2 incx=3 (result=4).  CORAL returns 0 for n == 0 or incx <= 0 (isamax.rs:49)
  if (n < 0) return -1;
  if (!result) return -4;
  CORAL_TRY(ensure_init());
  hipStream_t s = resolve(stream);
  if (n == 0 || incx <= 0) {
    hipLaunchKernelGGL(store_zero_kernel<int64_t>, dim3(1), dim3(1), 0, s, result);
    CORAL_LAUNCH_CHECK();
    return CORAL_OK;
  }
  if (!x) return -2;
  using R = typename elem_traits<E>::real;
  const int64_t g = reduce_grid(n);
  void *part = nullptr;
  CORAL_TRY(workspace(s, WS_SMALL, sizeof(ValIdx<R>) * (size_t)g, &part));
  hipLaunchKernelGGL(iamax_partial_kernel<E>, dim3((unsigned)g), dim3(kThreads), 0, s, n, x, incx, (ValIdx<R> *)part,
                     (int)contiguous16(x, incx));
  CORAL_LAUNCH_CHECK();
  hipLaunchKernelGGL(iamax_final_kernel<R>, dim3(1), dim3(kThreads), 0, s, (const ValIdx<R> *)part, g, result);
  CORAL_LAUNCH_CHECK();
  return CORAL_OK;
}

// ---------------------------------------------------------------------------
// asum: sum of |x_i| (complex |re| + |im|), sasum.rs:40-107, dasum.rs:41-110,
// scasum.rs:40-108, dzasum.rs:40-111.  Deterministic two-level sum as dot.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float eabs1(float v) { return fabsf(v); }
__device__ __forceinline__ double eabs1(double v) { return fabs(v); }
__device__ __forceinline__ float eabs1(float2 v) { return fabsf(v.x) + fabsf(v.y); }
__device__ __forceinline__ double eabs1(double2 v) { return fabs(v.x) + fabs(v.y); }

template <typename E, typename R>
__global__ __launch_bounds__(kThreads) void asum_partial_kernel(int64_t n, const E *__restrict__ x, int64_t incx,
                                                               R *part, int vec) {
  __shared__ R sh[4];
  R acc = 0;
  stream_visit<E, false>(n, x, incx, x, incx, vec, [&](int64_t, E a, E) { acc += eabs1(a); });
  R r = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = r;
}

template <typename E, typename R>
static int asum_impl(int64_t n, const E *x, int64_t incx, R *result, coral_stream_t stream) {
  // n=1 x=2 incx=3 (result=4)
  if (n < 0) return -1;
  if (incx <= 0) return -3;
  if (!result) return -4;
  if (n > 0 && !x) return -2;
  CORAL_TRY(ensure_init());
  hipStream_t s = resolve(stream);
  if (n == 0) {  // sasum.rs:46-48
    hipLaunchKernelGGL(store_zero_kernel<R>, dim3(1), dim3(1), 0, s, result);
    CORAL_LAUNCH_CHECK();
    return CORAL_OK;
  }
  const int64_t g = reduce_grid(n);
  void *part = nullptr;
  CORAL_TRY(workspace(s, WS_SMALL, sizeof(R) * (size_t)g, &part));
  hipLaunchKernelGGL((asum_partial_kernel<E, R>), dim3((unsigned)g), dim3(kThreads), 0, s, n, x, incx, (R *)part,
                     (int)contiguous16(x, incx));
  CORAL_LAUNCH_CHECK();
  hipLaunchKernelGGL(sum_final_kernel<R>, dim3(1), dim3(kThreads), 0, s, (const R *)part, g, result);
  CORAL_LAUNCH_CHECK();
  return CORAL_OK;
}

// ---------------------------------------------------------------------------
// Two-vector element maps: copy, swap, rot, rotm.  op(i, x_i, y_i) updates x_i
// and y_i in place (WX: x is written back).  Unit-stride 16-byte-aligned vectors
// stream with kU nontemporal 16-byte loads of each in flight per thread, every
// wave instruction covering one contiguous 1 KiB run; other strides go element
// by element.
// ---------------------------------------------------------------------------
template <typename E, bool WX, class Op>
__global__ __launch_bounds__(kThreads) void xy_map_kernel(int64_t n, E *x, int64_t incx, E *y, int64_t incy, int vec,
                                                         Op op) {
  constexpr int EV = 16 / sizeof(E);
  const int64_t nt = (int64_t)gridDim.x * kThreads;
  int64_t done = 0;
  if (vec) {
    const int64_t nv = n / EV;
    l1v4u *xv = reinterpret_cast<l1v4u *>(x);
    l1v4u *yv = reinterpret_cast<l1v4u *>(y);
    auto step = [&](int64_t base, auto guarded) {  // guards only on a partial step (see stream_visit)
      constexpr bool G = decltype(guarded)::value;
      l1v4u a[kU], b[kU];
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        const int64_t v = base + u * kThreads + threadIdx.x;
        if (!G || v < nv) {
          a[u] = __builtin_nontemporal_load(xv + v);
          if (WX) b[u] = __builtin_nontemporal_load(yv + v);
        }
      }
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        const int64_t v = base + u * kThreads + threadIdx.x;
        if (!G || v < nv) {
          E ea[EV], eb[EV];
          __builtin_memcpy(ea, &a[u], 16);
          if (WX) __builtin_memcpy(eb, &b[u], 16);
#pragma unroll
          for (int e = 0; e < EV; ++e) op(v * EV + e, ea[e], eb[e]);
          if (WX) {
            __builtin_memcpy(&a[u], ea, 16);
            __builtin_nontemporal_store(a[u], xv + v);
          }
          __builtin_memcpy(&b[u], eb, 16);
          __builtin_nontemporal_store(b[u], yv + v);
        }
      }
    };
    for (int64_t base = (int64_t)blockIdx.x * kThreads * kU; base < nv; base += nt * kU) {
      if (base + kU * kThreads <= nv) step(base, std::false_type{});
      else step(base, std::true_type{});
    }
    done = nv * EV;
  }
  for (int64_t i = done + (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += nt) {
    E xv = x[i * incx], yv = WX ? y[i * incy] : ezero<E>();
    op(i, xv, yv);
    if (WX) x[i * incx] = xv;
    y[i * incy] = yv;
  }
}

template <typename E> struct CopyOp {
  __device__ void operator()(int64_t, E &x, E &y) const { y = x; }
};
template <typename E> struct SwapOp {
  __device__ void operator()(int64_t, E &x, E &y) const {
    const E t = x;
    x = y;
    y = t;
  }
};

// rot: x' = c*x + s*y, y' = c*y - s*x (srot.rs:51-180, drot.rs:48-178, csrot.rs:50-166,
// zdrot.rs:51-166).  Element i < nf (the NEON body of a unit-stride call) gets
// x' = fma(y, s, c*x) and, for real vectors, y' = fma(-s, x, c*y); complex vectors'
// y' and all other elements round every product.
__device__ __forceinline__ float rot_x(float x, float y, float c, float s, bool f) {
#pragma clang fp contract(off)
  return f ? __builtin_fmaf(y, s, c * x) : c * x + s * y;
}
__device__ __forceinline__ double rot_x(double x, double y, double c, double s, bool f) {
#pragma clang fp contract(off)
  return f ? __builtin_fma(y, s, c * x) : c * x + s * y;
}
__device__ __forceinline__ float rot_y(float x, float y, float c, float s, bool f) {
#pragma clang fp contract(off)
  return f ? __builtin_fmaf(-s, x, c * y) : c * y - s * x;
}
__device__ __forceinline__ double rot_y(double x, double y, double c, double s, bool f) {
#pragma clang fp contract(off)
  return f ? __builtin_fma(-s, x, c * y) : c * y - s * x;
}
template <typename E> struct RotOp {
  using R = typename elem_traits<E>::real;
  R c, s;
  int64_t nf;
  __device__ void operator()(int64_t i, E &x, E &y) const {
    const bool f = i < nf;
    if constexpr (elem_traits<E>::cplx) {
      const E x0 = x, y0 = y;
      x.x = rot_x(x0.x, y0.x, c, s, f);
      x.y = rot_x(x0.y, y0.y, c, s, f);
      y.x = rot_y(x0.x, y0.x, c, s, false);
      y.y = rot_y(x0.y, y0.y, c, s, false);
    } else {
      const E x0 = x;
      x = rot_x(x0, y, c, s, f);
      y = rot_y(x0, y, c, s, f);
    }
  }
};

// rotm (srotm.rs:47-412, drotm.rs:47-410): H by flag; fused forms for i < nf as
// the NEON body: flag < 0 x' = fma(y, h12, h11*x), y' = fma(x, h21, h22*y);
// flag == 0 x' = fma(y, h12, x), y' = fma(x, h21, y); flag > 0 x' = fma(x, h11, y),
// y' = h22*y - x.  Other elements round every product.
template <typename E, int FLAG>  // FLAG: 0 full H (flag < 0), 1 flag == 0, 2 flag > 0
struct RotmOp {
  E h11, h21, h12, h22;
  int64_t nf;
  __device__ void operator()(int64_t i, E &x, E &y) const {
#pragma clang fp contract(off)
    const bool f = i < nf;
    const E x0 = x, y0 = y;
    if constexpr (FLAG == 0) {
      x = f ? efma(y0, h12, h11 * x0) : h11 * x0 + h12 * y0;
      y = f ? efma(x0, h21, h22 * y0) : h21 * x0 + h22 * y0;
    } else if constexpr (FLAG == 1) {
      x = f ? efma(y0, h12, x0) : x0 + h12 * y0;
      y = f ? efma(x0, h21, y0) : h21 * x0 + y0;
    } else {
      x = f ? efma(x0, h11, y0) : h11 * x0 + y0;
      y = h22 * y0 - x0;
    }
  }
};

template <typename E, bool WX, class Op>
static int launch_xy(int64_t n, E *x, int64_t incx, E *y, int64_t incy, Op op, coral_stream_t stream) {
  CORAL_TRY(ensure_init());
  const bool vec = contiguous16((const E *)x, incx) && contiguous16((const E *)y, incy);
  const int64_t g = vec ? std::max<int64_t>(1, std::min<int64_t>(cdiv(n, kThreads * 16), 2048))
                        : std::max<int64_t>(1, std::min<int64_t>(cdiv(n, kThreads * 4), 8192));
  hipLaunchKernelGGL((xy_map_kernel<E, WX, Op>), dim3((unsigned)g), dim3(kThreads), 0, resolve(stream), n, x, incx,
                     y, incy, (int)vec, op);
  CORAL_LAUNCH_CHECK();
  return CORAL_OK;
}

// argument checks shared by copy / swap / rot / rotm: n=1 x=2 incx=3 y=4 incy=5
template <typename E>
static int check_xy(int64_t n, const E *x, int64_t incx, const E *y, int64_t incy) {
  if (n < 0) return -1;
  if (incx <= 0) return -3;
  if (incy <= 0) return -5;
  if (n > 0 && !x) return -2;
  if (n > 0 && !y) return -4;
  return CORAL_OK;
}

template <typename E>
static int copy_impl(int64_t n, const E *x, int64_t incx, E *y, int64_t incy, coral_stream_t s) {
  CORAL_TRY(check_xy(n, x, incx, y, incy));
  if (n == 0) return CORAL_OK;  // scopy.rs:37
  return launch_xy<E, false>(n, const_cast<E *>(x), incx, y, incy, CopyOp<E>{}, s);
}
template <typename E>
static int swap_impl(int64_t n, E *x, int64_t incx, E *y, int64_t incy, coral_stream_t s) {
  CORAL_TRY(check_xy(n, x, incx, y, incy));
  if (n == 0) return CORAL_OK;  // sswap.rs:44
  return launch_xy<E, true>(n, x, incx, y, incy, SwapOp<E>{}, s);
}
// Q: values per NEON step of the unit-stride body, counted in real lanes (complex
// vectors run over their 2n floats), so nf in elements = floor(W*n/Q)*Q / W.
template <typename E>
static int rot_impl(int64_t n, E *x, int64_t incx, E *y, int64_t incy, typename elem_traits<E>::real c,
                    typename elem_traits<E>::real s, int64_t q, coral_stream_t st) {
  CORAL_TRY(check_xy(n, x, incx, y, incy));
  if (n == 0) return CORAL_OK;  // srot.rs:61
  constexpr int64_t W = elem_traits<E>::cplx ? 2 : 1;
  RotOp<E> op;
  op.c = c;
  op.s = s;
  op.nf = (incx == 1 && incy == 1) ? (W * n) / q * q / W : 0;
  return launch_xy<E, true>(n, x, incx, y, incy, op, st);
}
template <typename E>
static int rotm_impl(int64_t n, E *x, int64_t incx, E *y, int64_t incy, const E *param, int64_t q,
                     coral_stream_t st) {
  CORAL_TRY(check_xy(n, x, incx, y, incy));
  if (!param) return -6;
  if (n == 0) return CORAL_OK;  // srotm.rs:57
  const E flag = param[0];
  if (flag == (E)-2) return CORAL_OK;  // identity
  const int64_t nf = (incx == 1 && incy == 1) ? n / q * q : 0;
  if (flag < 0) {
    RotmOp<E, 0> op{param[1], param[2], param[3], param[4], nf};
    return launch_xy<E, true>(n, x, incx, y, incy, op, st);
  }
  if (flag == 0) {
    RotmOp<E, 1> op{param[1], param[2], param[3], param[4], nf};
    return launch_xy<E, true>(n, x, incx, y, incy, op, st);
  }
  RotmOp<E, 2> op{param[1], param[2], param[3], param[4], nf};
  return launch_xy<E, true>(n, x, incx, y, incy, op, st);
}

// ---------------------------------------------------------------------------
// rotg / rotmg: scalar constructions on device-resident scalars (one thread).
// srotg.rs:40-75, drotg.rs:40-71; srotmg.rs:42-199, drotmg.rs:41-195.
// Division and square root are IEEE correctly rounded here, as in CORAL.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void rotg_kernel(T *a, T *b, T *c, T *s) {
#pragma clang fp contract(off)
  const T av = *a, bv = *b;
  const T roe = fabs(av) > fabs(bv) ? av : bv;
  const T scale = fabs(av) + fabs(bv);
  if (scale == 0) {
    *c = 1;
    *s = 0;
    *a = 0;
    *b = 0;
    return;
  }
  const T as = av / scale, bs = bv / scale;
  T r = scale * sqrt(as * as + bs * bs);
  if (roe < 0) r = -r;
  const T cv = av / r, sv = bv / r;
  T z = 1;
  if (fabs(av) > fabs(bv)) z = sv;
  if (fabs(bv) >= fabs(av) && cv != 0) z = 1 / cv;
  *c = cv;
  *s = sv;
  *a = r;
  *b = z;
}

// The rescaling loops run only on finite d1 / d2: CORAL's would not end on an
// infinite one.  Each pass scales by GAM^2 = 2^24, so a finite value needs at
// most 90 passes (f64 exponent range); the bound below is a backstop.
template <typename T>
__global__ void rotmg_kernel(T *d1p, T *d2p, T *x1p, T y1, T *param) {
#pragma clang fp contract(off)
  const T gam = 4096, gamsq = gam * gam, rgamsq = 1 / gamsq;
  T d1 = *d1p, d2 = *d2p, x1 = *x1p;
  T flag, h11 = 0, h12 = 0, h21 = 0, h22 = 0;
  if (d1 < 0) {
    flag = -1;
    d1 = d2 = x1 = 0;
  } else {
    const T p2 = d2 * y1;
    if (p2 == 0) {
      param[0] = -2;
      return;
    }
    const T p1 = d1 * x1, q2 = p2 * y1, q1 = p1 * x1;
    if (fabs(q1) > fabs(q2)) {
      h21 = -y1 / x1;
      h12 = p2 / p1;
      const T u = 1 - h12 * h21;
      if (u <= 0) {
        flag = -1;
        d1 = d2 = x1 = 0;
        h11 = h12 = h21 = h22 = 0;
      } else {
        flag = 0;
        d1 /= u;
        d2 /= u;
        x1 *= u;
      }
    } else if (q2 < 0) {
      flag = -1;
      d1 = d2 = x1 = 0;
      h11 = h12 = h21 = h22 = 0;
    } else {
      flag = 1;
      h11 = p1 / p2;
      h22 = x1 / y1;
      const T u = 1 + h11 * h22;
      const T t = d2 / u;
      d2 = d1 / u;
      d1 = t;
      x1 = y1 * u;
    }
    if (d1 != 0 && isfinite(d1)) {
      for (int it = 0; it < 400 && (d1 <= rgamsq || d1 >= gamsq); ++it) {
        if (flag == 0) {
          h11 = 1;
          h22 = 1;
        } else {
          h21 = -1;
          h12 = 1;
        }
        flag = -1;
        if (d1 <= rgamsq) {
          d1 *= gamsq;
          x1 /= gam;
          h11 /= gam;
          h12 /= gam;
        } else {
          d1 /= gamsq;
          x1 *= gam;
          h11 *= gam;
          h12 *= gam;
        }
      }
    }
    if (d2 != 0 && isfinite(d2)) {
      for (int it = 0; it < 400 && (fabs(d2) <= rgamsq || fabs(d2) >= gamsq); ++it) {
        if (flag == 0) {
          h11 = 1;
          h22 = 1;
        } else {
          h21 = -1;
          h12 = 1;
        }
        flag = -1;
        if (fabs(d2) <= rgamsq) {
          d2 *= gamsq;
          h21 /= gam;
          h22 /= gam;
        } else {
          d2 /= gamsq;
          h21 *= gam;
          h22 *= gam;
        }
      }
    }
  }
  *d1p = d1;
  *d2p = d2;
  *x1p = x1;
  param[1] = param[2] = param[3] = param[4] = 0;
  if (flag < 0) {
    param[1] = h11;
    param[2] = h21;
    param[3] = h12;
    param[4] = h22;
  } else if (flag == 0) {
    param[2] = h21;
    param[3] = h12;
  } else {
    param[1] = h11;
    param[4] = h22;
  }
  param[0] = flag;
}

template <typename T>
static int rotg_impl(T *a, T *b, T *c, T *s, coral_stream_t st) {
  if (!a) return -1;
  if (!b) return -2;
  if (!c) return -3;
  if (!s) return -4;
  CORAL_TRY(ensure_init());
  hipLaunchKernelGGL(rotg_kernel<T>, dim3(1), dim3(1), 0, resolve(st), a, b, c, s);
  CORAL_LAUNCH_CHECK();
  return CORAL_OK;
}
template <typename T>
static int rotmg_impl(T *d1, T *d2, T *x1, T y1, T *param, coral_stream_t st) {
  if (!d1) return -1;
  if (!d2) return -2;
  if (!x1) return -3;
  if (!param) return -5;
  CORAL_TRY(ensure_init());
  hipLaunchKernelGGL(rotmg_kernel<T>, dim3(1), dim3(1), 0, resolve(st), d1, d2, x1, y1, param);
  CORAL_LAUNCH_CHECK();
  return CORAL_OK;
}

}  // namespace l1
}  // namespace coral

using namespace coral;
using namespace coral::l1;

extern "C" {

int coral_sdot(int64_t n, const float *x, int64_t incx, const float *y, int64_t incy, float *result, coral_stream_t s) {
  return dot_impl<float, false>(n, x, incx, y, incy, result, s);
}
int coral_ddot(int64_t n, const double *x, int64_t incx, const double *y, int64_t incy, double *result,
               coral_stream_t s) {
  return dot_impl<double, false>(n, x, incx, y, incy, result, s);
}
int coral_cdotu(int64_t n, const float *x, int64_t incx, const float *y, int64_t incy, float *result, coral_stream_t s) {
  return dot_impl<float2, false>(n, (const float2 *)x, incx, (const float2 *)y, incy, (float2 *)result, s);
}
int coral_cdotc(int64_t n, const float *x, int64_t incx, const float *y, int64_t incy, float *result, coral_stream_t s) {
  return dot_impl<float2, true>(n, (const float2 *)x, incx, (const float2 *)y, incy, (float2 *)result, s);
}
int coral_zdotu(int64_t n, const double *x, int64_t incx, const double *y, int64_t incy, double *result,
                coral_stream_t s) {
  return dot_impl<double2, false>(n, (const double2 *)x, incx, (const double2 *)y, incy, (double2 *)result, s);
}
int coral_zdotc(int64_t n, const double *x, int64_t incx, const double *y, int64_t incy, double *result,
                coral_stream_t s) {
  return dot_impl<double2, true>(n, (const double2 *)x, incx, (const double2 *)y, incy, (double2 *)result, s);
}

int coral_saxpy(int64_t n, float alpha, const float *x, int64_t incx, float *y, int64_t incy, coral_stream_t s) {
  return axpy_impl<float>(n, alpha, x, incx, y, incy, s);
}
int coral_daxpy(int64_t n, double alpha, const double *x, int64_t incx, double *y, int64_t incy, coral_stream_t s) {
  return axpy_impl<double>(n, alpha, x, incx, y, incy, s);
}
int coral_caxpy(int64_t n, const float alpha[2], const float *x, int64_t incx, float *y, int64_t incy,
                coral_stream_t s) {
  if (!alpha) return -2;
  return axpy_impl<float2>(n, make_float2(alpha[0], alpha[1]), (const float2 *)x, incx, (float2 *)y, incy, s);
}
int coral_zaxpy(int64_t n, const double alpha[2], const double *x, int64_t incx, double *y, int64_t incy,
                coral_stream_t s) {
  if (!alpha) return -2;
  return axpy_impl<double2>(n, make_double2(alpha[0], alpha[1]), (const double2 *)x, incx, (double2 *)y, incy, s);
}

int coral_snrm2(int64_t n, const float *x, int64_t incx, float *result, coral_stream_t s) {
  return nrm2_impl<float, float>(n, x, incx, result, s);
}
int coral_dnrm2(int64_t n, const double *x, int64_t incx, double *result, coral_stream_t s) {
  return nrm2_impl<double, double>(n, x, incx, result, s);
}
int coral_scnrm2(int64_t n, const float *x, int64_t incx, float *result, coral_stream_t s) {
  return nrm2_impl<float2, float>(n, (const float2 *)x, incx, result, s);
}
int coral_dznrm2(int64_t n, const double *x, int64_t incx, double *result, coral_stream_t s) {
  return nrm2_impl<double2, double>(n, (const double2 *)x, incx, result, s);
}

int coral_isamax(int64_t n, const float *x, int64_t incx, int64_t *result, coral_stream_t s) {
  return iamax_impl<float>(n, x, incx, result, s);
}
int coral_idamax(int64_t n, const double *x, int64_t incx, int64_t *result, coral_stream_t s) {
  return iamax_impl<double>(n, x, incx, result, s);
}
int coral_icamax(int64_t n, const float *x, int64_t incx, int64_t *result, coral_stream_t s) {
  return iamax_impl<float2>(n, (const float2 *)x, incx, result, s);
}
int coral_izamax(int64_t n, const double *x, int64_t incx, int64_t *result, coral_stream_t s) {
  return iamax_impl<double2>(n, (const double2 *)x, incx, result, s);
}

int coral_sscal(int64_t n, float alpha, float *x, int64_t incx, coral_stream_t s) {
  return scal_impl<float>(n, alpha, x, incx, s);
}
int coral_dscal(int64_t n, double alpha, double *x, int64_t incx, coral_stream_t s) {
  return scal_impl<double>(n, alpha, x, incx, s);
}
int coral_cscal(int64_t n, const float alpha[2], float *x, int64_t incx, coral_stream_t s) {
  if (!alpha) return -2;
  return scal_impl<float2, float2>(n, make_float2(alpha[0], alpha[1]), (float2 *)x, incx, s);
}
int coral_zscal(int64_t n, const double alpha[2], double *x, int64_t incx, coral_stream_t s) {
  if (!alpha) return -2;
  return scal_impl<double2, double2>(n, make_double2(alpha[0], alpha[1]), (double2 *)x, incx, s);
}
int coral_csscal(int64_t n, float alpha, float *x, int64_t incx, coral_stream_t s) {
  return scal_impl<float2, float>(n, alpha, (float2 *)x, incx, s);
}
int coral_zdscal(int64_t n, double alpha, double *x, int64_t incx, coral_stream_t s) {
  return scal_impl<double2, double>(n, alpha, (double2 *)x, incx, s);
}

int coral_sasum(int64_t n, const float *x, int64_t incx, float *result, coral_stream_t s) {
  return asum_impl<float, float>(n, x, incx, result, s);
}
int coral_dasum(int64_t n, const double *x, int64_t incx, double *result, coral_stream_t s) {
  return asum_impl<double, double>(n, x, incx, result, s);
}
int coral_scasum(int64_t n, const float *x, int64_t incx, float *result, coral_stream_t s) {
  return asum_impl<float2, float>(n, (const float2 *)x, incx, result, s);
}
int coral_dzasum(int64_t n, const double *x, int64_t incx, double *result, coral_stream_t s) {
  return asum_impl<double2, double>(n, (const double2 *)x, incx, result, s);
}

#define CORAL_COPY_SWAP(P, T, E)                                                                                  \
  int coral_##P##copy(int64_t n, const T *x, int64_t incx, T *y, int64_t incy, coral_stream_t s) {              \
    return copy_impl<E>(n, (const E *)x, incx, (E *)y, incy, s);                                                 \
  }                                                                                                               \
  int coral_##P##swap(int64_t n, T *x, int64_t incx, T *y, int64_t incy, coral_stream_t s) {                    \
    return swap_impl<E>(n, (E *)x, incx, (E *)y, incy, s);                                                       \
  }
CORAL_COPY_SWAP(s, float, float)
CORAL_COPY_SWAP(d, double, double)
CORAL_COPY_SWAP(c, float, float2)
CORAL_COPY_SWAP(z, double, double2)

int coral_srot(int64_t n, float *x, int64_t incx, float *y, int64_t incy, float c, float s, coral_stream_t st) {
  return rot_impl<float>(n, x, incx, y, incy, c, s, 4, st);  // srot.rs: 32- then 4-lane steps
}
int coral_drot(int64_t n, double *x, int64_t incx, double *y, int64_t incy, double c, double s, coral_stream_t st) {
  return rot_impl<double>(n, x, incx, y, incy, c, s, 2, st);  // drot.rs: 16- then 2-lane steps
}
int coral_csrot(int64_t n, float *x, int64_t incx, float *y, int64_t incy, float c, float s, coral_stream_t st) {
  return rot_impl<float2>(n, (float2 *)x, incx, (float2 *)y, incy, c, s, 4, st);  // csrot.rs: 16 then 4 floats
}
int coral_zdrot(int64_t n, double *x, int64_t incx, double *y, int64_t incy, double c, double s,
                coral_stream_t st) {
  return rot_impl<double2>(n, (double2 *)x, incx, (double2 *)y, incy, c, s, 2, st);  // zdrot.rs: 8 then 2
}
int coral_srotm(int64_t n, float *x, int64_t incx, float *y, int64_t incy, const float param[5], coral_stream_t st) {
  return rotm_impl<float>(n, x, incx, y, incy, param, 4, st);
}
int coral_drotm(int64_t n, double *x, int64_t incx, double *y, int64_t incy, const double param[5],
                coral_stream_t st) {
  return rotm_impl<double>(n, x, incx, y, incy, param, 2, st);
}
int coral_srotg(float *a, float *b, float *c, float *s, coral_stream_t st) { return rotg_impl<float>(a, b, c, s, st); }
int coral_drotg(double *a, double *b, double *c, double *s, coral_stream_t st) {
  return rotg_impl<double>(a, b, c, s, st);
}
int coral_srotmg(float *d1, float *d2, float *x1, float y1, float *param, coral_stream_t st) {
  return rotmg_impl<float>(d1, d2, x1, y1, param, st);
}
int coral_drotmg(double *d1, double *d2, double *x1, double y1, double *param, coral_stream_t st) {
  return rotmg_impl<double>(d1, d2, x1, y1, param, st);
}

}  // extern "C"
