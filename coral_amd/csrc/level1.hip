// Level-1 BLAS for gfx950: dot/dotu/dotc, axpy, nrm2, i?amax, scal.
//
// Reductions are deterministic: a fixed grid (a function of n only) writes one
// partial per workgroup, and a single workgroup combines the partials in a fixed
// tree order.  Results stay in device memory (stream ordered).
#include "coral_common.hpp"

#include <algorithm>

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
    for (int64_t base = (int64_t)blockIdx.x * kThreads * kU; base < nv; base += nt * kU) {
      l1v4u a[kU], b[kU];
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        const int64_t v = base + u * kThreads + t;
        if (v < nv) {
          a[u] = __builtin_nontemporal_load(xv + v);
          if constexpr (Y) b[u] = __builtin_nontemporal_load(yv + v);
        }
      }
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        const int64_t v = base + u * kThreads + t;
        if (v < nv) {
          E ea[EV], eb[EV];
          __builtin_memcpy(ea, &a[u], 16);
          if constexpr (Y) __builtin_memcpy(eb, &b[u], 16);
#pragma unroll
          for (int e = 0; e < EV; ++e) f(v * EV + e, ea[e], Y ? eb[e] : ea[e]);
        }
      }
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
  for (int64_t base = (int64_t)blockIdx.x * kThreads * kU; base < nv; base += nt * kU) {
    l1v4u a[kU], b[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int64_t v = base + u * kThreads + threadIdx.x;
      if (v < nv) { a[u] = __builtin_nontemporal_load(xv + v); b[u] = yv[v]; }
    }
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int64_t v = base + u * kThreads + threadIdx.x;
      if (v < nv) {
        E ea[EV], eb[EV];
        __builtin_memcpy(ea, &a[u], 16);
        __builtin_memcpy(eb, &b[u], 16);
#pragma unroll
        for (int e = 0; e < EV; ++e) eb[e] = axpy_upd(eb[e], alpha, ea[e], v * EV + e < nf);
        __builtin_memcpy(&b[u], eb, 16);
        yv[v] = b[u];
      }
    }
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
// keeps NaN/Inf as NaN)
// ---------------------------------------------------------------------------
template <typename E>
__global__ __launch_bounds__(kThreads) void scal_kernel(int64_t n, E alpha, E *x, int64_t incx) {
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
    E *p = x + i * incx;
    *p = emul(*p, alpha);
  }
}

template <typename E>
__global__ __launch_bounds__(kThreads) void scal_vec_kernel(int64_t n, E alpha, E *__restrict__ x) {
  constexpr int EV = 16 / sizeof(E);
  const int64_t nv = n / EV, nt = (int64_t)gridDim.x * kThreads;
  l1v4u *xv = reinterpret_cast<l1v4u *>(x);
  for (int64_t base = (int64_t)blockIdx.x * kThreads * kU; base < nv; base += nt * kU) {
    l1v4u a[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int64_t v = base + u * kThreads + threadIdx.x;
      if (v < nv) a[u] = xv[v];
    }
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int64_t v = base + u * kThreads + threadIdx.x;
      if (v < nv) {
        E ea[EV];
        __builtin_memcpy(ea, &a[u], 16);
#pragma unroll
        for (int e = 0; e < EV; ++e) ea[e] = emul(ea[e], alpha);
        __builtin_memcpy(&a[u], ea, 16);
        xv[v] = a[u];
      }
    }
  }
  for (int64_t i = nv * EV + (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += nt) x[i] = emul(x[i], alpha);
}

template <typename E>
static int scal_impl(int64_t n, E alpha, E *x, int64_t incx, coral_stream_t stream) {
  if (n < 0) return -1;
  if (incx <= 0) return -4;
  if (n == 0) return CORAL_OK;
  if (!x) return -3;
  CORAL_TRY(ensure_init());
  if (contiguous16((const E *)x, incx)) {
    const int64_t g = std::max<int64_t>(1, std::min<int64_t>(cdiv(n, kThreads * 16), 2048));
    hipLaunchKernelGGL(scal_vec_kernel<E>, dim3((unsigned)g), dim3(kThreads), 0, resolve(stream), n, alpha, x);
    CORAL_LAUNCH_CHECK();
    return CORAL_OK;
  }
  const int64_t g = std::max<int64_t>(1, std::min<int64_t>(cdiv(n, kThreads * 4), 8192));
  hipLaunchKernelGGL(scal_kernel<E>, dim3((unsigned)g), dim3(kThreads), 0, resolve(stream), n, alpha, x, incx);
  CORAL_LAUNCH_CHECK();
  return CORAL_OK;
}

// ---------------------------------------------------------------------------
// nrm2.  Single precision: squares summed in double (no overflow/underflow is
// possible for binary32 inputs), then sqrt and one rounding.  Double precision:
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
struct ValIdx { double v; int64_t i; };

__device__ __forceinline__ ValIdx vi_best(ValIdx a, ValIdx b) {
  if (b.v > a.v || (b.v == a.v && b.i < a.i)) return b;
  return a;
}

template <typename E>
__device__ __forceinline__ double mag(E v) {
  if constexpr (elem_traits<E>::cplx) {
    typename elem_traits<E>::real m = fabs(v.x) + fabs(v.y);  // rounded in element precision
    return (double)m;
  } else {
    return (double)fabs(v);
  }
}

template <typename E>
__global__ __launch_bounds__(kThreads) void iamax_partial_kernel(int64_t n, const E *__restrict__ x, int64_t incx,
                                                                ValIdx *part, int vec) {
  __shared__ ValIdx sh[kThreads];
  ValIdx b{-1.0, INT64_MAX};  // -1: nothing seen (magnitudes are >= 0)
  // each thread visits its indices in increasing order, so the strict '>' keeps the
  // first index of a maximum per thread
  stream_visit<E, false>(n, x, incx, x, incx, vec, [&](int64_t i, E v, E) {
    const double m = mag(v);
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

__global__ __launch_bounds__(kThreads) void iamax_final_kernel(const ValIdx *part, int64_t np, int64_t *result) {
  __shared__ ValIdx sh[kThreads];
  ValIdx b{-1.0, INT64_MAX};
  for (int64_t i = threadIdx.x; i < np; i += kThreads) b = vi_best(b, part[i]);
  sh[threadIdx.x] = b;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] = vi_best(sh[threadIdx.x], sh[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *result = (sh[0].v < 0.0) ? 0 : sh[0].i;
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
  const int64_t g = reduce_grid(n);
  void *part = nullptr;
  CORAL_TRY(workspace(s, WS_SMALL, sizeof(ValIdx) * (size_t)g, &part));
  hipLaunchKernelGGL(iamax_partial_kernel<E>, dim3((unsigned)g), dim3(kThreads), 0, s, n, x, incx, (ValIdx *)part,
                     (int)contiguous16(x, incx));
  CORAL_LAUNCH_CHECK();
  hipLaunchKernelGGL(iamax_final_kernel, dim3(1), dim3(kThreads), 0, s, (const ValIdx *)part, g, result);
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
  return scal_impl<float2>(n, make_float2(alpha[0], alpha[1]), (float2 *)x, incx, s);
}
int coral_zscal(int64_t n, const double alpha[2], double *x, int64_t incx, coral_stream_t s) {
  if (!alpha) return -2;
  return scal_impl<double2>(n, make_double2(alpha[0], alpha[1]), (double2 *)x, incx, s);
}

}  // extern "C"
