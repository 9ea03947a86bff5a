// Level-2 GEMV for gfx950: y := alpha*op(A)*x + beta*y for s/d/c/z, op in {N, T, C}.
//
// CORAL streams A through the fused `?axpyf` (N, coral-aarch64/src/level1_special/daxpyf.rs:45)
// and `?dotf` (T, ddotf.rs:37; C, cdotcf.rs:42) kernels after pre-scaling y by beta
// (dgemv_notranspose.rs:88) and x by alpha (:93).  Here both variants are one pass of
// HBM streaming over A with deterministic two-level reductions:
//
//  N: a workgroup owns 256*NV*EV rows x a chunk of columns (EV elements per 16-byte
//     vector); each thread keeps NV*EV row sums and reads NV vectors of every
//     column it visits, four columns in flight.  Column chunks write partial row
//     sums, which a second kernel adds in chunk order and combines with alpha/beta.
//  T/C: a wavefront owns four columns and reads each x vector once for all four;
//     the dot products are finished with a DPP/shuffle butterfly.  Long columns are
//     split into row chunks whose partial dots are reduced the same way.
//  A is streamed once, with nontemporal loads.
#include "coral_common.hpp"

#include <type_traits>

#include <algorithm>
#include <mutex>
#include <stdio.h>
#include <stdlib.h>

namespace coral {
namespace l2 {

typedef uint32_t gv4u __attribute__((ext_vector_type(4)));

// One 16-byte vector (EV = 16 / sizeof(E) elements) at p: a nontemporal load when
// the whole vector is in range and aligned, otherwise guarded element loads.  The
// lanes of a wave always cover one contiguous 1 KiB run per load instruction: with
// nontemporal loads a run split over two instructions is fetched twice.
template <typename E>
__device__ __forceinline__ void load_vec(E (&v)[16 / sizeof(E)], const E *p, bool vec, int64_t avail) {
  constexpr int EV = 16 / sizeof(E);
  if (vec && avail >= EV) {
    const gv4u pk = __builtin_nontemporal_load(reinterpret_cast<const gv4u *>(p));
    __builtin_memcpy(&v[0], &pk, 16);
  } else {
#pragma unroll
    for (int r = 0; r < EV; ++r) v[r] = (r < avail) ? p[r] : ezero<E>();
  }
}

// N: a workgroup owns 256 * NV * EV rows; lane l of wave w holds rows
// base_w + v*64*EV + l*EV + e (v < NV, e < EV) and reads them from every column of
// its chunk, four columns in flight.
template <typename E, int NV>
__global__ __launch_bounds__(256) void gemv_n_kernel(int64_t m, int64_t n, const E *__restrict__ A, int64_t lda,
                                                     const E *__restrict__ x, int64_t incx, E *__restrict__ P,
                                                     int64_t cpc, int vec_ok) {
  constexpr int EV = 16 / sizeof(E);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t wbase = (int64_t)blockIdx.x * (256 * NV * EV) + (int64_t)w * (64 * NV * EV);
  if (wbase >= m) return;  // uniform per wave; no barrier below
  const int64_t c0 = (int64_t)blockIdx.y * cpc;
  const int64_t c1 = c0 + cpc < n ? c0 + cpc : n;
  int64_t row[NV], avail[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    row[v] = wbase + v * 64 * EV + lane * EV;
    avail[v] = m - row[v];
  }
  E acc[NV][EV];
#pragma unroll
  for (int v = 0; v < NV; ++v)
#pragma unroll
    for (int e = 0; e < EV; ++e) acc[v][e] = ezero<E>();
  const bool vec = vec_ok;
  int64_t j = c0;
  for (; j + 4 <= c1; j += 4) {
    E a[4][NV][EV];
    E xv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      xv[u] = x[(j + u) * incx];
      const E *col = A + (j + u) * lda;
#pragma unroll
      for (int v = 0; v < NV; ++v) load_vec<E>(a[u][v], col + row[v], vec, avail[v]);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int v = 0; v < NV; ++v)
#pragma unroll
        for (int e = 0; e < EV; ++e) acc[v][e] = emad(acc[v][e], a[u][v][e], xv[u]);
  }
  for (; j < c1; ++j) {
    const E xv = x[j * incx];
    const E *col = A + j * lda;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      E a[EV];
      load_vec<E>(a, col + row[v], vec, avail[v]);
#pragma unroll
      for (int e = 0; e < EV; ++e) acc[v][e] = emad(acc[v][e], a[e], xv);
    }
  }
  E *out = P + (int64_t)blockIdx.y * m;
#pragma unroll
  for (int v = 0; v < NV; ++v)
#pragma unroll
    for (int e = 0; e < EV; ++e)
      if (e < avail[v]) out[row[v] + e] = acc[v][e];
}

// T/C: a wavefront owns CPW consecutive columns; per step lane l reads the
// vectors at i + v*64*EV + l*EV (v < NV) of x once and of each of its columns, so
// x is read once per CPW columns.  Measured at 16384^2 / 8192^2: double-precision
// elements stream faster with one vector of eight columns (DGEMV T 6.17 -> 6.38
// TB/s, ZGEMV C 5.84 -> 6.32; ZGEMV C 5.90-5.95 with sixteen columns or one vector
// of four), complex float with two vectors of four columns (CGEMV C 5.94-5.96 TB/s;
// 5.72 with one vector of eight columns, 5.68-5.73 with two of eight, 5.58 with one
// of sixteen, 5.88-5.90 with four of four, 5.57-5.62 with four of two), real float
// with one vector of sixteen columns (SGEMV T 6.20-6.25 -> 6.68 TB/s; 6.37 with one
// of eight, 6.39 with one of thirty-two, 6.24 with two of eight, 6.69 with two of
// sixteen).
template <typename E>
struct GemvTShape {
  static constexpr bool kDouble = std::is_same<typename elem_traits<E>::real, double>::value;
  static constexpr bool kReal32 = std::is_same<E, float>::value;
  static constexpr int NV = kDouble ? 1 : (kReal32 ? 1 : 2), CPW = kDouble ? 8 : (kReal32 ? 16 : 4);
};
template <typename E, bool CONJ>
__global__ __launch_bounds__(256) void gemv_t_kernel(int64_t m, int64_t n, const E *__restrict__ A, int64_t lda,
                                                     const E *__restrict__ x, int64_t incx, E *__restrict__ P,
                                                     int64_t rpc, int vec_ok, int xvec_ok) {
  constexpr int EV = 16 / sizeof(E), NV = GemvTShape<E>::NV, CPW = GemvTShape<E>::CPW;
  const int lane = threadIdx.x & 63;
  const int64_t j0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * CPW;
  if (j0 >= n) return;  // uniform per wavefront; no block-level barrier below
  const int64_t r0 = (int64_t)blockIdx.y * rpc;
  const int64_t r1 = r0 + rpc < m ? r0 + rpc : m;
  const int ncols = n - j0 < CPW ? (int)(n - j0) : CPW;
  E acc[CPW];
#pragma unroll
  for (int u = 0; u < CPW; ++u) acc[u] = ezero<E>();
  for (int64_t i = r0; i < r1; i += 64 * EV * NV) {
    E xv[NV][EV], a[NV][CPW][EV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      const int64_t rr = i + v * 64 * EV + lane * EV;
      const int64_t avail = r1 - rr;
      if (incx == 1) {
        load_vec<E>(xv[v], x + rr, xvec_ok, avail);
      } else {
#pragma unroll
        for (int e = 0; e < EV; ++e) xv[v][e] = (e < avail) ? x[(rr + e) * incx] : ezero<E>();
      }
#pragma unroll
      for (int u = 0; u < CPW; ++u) {
        if (u < ncols) load_vec<E>(a[v][u], A + (j0 + u) * lda + rr, vec_ok, avail);
        else {
#pragma unroll
          for (int e = 0; e < EV; ++e) a[v][u][e] = ezero<E>();
        }
      }
    }
#pragma unroll
    for (int v = 0; v < NV; ++v)
#pragma unroll
      for (int u = 0; u < CPW; ++u)
#pragma unroll
        for (int e = 0; e < EV; ++e) acc[u] = emad<CONJ>(acc[u], a[v][u][e], xv[v][e]);
  }
#pragma unroll
  for (int u = 0; u < CPW; ++u) {
    const E sum = wave_esum(acc[u]);
    if (lane == 0 && u < ncols) P[(int64_t)blockIdx.y * n + j0 + u] = sum;
  }
}

// y[i] := alpha * sum_c P[c][i] + beta * y[i].  A workgroup takes 64 rows; its four
// wavefronts sum the chunk ranges [g*q, (g+1)*q) in order (eight loads in flight),
// and the four group sums are added in group order: a fixed order, so the result
// does not depend on scheduling.
template <typename E>
__global__ __launch_bounds__(256) void gemv_reduce_kernel(int64_t len, int64_t nchunks, const E *__restrict__ P,
                                                          E alpha, E beta, int beta_zero, E *y, int64_t incy) {
  __shared__ E part[4][64];
  const int lane = threadIdx.x & 63, g = threadIdx.x >> 6;
  const int64_t i = (int64_t)blockIdx.x * 64 + lane;
  const int64_t q = (nchunks + 3) / 4;
  const int64_t c0 = g * q, c1 = c0 + q < nchunks ? c0 + q : nchunks;
  E s = ezero<E>();
  if (i < len) {
    int64_t c = c0;
    for (; c + 8 <= c1; c += 8) {
      E v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = P[(c + u) * len + i];
#pragma unroll
      for (int u = 0; u < 8; ++u) s = eadd(s, v[u]);
    }
    for (; c < c1; ++c) s = eadd(s, P[c * len + i]);
  }
  part[g][lane] = s;
  __syncthreads();
  if (g == 0 && i < len) {
    const E t = eadd(eadd(part[0][lane], part[1][lane]), eadd(part[2][lane], part[3][lane]));
    E *yp = y + i * incy;
    *yp = axpby(alpha, t, beta, yp, beta_zero);
  }
}

// y := beta*y (alpha == 0 path); beta == 0 stores zeros without reading y.
template <typename E>
__global__ void beta_only_kernel(int64_t len, E beta, int beta_zero, E *y, int64_t incy) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += (int64_t)gridDim.x * blockDim.x) {
    E *yp = y + i * incy;
    *yp = beta_zero ? ezero<E>() : emul(beta, *yp);
  }
}

template <typename E>
static bool is_one(E v) {
  if constexpr (elem_traits<E>::cplx) return v.x == 1 && v.y == 0;
  else return v == 1;
}
template <typename E>
static bool is_zero_h(E v) {
  if constexpr (elem_traits<E>::cplx) return v.x == 0 && v.y == 0;
  else return v == 0;
}

template <typename E>
int launch_beta_only(int64_t len, E beta, E *y, int64_t incy, hipStream_t s) {
  if (is_one(beta)) return CORAL_OK;
  int64_t blocks = std::min<int64_t>(cdiv(len, 256), 1024);
  hipLaunchKernelGGL(beta_only_kernel<E>, dim3((unsigned)blocks), dim3(256), 0, s, len, beta, (int)is_zero_h(beta), y,
                     incy);
  CORAL_LAUNCH_CHECK();
  return CORAL_OK;
}

// Partial-sum buffer of a GEMV call: the stream's workspace, or (fixed != nullptr) a
// caller-provided buffer of cap bytes, in which case the chunk count is cut to fit
// and nothing is allocated.
struct GemvScratch {
  void *fixed = nullptr;
  size_t cap = 0;
};

// CORAL argument order: trans=1 m=2 n=3 alpha=4 a=5 lda=6 x=7 incx=8 beta=9 y=10 incy=11
template <typename E>
int gemv_impl(int32_t trans, int64_t m, int64_t n, E alpha, const E *A, int64_t lda, const E *x, int64_t incx,
              E beta, E *y, int64_t incy, coral_stream_t stream, GemvScratch fs = GemvScratch()) {
  if (!valid_op(trans)) return -1;
  if (m < 0) return -2;
  if (n < 0) return -3;
  if (lda < imax64(1, m)) return -6;
  if (incx <= 0) return -8;
  if (incy <= 0) return -11;
  // quick return, coral-aarch64/src/level2/dgemv_notranspose.rs:75-76
  if (m == 0 || n == 0) return CORAL_OK;
  if (is_zero_h(alpha) && is_one(beta)) return CORAL_OK;
  const bool notrans = trans == CORAL_NO_TRANSPOSE;
  const int64_t ylen = notrans ? m : n;
  const int64_t xlen = notrans ? n : m;
  if (!y) return -10;
  CORAL_TRY(ensure_init());
  hipStream_t s = resolve(stream);
  if (is_zero_h(alpha)) return launch_beta_only<E>(ylen, beta, y, incy, s);
  if (!A) return -5;
  if (!x) return -7;
  (void)xlen;
  const int cus = num_cus();
  const bool a_aligned = ((uintptr_t)A % 16 == 0) && ((lda * (int64_t)sizeof(E)) % 16 == 0);
  if (notrans) {
    // 16-byte vectors per column a thread streams (2 or 4) and the workgroup-count
    // target per CU; CORAL_GEMV_CFG=<vectors>,<wg_per_cu> overrides.  Measured on
    // 16384^2 f64 with nontemporal loads: 2 vectors x 1 workgroup/CU 6.39 TB/s,
    // x 2 6.18, x 4 6.04 (more column chunks also add partial-sum traffic).
    int cfg_nv = 2, cfg_wg = 1;
    if (const char *e = getenv("CORAL_GEMV_CFG")) sscanf(e, "%d,%d", &cfg_nv, &cfg_wg);
    if (cfg_nv != 4) cfg_nv = 2;
    if (cfg_wg < 1 || cfg_wg > 64) cfg_wg = 1;
    const int64_t rows_per_wg = 256 * cfg_nv * (16 / (int)sizeof(E));
    const int64_t row_blocks = cdiv(m, rows_per_wg);
    // enough column chunks for cfg_wg workgroups per CU, each chunk >= 32 columns
    int64_t chunks = std::max<int64_t>(1, std::min<int64_t>(cdiv((int64_t)cfg_wg * cus, row_blocks), cdiv(n, 32)));
    if (fs.fixed) chunks = std::min<int64_t>(chunks, (int64_t)(fs.cap / (sizeof(E) * (size_t)m)));
    const int64_t cpc = cdiv(n, chunks);
    chunks = cdiv(n, cpc);
    void *P = fs.fixed;
    if (!P) CORAL_TRY(workspace(s, WS_AUX2, sizeof(E) * (size_t)chunks * (size_t)m, &P));
    const int slot = fs.fixed ? -1 : timing_begin(TK_GEMV, s);
    if (cfg_nv == 4)
      hipLaunchKernelGGL((gemv_n_kernel<E, 4>), dim3((unsigned)row_blocks, (unsigned)chunks), dim3(256), 0, s, m, n,
                         A, lda, x, incx, (E *)P, cpc, (int)a_aligned);
    else
      hipLaunchKernelGGL((gemv_n_kernel<E, 2>), dim3((unsigned)row_blocks, (unsigned)chunks), dim3(256), 0, s, m, n,
                         A, lda, x, incx, (E *)P, cpc, (int)a_aligned);
    CORAL_LAUNCH_CHECK();
    timing_end(slot, s);
    hipLaunchKernelGGL(gemv_reduce_kernel<E>, dim3((unsigned)cdiv(m, 64)), dim3(256), 0, s, m, chunks,
                       (const E *)P, alpha, beta, (int)is_zero_h(beta), y, incy);
    CORAL_LAUNCH_CHECK();
  } else {
    const int64_t col_blocks = cdiv(n, 4 * GemvTShape<E>::CPW);
    const int64_t unit = 64 * GemvTShape<E>::NV * (16 / (int64_t)sizeof(E));  // rows per wave step
    int64_t chunks = std::max<int64_t>(1, std::min<int64_t>(cdiv(8 * cus, col_blocks), cdiv(m, 4 * unit)));
    if (fs.fixed) chunks = std::min<int64_t>(chunks, (int64_t)(fs.cap / (sizeof(E) * (size_t)n)));
    const int64_t rpc = cdiv(cdiv(m, chunks), unit) * unit;
    chunks = cdiv(m, rpc);
    const bool x_aligned = ((uintptr_t)x % 16 == 0);
    void *P = fs.fixed;
    if (!P) CORAL_TRY(workspace(s, WS_AUX2, sizeof(E) * (size_t)chunks * (size_t)n, &P));
    const int slot = fs.fixed ? -1 : timing_begin(TK_GEMV, s);
    if (trans == CORAL_CONJUGATE_TRANSPOSE && elem_traits<E>::cplx)
      hipLaunchKernelGGL((gemv_t_kernel<E, true>), dim3((unsigned)col_blocks, (unsigned)chunks), dim3(256), 0, s, m, n,
                         A, lda, x, incx, (E *)P, rpc, (int)a_aligned, (int)x_aligned);
    else
      hipLaunchKernelGGL((gemv_t_kernel<E, false>), dim3((unsigned)col_blocks, (unsigned)chunks), dim3(256), 0, s, m,
                         n, A, lda, x, incx, (E *)P, rpc, (int)a_aligned, (int)x_aligned);
    CORAL_LAUNCH_CHECK();
    timing_end(slot, s);
    hipLaunchKernelGGL(gemv_reduce_kernel<E>, dim3((unsigned)cdiv(n, 64)), dim3(256), 0, s, n, chunks,
                       (const E *)P, alpha, beta, (int)is_zero_h(beta), y, incy);
    CORAL_LAUNCH_CHECK();
  }
  return CORAL_OK;
}

// Scratch of the allocation-free DGEMV (coral_dgemv_noalloc): the partial sums of its
// column chunks (N) or row chunks (T), at most kNoAllocScratch bytes, so y may hold
// at most kNoAllocScratch / 8 = 2^19 elements.  One copy per GPU, held until the
// process ends; every such call on a GPU shares it (calls that overlap on one GPU
// race on it).  Host code uses only the device address hipGetSymbolAddress returns
// for the current device, never the array's name.
constexpr size_t kNoAllocScratch = size_t(4) << 20;
static __device__ __align__(256) char scratch[kNoAllocScratch];
constexpr int kMaxDevices = 64;
static void *g_scratch_ptr[kMaxDevices];
static std::mutex g_scratch_mu;

static int scratch_address(void **ptr) {
  int dev = 0;
  CORAL_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    set_last_error("device %d is beyond the %d devices the allocation-free GEMV supports", dev, kMaxDevices);
    return CORAL_ERR_UNSUPPORTED;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    CORAL_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = p;
  }
  *ptr = g_scratch_ptr[dev];
  return CORAL_OK;
}

int dgemv_noalloc(int32_t trans, int64_t m, int64_t n, double alpha, const double *a, int64_t lda,
                  const double *x, int64_t incx, double beta, double *y, int64_t incy, coral_stream_t stream) {
  if (!valid_op(trans)) return -1;
  const bool notrans = trans == CORAL_NO_TRANSPOSE;
  // the partial sums of one chunk cover all of y
  if (notrans && m > (int64_t)(kNoAllocScratch / sizeof(double))) {
    set_last_error("coral_dgemv_noalloc: m = %lld exceeds the %zu rows its fixed scratch holds", (long long)m,
                   kNoAllocScratch / sizeof(double));
    return -2;
  }
  if (!notrans && n > (int64_t)(kNoAllocScratch / sizeof(double))) {
    set_last_error("coral_dgemv_noalloc: n = %lld exceeds the %zu columns its fixed scratch holds", (long long)n,
                   kNoAllocScratch / sizeof(double));
    return -3;
  }
  CORAL_TRY(ensure_init());
  GemvScratch fs;
  CORAL_TRY(scratch_address(&fs.fixed));
  fs.cap = kNoAllocScratch;
  return gemv_impl<double>(trans, m, n, alpha, a, lda, x, incx, beta, y, incy, stream, fs);
}

template int launch_beta_only<float>(int64_t, float, float *, int64_t, hipStream_t);
template int launch_beta_only<double>(int64_t, double, double *, int64_t, hipStream_t);
template int launch_beta_only<float2>(int64_t, float2, float2 *, int64_t, hipStream_t);
template int launch_beta_only<double2>(int64_t, double2, double2 *, int64_t, hipStream_t);

}  // namespace l2
}  // namespace coral

using namespace coral;
using namespace coral::l2;

extern "C" {

int coral_sgemv(int32_t trans, int64_t m, int64_t n, float alpha, const float *a, int64_t lda, const float *x,
                int64_t incx, float beta, float *y, int64_t incy, coral_stream_t stream) {
  return gemv_impl<float>(trans, m, n, alpha, a, lda, x, incx, beta, y, incy, stream);
}
int coral_dgemv(int32_t trans, int64_t m, int64_t n, double alpha, const double *a, int64_t lda, const double *x,
                int64_t incx, double beta, double *y, int64_t incy, coral_stream_t stream) {
  return gemv_impl<double>(trans, m, n, alpha, a, lda, x, incx, beta, y, incy, stream);
}
int coral_dgemv_noalloc(int32_t trans, int64_t m, int64_t n, double alpha, const double *a, int64_t lda,
                        const double *x, int64_t incx, double beta, double *y, int64_t incy, coral_stream_t stream) {
  return dgemv_noalloc(trans, m, n, alpha, a, lda, x, incx, beta, y, incy, stream);
}
int coral_cgemv(int32_t trans, int64_t m, int64_t n, const float alpha[2], const float *a, int64_t lda,
                const float *x, int64_t incx, const float beta[2], float *y, int64_t incy, coral_stream_t stream) {
  if (!alpha) return -4;
  if (!beta) return -9;
  return gemv_impl<float2>(trans, m, n, make_float2(alpha[0], alpha[1]), (const float2 *)a, lda, (const float2 *)x,
                           incx, make_float2(beta[0], beta[1]), (float2 *)y, incy, stream);
}
int coral_zgemv(int32_t trans, int64_t m, int64_t n, const double alpha[2], const double *a, int64_t lda,
                const double *x, int64_t incx, const double beta[2], double *y, int64_t incy, coral_stream_t stream) {
  if (!alpha) return -4;
  if (!beta) return -9;
  return gemv_impl<double2>(trans, m, n, make_double2(alpha[0], alpha[1]), (const double2 *)a, lda,
                            (const double2 *)x, incx, make_double2(beta[0], beta[1]), (double2 *)y, incy, stream);
}

}  // extern "C"
