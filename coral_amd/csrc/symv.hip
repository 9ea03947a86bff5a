// Level-2 SYMV / HEMV for gfx950: y := alpha*A*x + beta*y with A symmetric
// (s/d) or Hermitian (c/z), read from the `uplo` triangle only.
//
// CORAL walks the stored triangle once per column (lda == n path,
// coral-aarch64/src/level2/chemv.rs:142-231): every stored a_ij feeds y_i (axpy)
// and y_j (conjugated dot), the diagonal contributes Re(a_jj) only (:178-182), and
// the other triangle is never touched.  Here the stored triangle is cut into
// NB x NB blocks, one workgroup per block, and each block is streamed from HBM
// exactly once:
//   off-diagonal block (r, c):  u = A_rc * x_c  -> partial for row block r
//                               v = A_rc^H * x_r -> partial for row block c
//   diagonal block (b, b):      the stored half of the block, both roles.
// Every row block receives exactly nb partials (one per partner block), stored in
// Q[r][c][NB]; a second kernel adds them in c order and applies alpha/beta
// (beta == 0 overwrites y, as chemv.rs:135-139 does).
#include "coral_common.hpp"

#include <algorithm>
#include <math.h>

namespace coral {
namespace sym {

constexpr int NB = 128;  // block edge; 4 wavefronts x 32 columns, 2 rows per lane

template <typename E> __device__ __forceinline__ E real_part(E a) {
  if constexpr (elem_traits<E>::cplx) { a.y = 0; return a; }
  else return a;
}

// blocks enumerated column by column over the upper block triangle: t -> (lo, hi), lo <= hi
__device__ __forceinline__ void block_of(int64_t t, int64_t &lo, int64_t &hi) {
  int64_t c = (int64_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
  while (c * (c + 1) / 2 > t) --c;
  while ((c + 1) * (c + 2) / 2 <= t) ++c;
  hi = c;
  lo = t - c * (c + 1) / 2;
}

template <typename E, bool HERM>
__global__ __launch_bounds__(256) void symv_block_kernel(int64_t n, int upper, const E *__restrict__ A, int64_t lda,
                                                         const E *__restrict__ x, int64_t incx, E *__restrict__ Q,
                                                         int64_t nb) {
  __shared__ E xr_s[NB], xc_s[NB];
  __shared__ E vd_s[4][NB];  // v of diagonal-block columns, by owning wavefront
  int64_t lo, hi;
  block_of(blockIdx.x, lo, hi);
  // stored block (br, bc): upper -> br <= bc, lower -> br >= bc
  const int64_t br = upper ? lo : hi, bc = upper ? hi : lo;
  const int64_t r0 = br * NB, c0 = bc * NB;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (tid < NB) {
    int64_t i = r0 + tid;
    xr_s[tid] = i < n ? x[i * incx] : ezero<E>();
  } else {
    int64_t j = c0 + tid - NB;
    xc_s[tid - NB] = j < n ? x[j * incx] : ezero<E>();
  }
  __syncthreads();
  const bool diag = br == bc;
  const int64_t ra = r0 + lane, rb = r0 + lane + 64;  // the two rows of this lane
  E ua = ezero<E>(), ub = ezero<E>();
  E xa = xr_s[lane], xb = xr_s[lane + 64];
  for (int jj = wave; jj < NB; jj += 4) {
    const int64_t j = c0 + jj;
    E v = ezero<E>();
    if (j < n) {
      const E *col = A + j * lda;
      const E xj = xc_s[jj];
      if (!diag) {
        E a0 = ra < n ? col[ra] : ezero<E>();
        E a1 = rb < n ? col[rb] : ezero<E>();
        ua = emad(ua, a0, xj);
        ub = emad(ub, a1, xj);
        v = emad<HERM>(v, a0, xa);
        v = emad<HERM>(v, a1, xb);
      } else {
        // stored half only: upper keeps row <= col, lower keeps row >= col
        const int64_t rl[2] = {ra, rb};
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int64_t r = rl[h];
          if (r >= n) continue;
          const bool stored = upper ? (r <= j) : (r >= j);
          if (!stored) continue;
          E a = col[r];
          if (r == j) {
            E ad = HERM ? real_part(a) : a;
            if (h == 0) ua = emad(ua, ad, xj); else ub = emad(ub, ad, xj);
          } else {
            if (h == 0) ua = emad(ua, a, xj); else ub = emad(ub, a, xj);
            v = emad<HERM>(v, a, h == 0 ? xa : xb);
          }
        }
      }
    }
    v = wave_esum(v);
    if (lane == 0) {
      // v is the column-block contribution for row jj of block bc
      if (!diag) Q[(bc * nb + br) * NB + jj] = v;
      else vd_s[wave][jj] = v;  // staged; combined with u below (diag block: both into Q[b][b])
    }
  }
  // u from the four wavefronts -> partial for row block br
  __syncthreads();
  __shared__ E uw[4][NB];
  uw[wave][lane] = ua;
  uw[wave][lane + 64] = ub;
  __syncthreads();
  if (tid < NB) {
    E s = eadd(eadd(uw[0][tid], uw[1][tid]), eadd(uw[2][tid], uw[3][tid]));
    if (diag) {
      // v of column tid was written by wave (tid % 4)
      s = eadd(s, vd_s[tid & 3][tid]);
    }
    Q[(br * nb + bc) * NB + tid] = s;
  }
}

template <typename E>
__global__ void symv_reduce_kernel(int64_t n, int64_t nb, const E *__restrict__ Q, E alpha, E beta, int beta_zero,
                                   E *y, int64_t incy) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t r = i / NB, w = i % NB;
  E s = ezero<E>();
  for (int64_t c = 0; c < nb; ++c) s = eadd(s, Q[(r * nb + c) * NB + w]);
  E *yp = y + i * incy;
  *yp = axpby(alpha, s, beta, yp, beta_zero);
}

template <typename E>
__global__ void symv_beta_kernel(int64_t n, E beta, int beta_zero, E *y, int64_t incy) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    E *yp = y + i * incy;
    *yp = beta_zero ? ezero<E>() : emul(beta, *yp);
  }
}

// CORAL argument order: uplo=1 n=2 alpha=3 a=4 lda=5 x=6 incx=7 beta=8 y=9 incy=10
template <typename E, bool HERM>
static int symv_impl(int32_t uplo, int64_t n, E alpha, const E *A, int64_t lda, const E *x, int64_t incx, E beta,
                     E *y, int64_t incy, coral_stream_t stream) {
  if (!valid_uplo(uplo)) return -1;
  if (n < 0) return -2;
  if (lda < imax64(1, n)) return -5;
  if (incx <= 0) return -7;
  if (incy <= 0) return -10;
  // quick return, coral-aarch64/src/level2/chemv.rs:109-110
  if (n == 0) return CORAL_OK;
  if (eiszero(alpha) && eisone(beta)) return CORAL_OK;
  if (!y) return -9;
  CORAL_TRY(ensure_init());
  hipStream_t s = resolve(stream);
  if (eiszero(alpha)) {
    int64_t blocks = std::min<int64_t>(cdiv(n, 256), 1024);
    hipLaunchKernelGGL(symv_beta_kernel<E>, dim3((unsigned)blocks), dim3(256), 0, s, n, beta, (int)eiszero(beta), y,
                       incy);
    CORAL_LAUNCH_CHECK();
    return CORAL_OK;
  }
  if (!A) return -4;
  if (!x) return -6;
  const int64_t nb = cdiv(n, NB);
  void *Q = nullptr;
  CORAL_TRY(workspace(s, WS_AUX2, sizeof(E) * (size_t)(nb * nb * NB), &Q));
  const int64_t nblocks = nb * (nb + 1) / 2;
  hipLaunchKernelGGL((symv_block_kernel<E, HERM>), dim3((unsigned)nblocks), dim3(256), 0, s, n,
                     (int)(uplo == CORAL_UPPER_TRIANGULAR), A, lda, x, incx, (E *)Q, nb);
  CORAL_LAUNCH_CHECK();
  hipLaunchKernelGGL(symv_reduce_kernel<E>, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, s, n, nb, (const E *)Q, alpha,
                     beta, (int)eiszero(beta), y, incy);
  CORAL_LAUNCH_CHECK();
  return CORAL_OK;
}

}  // namespace sym
}  // namespace coral

using namespace coral;
using namespace coral::sym;

extern "C" {

int coral_ssymv(int32_t uplo, int64_t n, float alpha, const float *a, int64_t lda, const float *x, int64_t incx,
                float beta, float *y, int64_t incy, coral_stream_t s) {
  return symv_impl<float, false>(uplo, n, alpha, a, lda, x, incx, beta, y, incy, s);
}
int coral_dsymv(int32_t uplo, int64_t n, double alpha, const double *a, int64_t lda, const double *x, int64_t incx,
                double beta, double *y, int64_t incy, coral_stream_t s) {
  return symv_impl<double, false>(uplo, n, alpha, a, lda, x, incx, beta, y, incy, s);
}
int coral_chemv(int32_t uplo, int64_t n, const float alpha[2], const float *a, int64_t lda, const float *x,
                int64_t incx, const float beta[2], float *y, int64_t incy, coral_stream_t s) {
  if (!alpha) return -3;
  if (!beta) return -8;
  return symv_impl<float2, true>(uplo, n, make_float2(alpha[0], alpha[1]), (const float2 *)a, lda, (const float2 *)x,
                                 incx, make_float2(beta[0], beta[1]), (float2 *)y, incy, s);
}
int coral_zhemv(int32_t uplo, int64_t n, const double alpha[2], const double *a, int64_t lda, const double *x,
                int64_t incx, const double beta[2], double *y, int64_t incy, coral_stream_t s) {
  if (!alpha) return -3;
  if (!beta) return -8;
  return symv_impl<double2, true>(uplo, n, make_double2(alpha[0], alpha[1]), (const double2 *)a, lda,
                                  (const double2 *)x, incx, make_double2(beta[0], beta[1]), (double2 *)y, incy, s);
}

}  // extern "C"
