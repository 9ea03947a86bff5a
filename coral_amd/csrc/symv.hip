// Level-2 SYMV / HEMV for gfx950: y := alpha*A*x + beta*y with A symmetric
// (s/d) or Hermitian (c/z), read from the `uplo` triangle only.
//
// CORAL walks the stored triangle once per column (lda == n path,
// coral-aarch64/src/level2/chemv.rs:142-231): every stored a_ij feeds y_i (axpy)
// and y_j (conjugated dot), the diagonal contributes Re(a_jj) only (:178-182), and
// the other triangle is never used.  Here the stored triangle is cut into
// NB x NB blocks, one workgroup per block, and each block is streamed from HBM
// exactly once (a diagonal block is read whole; its unstored half is masked out):
//   off-diagonal block (r, c):  u = A_rc * x_c  -> partial for row block r
//                               v = A_rc^H * x_r -> partial for row block c
//   diagonal block (b, b):      the stored half of the block, both roles.
// Every row block receives exactly nb partials (one per partner block), stored in
// Q[r][c][NB]; a second kernel adds them in c order and applies alpha/beta
// (beta == 0 overwrites y, as chemv.rs:135-139 does).
//
// TRMV (x := op(A) x, A triangular; coral-aarch64/src/level2/dtrmv.rs:58,
// dtrlmv.rs:47-469, dtrumv.rs:48-460, ctrmv.rs, ztrmv.rs) streams the stored
// triangle the same way and keeps one of the two products: op = N takes the
// block's rows (partial of row block r from partner c), op = T / C its columns
// (partial of row block c from partner r).  A row block's partners are then
// the blocks on one side of the diagonal, which the reduce adds in order.
#include "coral_common.hpp"

#include <algorithm>
#include <math.h>

namespace coral {
namespace sym {

constexpr int NB = 64;  // block edge

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

typedef uint32_t sv4u __attribute__((ext_vector_type(4)));

// Streams the NB x NB block at (r0, c0) into LDS, s[c * (NB + 1) + r], with all of
// the block's loads issued before the first LDS store: 16-byte nontemporal loads
// when the block lies inside the matrix and the columns are 16-byte aligned (each
// half-wave or wave covers one contiguous column run), element loads with zero
// fill otherwise.
template <typename E>
__device__ __forceinline__ void stream_block(const E *__restrict__ A, int64_t lda, int64_t n, int64_t r0, int64_t c0,
                                             int vec_ok, E *s) {
  constexpr int EV = 16 / sizeof(E);  // elements per 16-byte vector
  constexpr int VPC = NB / EV;        // vectors per block column
  constexpr int LDP = NB + 1;         // padded column length: column reads hit distinct banks
  constexpr int IT = NB * VPC / 256;  // vectors per thread
  const int t = threadIdx.x;
  const bool full = vec_ok && c0 + NB <= n && r0 + NB <= n;
  E blk[IT][EV];
  if (full) {
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int q = t + 256 * it;
      const int c = q / VPC, rv = (q % VPC) * EV;
      const sv4u pk = __builtin_nontemporal_load(reinterpret_cast<const sv4u *>(A + (c0 + c) * lda + r0 + rv));
      __builtin_memcpy(&blk[it][0], &pk, 16);
    }
  } else {
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int q = t + 256 * it;
      const int c = q / VPC, rv = (q % VPC) * EV;
      const int64_t gr = r0 + rv, gc = c0 + c;
#pragma unroll
      for (int e = 0; e < EV; ++e) blk[it][e] = (gc < n && gr + e < n) ? A[gc * lda + gr + e] : ezero<E>();
    }
  }
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int q = t + 256 * it;
    const int c = q / VPC, rv = (q % VPC) * EV;
#pragma unroll
    for (int e = 0; e < EV; ++e) s[c * LDP + rv + e] = blk[it][e];
  }
}

// One stored NB x NB block per workgroup, streamed into LDS; then u = A_blk x_c
// (by rows) and v = A_blk^H x_r (by columns) are formed from LDS by four row/column
// segments each and combined in segment order.
template <typename E, bool HERM>
__global__ __launch_bounds__(256) void symv_block_kernel(int64_t n, int upper, const E *__restrict__ A, int64_t lda,
                                                         const E *__restrict__ x, int64_t incx, E *__restrict__ Q,
                                                         int64_t nb, int vec_ok) {
  constexpr int LDP = NB + 1;
  __shared__ E s[NB * LDP];                  // s[c * LDP + r]
  __shared__ E xr_s[NB], xc_s[NB];
  __shared__ E up[4][NB], vp[4][NB];
  int64_t lo, hi;
  block_of(blockIdx.x, lo, hi);
  // stored block (br, bc).  Consecutive workgroups take consecutive row blocks of
  // one block column (upper: rows 0..hi of column hi; lower: the mirror image), so
  // concurrently running workgroups read neighbouring runs of the same columns.
  const int64_t br = upper ? lo : nb - 1 - lo, bc = upper ? hi : nb - 1 - hi;
  const int64_t r0 = br * NB, c0 = bc * NB;
  const int t = threadIdx.x;
  const bool diag = br == bc;
  stream_block(A, lda, n, r0, c0, vec_ok, s);
  if (t < NB) {
    const int64_t i = r0 + t;
    xr_s[t] = i < n ? x[i * incx] : ezero<E>();
  } else if (t < 2 * NB) {
    const int64_t j = c0 + t - NB;
    xc_s[t - NB] = j < n ? x[j * incx] : ezero<E>();
  }
  __syncthreads();
  // ---- u (rows) and v (columns), four segments of NB/4 each ----
  constexpr int SEG = NB / 4;
  const int idx = t % NB, seg = t / NB, k0 = seg * SEG;
  E xcv[SEG], xrv[SEG];
#pragma unroll
  for (int kk = 0; kk < SEG; ++kk) {  // broadcast reads, once
    xcv[kk] = xc_s[k0 + kk];
    xrv[kk] = xr_s[k0 + kk];
  }
  E u = ezero<E>(), v = ezero<E>();
  if (!diag) {
#pragma unroll
    for (int kk = 0; kk < SEG; ++kk) {
      u = emad(u, s[(k0 + kk) * LDP + idx], xcv[kk]);        // row idx, column k
      v = emad<HERM>(v, s[idx * LDP + k0 + kk], xrv[kk]);    // column idx, row k
    }
  } else {
    // stored half only (upper: row <= column); the diagonal counts once, in u,
    // with its imaginary part dropped for Hermitian matrices
#pragma unroll
    for (int kk = 0; kk < SEG; ++kk) {
      const int k = k0 + kk;
      if (upper ? idx <= k : idx >= k) {
        E a = s[k * LDP + idx];
        if (HERM && idx == k) a = real_part(a);
        u = emad(u, a, xcv[kk]);
      }
      if (upper ? k < idx : k > idx) v = emad<HERM>(v, s[idx * LDP + k], xrv[kk]);
    }
  }
  up[seg][idx] = u;
  vp[seg][idx] = v;
  __syncthreads();
  if (t < NB) {
    const E uu = eadd(eadd(up[0][t], up[1][t]), eadd(up[2][t], up[3][t]));
    const E vv = eadd(eadd(vp[0][t], vp[1][t]), eadd(vp[2][t], vp[3][t]));
    if (diag) {
      Q[(br * nb + bc) * NB + t] = eadd(uu, vv);
    } else {
      Q[(br * nb + bc) * NB + t] = uu;  // partial of row block br from partner bc
      Q[(bc * nb + br) * NB + t] = vv;  // partial of row block bc from partner br
    }
  }
}

// The same block product with the block staged in LDS one half (32 columns) at a
// time.  Every thread's loads are issued into registers first, as in stream_block;
// only the LDS copy is split, which halves the block's LDS (complex double: 47 KB
// per workgroup instead of 76 KB, three workgroups per CU instead of two).  Thread t
// holds rows (t % VPC)*EV + e of column t / VPC + (256 / VPC) * it (it < IT), so
// half h holds it in [h*IT/2, (h+1)*IT/2).  u (rows) accumulates over both halves in
// four column segments of 8 per half; v (columns) takes each half's 32 columns in
// eight row segments of 8.
template <typename E, bool HERM>
__global__ __launch_bounds__(256) void symv_block_halves_kernel(int64_t n, int upper, const E *__restrict__ A,
                                                                int64_t lda, const E *__restrict__ x, int64_t incx,
                                                                E *__restrict__ Q, int64_t nb, int vec_ok) {
  constexpr int EV = 16 / sizeof(E), VPC = NB / EV, CPI = 256 / VPC;  // columns per it step
  constexpr int HC = NB / 2, LDP = NB + 1, IT = NB * VPC / 256;
  static_assert(CPI * IT / 2 == HC, "each half is half of a thread's vectors");
  __shared__ E s[HC * LDP];  // s[(c - h*HC) * LDP + r]
  __shared__ E xr_s[NB], xc_s[NB];
  __shared__ E up[4][NB], vp[8][NB];
  int64_t lo, hi;
  block_of(blockIdx.x, lo, hi);
  const int64_t br = upper ? lo : nb - 1 - lo, bc = upper ? hi : nb - 1 - hi;
  const int64_t r0 = br * NB, c0 = bc * NB;
  const int t = threadIdx.x, tc = t / VPC, rv = (t % VPC) * EV;
  const bool diag = br == bc;
  E blk[IT][EV];
  if (vec_ok && c0 + NB <= n && r0 + NB <= n) {
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const sv4u pk =
          __builtin_nontemporal_load(reinterpret_cast<const sv4u *>(A + (c0 + tc + CPI * it) * lda + r0 + rv));
      __builtin_memcpy(&blk[it][0], &pk, 16);
    }
  } else {
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int64_t gc = c0 + tc + CPI * it;
#pragma unroll
      for (int e = 0; e < EV; ++e) {
        const int64_t gr = r0 + rv + e;
        blk[it][e] = (gc < n && gr < n) ? A[gc * lda + gr] : ezero<E>();
      }
    }
  }
  if (t < NB) {
    const int64_t i = r0 + t;
    xr_s[t] = i < n ? x[i * incx] : ezero<E>();
  } else if (t < 2 * NB) {
    const int64_t j = c0 + t - NB;
    xc_s[t - NB] = j < n ? x[j * incx] : ezero<E>();
  }
  const int idx = t % NB, seg = t / NB;  // u: row idx, columns h*HC + seg*8 + kk
  const int cl = t % HC, rs = t / HC;    // v: column h*HC + cl, rows rs*8 + kk
  E u = ezero<E>();
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    if (h) __syncthreads();  // every read of the first half is done
#pragma unroll
    for (int it = h * IT / 2; it < (h + 1) * IT / 2; ++it)
#pragma unroll
      for (int e = 0; e < EV; ++e) s[(tc + CPI * it - h * HC) * LDP + rv + e] = blk[it][e];
    __syncthreads();
    E v = ezero<E>();
    const int c = h * HC + cl;
#pragma unroll
    for (int kk = 0; kk < 8; ++kk) {
      const int kc = seg * 8 + kk, k = h * HC + kc;  // column of u
      const int r = rs * 8 + kk;                     // row of v
      if (!diag) {
        u = emad(u, s[kc * LDP + idx], xc_s[k]);
        v = emad<HERM>(v, s[cl * LDP + r], xr_s[r]);
      } else {
        // stored half only (upper: row <= column); the diagonal counts once, in u,
        // with its imaginary part dropped for Hermitian matrices
        if (upper ? idx <= k : idx >= k) {
          E a = s[kc * LDP + idx];
          if (HERM && idx == k) a = real_part(a);
          u = emad(u, a, xc_s[k]);
        }
        if (upper ? r < c : r > c) v = emad<HERM>(v, s[cl * LDP + r], xr_s[r]);
      }
    }
    vp[rs][c] = v;
  }
  up[seg][idx] = u;
  __syncthreads();
  if (t < NB) {
    const E uu = eadd(eadd(up[0][t], up[1][t]), eadd(up[2][t], up[3][t]));
    const E vv = eadd(eadd(eadd(vp[0][t], vp[1][t]), eadd(vp[2][t], vp[3][t])),
                      eadd(eadd(vp[4][t], vp[5][t]), eadd(vp[6][t], vp[7][t])));
    if (diag) {
      Q[(br * nb + bc) * NB + t] = eadd(uu, vv);
    } else {
      Q[(br * nb + bc) * NB + t] = uu;
      Q[(bc * nb + br) * NB + t] = vv;
    }
  }
}

// y[i] := alpha * sum_c Q[r][c][w] + beta * y[i] with i = r*NB + w.  A workgroup
// takes 64 rows; its four wavefronts sum partner ranges [g*q, (g+1)*q) in order
// (eight loads in flight) and the group sums are added in group order, so the
// result does not depend on scheduling.
template <typename E>
__global__ __launch_bounds__(256) void symv_reduce_kernel(int64_t n, int64_t nb, const E *__restrict__ Q, E alpha,
                                                          E beta, int beta_zero, E *y, int64_t incy) {
  __shared__ E part[4][64];
  const int lane = threadIdx.x & 63, g = threadIdx.x >> 6;
  const int64_t i = (int64_t)blockIdx.x * 64 + lane;
  const int64_t r = i / NB, w = i % NB;
  const int64_t q = (nb + 3) / 4;
  const int64_t c0 = g * q, c1 = c0 + q < nb ? c0 + q : nb;
  E acc = ezero<E>();
  if (i < n) {
    const E *base = Q + r * nb * NB + w;
    int64_t c = c0;
    for (; c + 8 <= c1; c += 8) {
      E v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = base[(c + u) * NB];
#pragma unroll
      for (int u = 0; u < 8; ++u) acc = eadd(acc, v[u]);
    }
    for (; c < c1; ++c) acc = eadd(acc, base[c * NB]);
  }
  part[g][lane] = acc;
  __syncthreads();
  if (g == 0 && i < n) {
    const E tsum = eadd(eadd(part[0][lane], part[1][lane]), eadd(part[2][lane], part[3][lane]));
    E *yp = y + i * incy;
    *yp = axpby(alpha, tsum, beta, yp, beta_zero);
  }
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
  const int vec_ok = ((uintptr_t)A % 16 == 0) && ((lda * (int64_t)sizeof(E)) % 16 == 0);
  // LDS per workgroup decides how many run on a CU.  Measured, alternating runs on
  // one box: DSYMV upper 16384 5.29 TB/s whole block in LDS, 5.58 in halves; ZHEMV
  // lower 8192 5.0 / 5.40-5.47; SSYMV upper 16384 5.73-5.75 / 5.63-5.72 (its
  // whole-block LDS is small already).
  if constexpr (sizeof(E) >= 8)
    hipLaunchKernelGGL((symv_block_halves_kernel<E, HERM>), dim3((unsigned)nblocks), dim3(256), 0, s, n,
                       (int)(uplo == CORAL_UPPER_TRIANGULAR), A, lda, x, incx, (E *)Q, nb, vec_ok);
  else
    hipLaunchKernelGGL((symv_block_kernel<E, HERM>), dim3((unsigned)nblocks), dim3(256), 0, s, n,
                       (int)(uplo == CORAL_UPPER_TRIANGULAR), A, lda, x, incx, (E *)Q, nb, vec_ok);
  CORAL_LAUNCH_CHECK();
  hipLaunchKernelGGL(symv_reduce_kernel<E>, dim3((unsigned)cdiv(n, 64)), dim3(256), 0, s, n, nb, (const E *)Q, alpha,
                     beta, (int)eiszero(beta), y, incy);
  CORAL_LAUNCH_CHECK();
  return CORAL_OK;
}

// ---- TRMV ----
// OP 0: y = A x (rows of the block times x_c); OP 1: A^T x, OP 2: A^H x (columns
// times x_r).  In a diagonal block only the stored half counts; a unit diagonal
// contributes x itself and its stored value is never used.
template <typename E, int OP>
__global__ __launch_bounds__(256) void trmv_block_kernel(int64_t n, int upper, int unit, const E *__restrict__ A,
                                                         int64_t lda, const E *__restrict__ x, int64_t incx,
                                                         E *__restrict__ Q, int64_t nb, int vec_ok) {
  constexpr int LDP = NB + 1;
  __shared__ E s[NB * LDP];
  __shared__ E xs[NB];
  __shared__ E part[4][NB];
  int64_t lo, hi;
  block_of(blockIdx.x, lo, hi);
  const int64_t br = upper ? lo : nb - 1 - lo, bc = upper ? hi : nb - 1 - hi;
  const int64_t r0 = br * NB, c0 = bc * NB;
  const int t = threadIdx.x;
  const bool diag = br == bc;
  stream_block(A, lda, n, r0, c0, vec_ok, s);
  if (t < NB) {
    const int64_t j = (OP == 0 ? c0 : r0) + t;
    xs[t] = j < n ? x[j * incx] : ezero<E>();
  }
  __syncthreads();
  constexpr int SEG = NB / 4;
  const int idx = t % NB, seg = t / NB, k0 = seg * SEG;
  E acc = ezero<E>();
  if (!diag) {
#pragma unroll
    for (int kk = 0; kk < SEG; ++kk) {
      const int k = k0 + kk;
      acc = emad<OP == 2>(acc, OP == 0 ? s[k * LDP + idx] : s[idx * LDP + k], xs[k]);
    }
  } else {
#pragma unroll
    for (int kk = 0; kk < SEG; ++kk) {
      const int k = k0 + kk;
      const int row = OP == 0 ? idx : k, col = OP == 0 ? k : idx;
      if (row == col) {
        acc = unit ? eadd(acc, xs[k]) : emad<OP == 2>(acc, s[k * LDP + k], xs[k]);
      } else if (upper ? row < col : row > col) {
        acc = emad<OP == 2>(acc, OP == 0 ? s[k * LDP + idx] : s[idx * LDP + k], xs[k]);
      }
    }
  }
  part[seg][idx] = acc;
  __syncthreads();
  if (t < NB) {
    const E sum = eadd(eadd(part[0][t], part[1][t]), eadd(part[2][t], part[3][t]));
    const int64_t dest = OP == 0 ? br : bc, partner = OP == 0 ? bc : br;
    Q[(dest * nb + partner) * NB + t] = sum;
  }
}

// x[i] := sum of row block r's partials over its partners, which are c = 0..r
// (from_zero) or c = r..nb-1.  A workgroup takes one row block (64 rows); its four
// wavefronts sum consecutive quarters of the partner range in order and the
// quarter sums are added in order, so the result does not depend on scheduling.
template <typename E>
__global__ __launch_bounds__(256) void trmv_reduce_kernel(int64_t n, int64_t nb, const E *__restrict__ Q, int from_zero,
                                                          E *x, int64_t incx) {
  __shared__ E part[4][64];
  const int lane = threadIdx.x & 63, g = threadIdx.x >> 6;
  const int64_t r = blockIdx.x, i = r * NB + lane;
  const int64_t lo = from_zero ? 0 : r, cnt = from_zero ? r + 1 : nb - r;
  const int64_t q = (cnt + 3) / 4;
  const int64_t c0 = lo + g * q, c1 = lo + (g * q + q < cnt ? g * q + q : cnt);
  E acc = ezero<E>();
  const E *base = Q + r * nb * NB + lane;
  int64_t c = c0;
  for (; c + 8 <= c1; c += 8) {
    E v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = base[(c + u) * NB];
#pragma unroll
    for (int u = 0; u < 8; ++u) acc = eadd(acc, v[u]);
  }
  for (; c < c1; ++c) acc = eadd(acc, base[c * NB]);
  part[g][lane] = acc;
  __syncthreads();
  if (g == 0 && i < n) x[i * incx] = eadd(eadd(part[0][lane], part[1][lane]), eadd(part[2][lane], part[3][lane]));
}

// CORAL argument order: uplo=1 trans=2 diag=3 n=4 a=5 lda=6 x=7 incx=8
template <typename E>
static int trmv_impl(int32_t uplo, int32_t trans, int32_t diag, int64_t n, const E *a, int64_t lda, E *x,
                     int64_t incx, coral_stream_t stream) {
  if (!valid_uplo(uplo)) return -1;
  if (!valid_op(trans)) return -2;
  if (!valid_diag(diag)) return -3;
  if (n < 0) return -4;
  if (lda < imax64(1, n)) return -6;
  if (incx <= 0) return -8;
  if (n == 0) return CORAL_OK;
  if (!a) return -5;
  if (!x) return -7;
  CORAL_TRY(ensure_init());
  hipStream_t s = resolve(stream);
  const int64_t nb = cdiv(n, NB);
  void *Q = nullptr;
  CORAL_TRY(workspace(s, WS_AUX2, sizeof(E) * (size_t)(nb * nb * NB), &Q));
  // real routines map ConjugateTranspose to Transpose (coral-aarch64/src/level2/dtrmv.rs:58-79)
  const int op = (trans == CORAL_CONJUGATE_TRANSPOSE && !elem_traits<E>::cplx) ? 1 : trans;
  const int upper = uplo == CORAL_UPPER_TRIANGULAR, unit = diag == CORAL_UNIT_DIAGONAL;
  const int vec_ok = ((uintptr_t)a % 16 == 0) && ((lda * (int64_t)sizeof(E)) % 16 == 0);
  const dim3 grid((unsigned)(nb * (nb + 1) / 2));
  if (op == 0)
    hipLaunchKernelGGL((trmv_block_kernel<E, 0>), grid, dim3(256), 0, s, n, upper, unit, a, lda, (const E *)x, incx,
                       (E *)Q, nb, vec_ok);
  else if (op == 1)
    hipLaunchKernelGGL((trmv_block_kernel<E, 1>), grid, dim3(256), 0, s, n, upper, unit, a, lda, (const E *)x, incx,
                       (E *)Q, nb, vec_ok);
  else
    hipLaunchKernelGGL((trmv_block_kernel<E, 2>), grid, dim3(256), 0, s, n, upper, unit, a, lda, (const E *)x, incx,
                       (E *)Q, nb, vec_ok);
  CORAL_LAUNCH_CHECK();
  // partners on the diagonal's near side: N lower and T/C upper sum c = 0..r
  const int from_zero = (!upper) == (op == 0);
  hipLaunchKernelGGL(trmv_reduce_kernel<E>, dim3((unsigned)nb), dim3(256), 0, s, n, nb, (const E *)Q, from_zero, x,
                     incx);
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

#define TRMV_ENTRY(NAME, E, T)                                                                               \
  int coral_##NAME(int32_t uplo, int32_t trans, int32_t diag, int64_t n, const T *a, int64_t lda, T *x,     \
                   int64_t incx, coral_stream_t s) {                                                         \
    return trmv_impl<E>(uplo, trans, diag, n, (const E *)a, lda, (E *)x, incx, s);                          \
  }
TRMV_ENTRY(strmv, float, float)
TRMV_ENTRY(dtrmv, double, double)
TRMV_ENTRY(ctrmv, float2, float)
TRMV_ENTRY(ztrmv, double2, double)

}  // extern "C"
